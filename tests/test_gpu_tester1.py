"""Tester 1 on the device (gpu_t1_run): the per-size summaries equal the reference's own t1_run on
one host thread (oracle/_ref, the reference's sources built as they lie) and the oracle's Brent
schedule program by program; a device run and a host run of other programs merge through the raw
form like two host runs."""
import ctypes as C

import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from oracle import pyref
from tests.helpers import golden

pytestmark = pytest.mark.gpu


def bucket(lam):
    b = 0
    while b < 62 and (1 << b) < lam:
        b += 1
    return b


def oracle_t1(cfg):
    """Per-size summaries from the oracle's Brent on group 0 of every program, in the reference's
    program order (sizes take turns, src/tester.c:190-192)."""
    n = cfg.nsizes
    out = []
    for si in range(n):
        cells = cfg.cells[si]
        out.append({"cells": cells, "len": t.t1_program_len(cfg, cells), "programs": 0,
                    "starts": 0, "passes": 0, "unfinished": 0, "hist": [0] * 64, "longest": 0,
                    "longest_seed": 0})
    for j in range(cfg.programs * n):
        si, k = j % n, j // n
        d = out[si]
        ps = t.t1_program_seed(cfg.seed, d["cells"], k)
        prog = t.random_program(d["cells"], d["len"], ps)
        mem, rr = t.t1_start(prog, ps, 1)
        lams, ran = O.brent_lines(prog, mem[0], rr[0], cfg.max_passes, with_passes=True)
        d["programs"] += 1
        d["starts"] += 512
        d["passes"] += ran
        for lam in (int(x) for x in lams):
            if lam == 0:
                d["hist"][63] += 1
                continue
            d["hist"][bucket(lam)] += 1
            if lam > d["longest"]:
                d["longest"], d["longest_seed"] = lam, ps
    return out


def ref_t1(cfg, threads=1):
    c = t.t1_config(list(cfg.cells[:cfg.nsizes]), cfg.ratio, cfg.programs, cfg.max_passes, cfg.seed)
    c.threads = threads
    out = (t.T1SizeStruct * cfg.nsizes)()
    pyref.t1_run(C.byref(c), out)
    return [t.t1_size_dict(x) for x in out]


def test_t1_matches_oracle_and_reference(engine):
    cfg = t.t1_config([1, 3, 8, 16, 40, 64], ratio=4.0, programs=6, max_passes=1 << 12, seed=7)
    print("gpu_t1_run: 6 sizes x 6 programs, 2^12 passes", flush=True)
    got, sec = engine.t1_run(cfg)
    assert sec > 0
    assert got == oracle_t1(cfg)
    if pyref.available():
        assert got == ref_t1(cfg)
    for d in got:
        assert d["unfinished"] == 0 and sum(d["hist"]) == d["starts"] == 512 * d["programs"]


def test_t1_ratio_and_long_search(engine):
    # ratio 0.5 makes programs shorter than their cells; 2^16 passes finds longer cycles
    cfg = t.t1_config([2, 5, 12, 30], ratio=0.5, programs=5, max_passes=1 << 16, seed=3)
    print("gpu_t1_run: ratio 0.5, 2^16 passes", flush=True)
    got, _ = engine.t1_run(cfg)
    assert [d["len"] for d in got] == [1, 3, 6, 15]
    assert got == oracle_t1(cfg)


def test_t1_seconds_mode(engine):
    # small programs: rounds until 0.3 s are up; every finished program adds 512 starts, the
    # searches cut off add their passes and count as unfinished
    cfg = t.t1_config([8, 16], ratio=4.0, programs=0, max_passes=1 << 10, seed=11, seconds=0.3)
    print("gpu_t1_run: rounds for 0.3 s", flush=True)
    got, sec = engine.t1_run(cfg)
    assert 0.3 <= sec < 5.0
    for d in got:
        assert d["programs"] > 0 and sum(d["hist"]) == d["starts"] == 512 * d["programs"]
        assert d["passes"] >= d["programs"]
    assert [c["cells"] for c in engine.t1_choices()] == [8, 16]


def test_t1_seconds_mode_stops_long_searches(engine):
    # 600-cell programs searched up to 2^30 passes: the stop after 1 s cuts the searches in
    # flight at a 64th pass; they add their passes but no cycle lengths
    cfg = t.t1_config([600], ratio=4.0, programs=0, max_passes=1 << 30, seed=12, seconds=1.0)
    print("gpu_t1_run: 600 cells for 1 s", flush=True)
    got, sec = engine.t1_run(cfg)
    d = got[0]
    assert 1.0 <= sec < 10.0
    assert d["unfinished"] > 0 and d["passes"] > 0
    assert sum(d["hist"]) == d["starts"] == 512 * d["programs"]


def test_t1_errors(engine):
    bad = [t.t1_config([8], programs=0, seconds=0.0),
           t.t1_config([8], ratio=0.0),
           t.t1_config([8], max_passes=0),
           t.t1_config([8], max_passes=1 << 33),
           t.t1_config([0])]
    for cfg in bad:
        with pytest.raises(ValueError):
            engine.t1_run(cfg)


@pytest.mark.skipif(not pyref.available(), reason="oracle/_ref is not built")
def test_t1_device_and_host_runs_merge(engine, tmp_path):
    gcfg = t.t1_config([8, 16], 4.0, 8, 1 << 10, 21)
    hcfg = t.t1_config([16, 8], 4.0, 8, 1 << 10, 22)
    dev, dsec = engine.t1_run(gcfg)
    host = ref_t1(hcfg)
    a, b = tmp_path / "gpu.raw", tmp_path / "cpu.raw"
    a.write_text(t.t1_raw(dev, "gpu-box", dsec, 1, 1 << 10))
    b.write_text(t.t1_raw(host, "cpu-box", 2.5, 16, 1 << 10))
    sizes, sec, th, mp = t.t1_merge([str(a), str(b)])
    ref_out = (t.T1SizeStruct * 32)()
    n, rsec, rth, rmp = pyref.t1_merge([str(a), str(b)], ref_out)
    assert sizes == [t.t1_size_dict(ref_out[i]) for i in range(n)]
    assert (th, mp) == (17, 1 << 10) and sec == max(2.5, round(dsec, 6))
    assert [d["programs"] for d in sizes] == [16, 16]


def test_enum_shares_on_device_merge_to_the_whole(engine, tmp_path):
    prog = golden("mul10")
    whole = engine.enum_run(t.EN_VERIFY, prog, spec="p = a * b + 3")
    paths = []
    for k in range(3):
        print("verify share %d of 3" % k, flush=True)
        r = engine.enum_run(t.EN_VERIFY, prog, spec="p = a * b + 3", part=k, part_end=k, parts=3)
        p = tmp_path / ("share%d.raw" % k)
        p.write_text(t.enum_raw(t.EN_VERIFY, r))
        paths.append(str(p))
    got = t.enum_merge(paths)
    for key in ("space", "tried", "hits", "hit", "nhit", "ops"):
        assert got[key] == whole[key], key
