"""Raw run summaries and their merge, against the reference's own tester 1 output and merge
(src/tester.c:355-431, built from its sources into oracle/_ref) and, for enumeration shares, against
the reference enumerator's whole run: shares written raw and merged give the whole
(the property of tests/test_work.c:269-291). CPU only."""
import ctypes as C
import socket

import pytest

import tesseract_amd as t
from oracle import pyref

needs_ref = pytest.mark.skipif(not pyref.available(), reason="oracle/_ref is not built")


def ref_t1(cells, programs, max_passes, seed, ratio=4.0):
    cfg = t.t1_config(cells, ratio, programs, max_passes, seed)
    cfg.threads = 1  # one worker takes the programs in order: a deterministic longest_seed
    out = (t.T1SizeStruct * len(cells))()
    wall, threads = pyref.t1_run(C.byref(cfg), out)
    return out, wall, threads


@needs_ref
def test_t1_raw_text_is_the_references():
    out, wall, threads = ref_t1([8, 16, 33], 6, 1 << 12, 5)
    host = socket.gethostname()
    want = pyref.t1_print_raw(out, 3, wall, threads, 1 << 12)
    got = t.t1_raw([t.t1_size_dict(x) for x in out], host, wall, threads, 1 << 12)
    assert got == want
    assert got.splitlines()[1].startswith("T1 8 32 6 3072 ")


@needs_ref
def test_t1_merge_matches_the_references(tmp_path):
    paths = []
    for k, (seed, cells) in enumerate([(1, [8, 16]), (2, [16, 8, 32]), (3, [8])]):
        out, wall, threads = ref_t1(cells, 4, 1 << 10, seed)
        text = t.t1_raw([t.t1_size_dict(x) for x in out], "host%d" % k, wall + k, threads + k,
                        1 << 10)
        p = tmp_path / ("run%d.raw" % k)
        p.write_text(text + "a line of something else\n")
        paths.append(str(p))
    sizes, sec, th, mp = t.t1_merge(paths)
    ref_out = (t.T1SizeStruct * 32)()
    n, rsec, rth, rmp = pyref.t1_merge(paths, ref_out)
    assert len(sizes) == n == 3
    assert sizes == [t.t1_size_dict(ref_out[i]) for i in range(n)]
    assert (sec, th, mp) == (rsec, rth, rmp)
    # the sums themselves: 8 cells appear in three runs of 4 programs each
    assert sizes[0]["cells"] == 8 and sizes[0]["programs"] == 12 and sizes[0]["starts"] == 12 * 512
    assert sum(sizes[0]["hist"]) == 12 * 512


def test_t1_merge_missing_file(tmp_path):
    with pytest.raises(OSError):
        t.t1_merge([str(tmp_path / "absent.raw")])


def test_t1_program_len():
    cfg = t.t1_config(ratio=4.0)
    assert [t.t1_program_len(cfg, c) for c in (1, 8, 256)] == [4, 32, 1024]
    cfg = t.t1_config(ratio=0.1)
    assert t.t1_program_len(cfg, 1) == 1  # never 0 instructions
    assert t.t1_program_len(cfg, 5) == 1  # 0.5 rounds up
    assert t.t1_program_len(cfg, 4) == 1


@needs_ref
@pytest.mark.parametrize("spec,parts", [("p = a * b", 3), ("p = a * b + 1", 3), ("p = a ^ b", 5)])
def test_enum_shares_merge_to_the_whole(tmp_path, spec, parts):
    prog = pyref.make("gen_mul", 10)  # 20 inputs: 32 chunks of 2^15 assignments, shared round robin
    whole = pyref.enum_run(t.EN_VERIFY, prog.ptr, spec_text=spec, threads=2)
    paths = []
    for k in range(parts):
        r = pyref.enum_run(t.EN_VERIFY, prog.ptr, spec_text=spec, part=k, part_end=k, parts=parts,
                           threads=2)
        assert 0 < r["tried"] < whole["tried"]
        p = tmp_path / ("share%d.raw" % k)
        p.write_text(t.enum_raw(t.EN_VERIFY, r))
        paths.append(str(p))
    got = t.enum_merge(paths)
    for key in ("space", "tried", "hits", "hit", "nhit", "ops"):
        assert got[key] == whole[key], key
    assert got["threads"] == 2 * parts


def test_enum_raw_round_trip(tmp_path):
    r = {"space": 1 << 40, "tried": 123456789, "hits": 40, "hit": list(range(5, 85, 5))[:16],
         "seconds": 1.25, "ops": 98765432101, "threads": 16}
    text = t.enum_raw(t.EN_SEARCH, r)
    assert text.splitlines()[0] == "EN 2 1099511627776 123456789 40 1.250000 16 98765432101"
    assert text.splitlines()[1] == "ENHIT 5"
    p = tmp_path / "one.raw"
    p.write_text(text)
    got = t.enum_merge([str(p), str(p)])
    assert got["tried"] == 2 * r["tried"] and got["hits"] == 80 and got["ops"] == 2 * r["ops"]
    assert got["hit"] == sorted(r["hit"] * 2)[:16]
    assert got["seconds"] == 1.25 and got["threads"] == 32 and got["space"] == r["space"]
    with pytest.raises(OSError):
        t.enum_merge([str(tmp_path / "absent.raw")])
