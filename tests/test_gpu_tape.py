"""The tape interpreter tier on the GPU: the same results as the compiled kernel and the oracle."""
import numpy as np
import pytest

import tesseract_amd as t
from tesseract_amd import netlists as N
from oracle import pyoracle as O
from tests.helpers import golden

pytestmark = pytest.mark.gpu

TAPE = {"tape": 1}
KEY = ("tried", "hits", "hit", "ops", "space")


def _same(x, y):
    return all(x[k] == y[k] for k in KEY)


def test_tape_equiv_requests(engine):
    checked = 0
    for k in range(24):
        c, r, n_in, faulty = N.request(5, k)
        if n_in > 24:
            continue
        a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
        print("request", k, "inputs", n_in, "gates", len(c.gates), flush=True)
        got = engine.enum_run(t.EN_EQUIV, a, b, tune=TAPE)
        assert _same(got, engine.enum_run(t.EN_EQUIV, a, b))
        if n_in <= 16:
            o = O.enum(t.EN_EQUIV, a, b, b_in=b.inputs, b_out=b.outputs)
            assert (got["hits"], got["hit"], got["tried"]) == (o["hits"], o["hit"], o["tried"])
        checked += 1
    assert checked >= 10


def test_tape_large_space_and_parts(engine):
    rng = np.random.default_rng(21)
    c = N.circuit(30, 6, 400, rng)
    r = N.fault(N.rewrite(c, rng), rng)
    a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
    print("2^30 equiv, 400 gates", flush=True)
    got = engine.enum_run(t.EN_EQUIV, a, b, tune=TAPE)
    assert _same(got, engine.enum_run(t.EN_EQUIV, a, b))
    for x in got["hit"]:
        assert c.evaluate(x) != r.evaluate(x)
    for part in range(3):
        print("share", part, "of 3", flush=True)
        assert _same(engine.enum_run(t.EN_EQUIV, a, b, part=part, part_end=part, parts=3, tune=TAPE),
                     engine.enum_run(t.EN_EQUIV, a, b, part=part, part_end=part, parts=3))


def test_tape_search(engine):
    for bits, target in ((16, 0x1234), (24, 0xABCDEF)):
        prog = golden("hash%d_5" % bits)
        spec = "h = %d" % target
        print("search", bits, flush=True)
        assert _same(engine.enum_run(t.EN_SEARCH, prog, spec=spec, tune=TAPE),
                     engine.enum_run(t.EN_SEARCH, prog, spec=spec))


def test_tape_hit_list_overflow(engine):
    # every assignment differs: 2^17 hit words overflow the 2^16-entry list, so the host lists
    # the smallest hits again by halves
    rng = np.random.default_rng(8)
    c = N.circuit(22, 1, 120, rng)
    y = c.outputs[0]
    inv = N.Circuit(c.inputs, c.gates + [("y", y, y)], ["y"])
    same = N.Circuit(c.inputs, c.gates + [("y_t", y, y), ("y", "y_t", "y_t")], ["y"])
    a, b = t.net_parse(inv.bench()), t.net_parse(same.bench())
    got = engine.enum_run(t.EN_EQUIV, a, b, tune=TAPE)
    assert got["hits"] == 1 << 22 and got["hit"] == list(range(16))
    assert _same(got, engine.enum_run(t.EN_EQUIV, a, b))


def test_tape_enum_code_reruns(engine):
    rng = np.random.default_rng(2)
    c = N.circuit(20, 8, 2000, rng)
    r = N.fault(N.rewrite(c, rng), rng)
    a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
    code = t.EnumCode(engine, t.EN_EQUIV, a, b, tune=TAPE)
    first = code.run()
    assert _same(first, code.run())
    assert _same(first, engine.enum_run(t.EN_EQUIV, a, b))


@pytest.mark.parametrize("n_in", [5, 7, 9, 10, 11, 13])
def test_tape_partial_rounds(engine, n_in):
    # 1 to 256 words: the last round of a wave has fewer lanes with a word than the 16 that hold
    # a group's descriptors, and every such lane must still run the tape
    rng = np.random.default_rng(300 + n_in)
    c = N.circuit(n_in, 3, 150, rng)
    r = N.fault(N.rewrite(c, rng), rng)
    a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
    print("partial rounds, inputs", n_in, flush=True)
    got = engine.enum_run(t.EN_EQUIV, a, b, tune=TAPE)
    o = O.enum(t.EN_EQUIV, a, b, b_in=b.inputs, b_out=b.outputs)
    assert (got["hits"], got["hit"], got["tried"]) == (o["hits"], o["hit"], o["tried"])
    assert _same(got, engine.enum_run(t.EN_EQUIV, a, b))


@pytest.mark.parametrize("name,spec", [("mul8", "p = a * b"), ("mul8", "p = a * b + (a & b & 1)"),
                                       ("mul10", "p = a * b + 3"), ("adder8", "c:s = a + b")])
def test_tape_verify(engine, name, spec):
    # a verify's spec bit-blasted into the tape: the compiled kernel's hits, count and 16 smallest
    prog = golden(name)
    print("tape verify", name, spec, flush=True)
    got = engine.enum_run(t.EN_VERIFY, prog, spec=spec, tune=TAPE)
    assert _same(got, engine.enum_run(t.EN_VERIFY, prog, spec=spec))
