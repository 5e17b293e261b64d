"""Parity of gpu_enum_run with the reference's enumeration: its known answers
(tests/test_work.c:239-318 in the reference), the oracle's enumeration on small programs, and
closed forms at full size (mul16 over 2^32 with 0 mismatches; toyhash32 preimages)."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def test_adder8_known_answers(engine):
    add = t.gen_adder(8)
    print("adder8 verify", flush=True)
    r = engine.enum_run(t.EN_VERIFY, add, spec="c:s = a + b + c")
    assert r["tried"] == 131072 and r["hits"] == 0 and r["space"] == 131072
    assert r["ops"] == 152 * 131072
    r = engine.enum_run(t.EN_VERIFY, add, spec="c:s = a + b")
    assert r["hits"] == 65536 and r["hit"] == [2 * h + 1 for h in range(16)]
    total = {"tried": 0, "hits": 0, "hit": []}
    for part in range(3):
        s = engine.enum_run(t.EN_VERIFY, add, spec="c:s = a + b", part=part, part_end=part, parts=3)
        total["tried"] += s["tried"]
        total["hits"] += s["hits"]
        total["hit"] = sorted(total["hit"] + s["hit"])[:16]
    assert total["tried"] == 131072 and total["hits"] == 65536
    assert total["hit"] == [2 * h + 1 for h in range(16)]
    a = engine.enum_run(t.EN_VERIFY, add, spec="c:s = a + b", part=0, part_end=1, parts=3)
    b = engine.enum_run(t.EN_VERIFY, add, spec="c:s = a + b", part=2, part_end=2, parts=3)
    assert a["tried"] + b["tried"] == 131072 and a["hits"] + b["hits"] == 65536


@pytest.mark.parametrize("bits", range(1, 9))
def test_small_multipliers(engine, bits):
    r = engine.enum_run(t.EN_VERIFY, t.gen_mul(bits), spec="p = a * b")
    assert r["hits"] == 0 and r["tried"] == 1 << (2 * bits)


def _brute_preimages(bits, rounds, target):
    xs = np.arange(1 << bits, dtype=np.uint64)
    hs = np.array([O.toyhash(int(x), bits, rounds) for x in xs], dtype=np.uint64)
    return xs[hs == target]


def _rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


@pytest.mark.parametrize("bits", [3, 16])
def test_hash_search(engine, bits):
    prog = t.gen_hash(bits, 5)
    target = O.toyhash(12345 % (1 << bits), bits, 5)
    want = _brute_preimages(bits, 5, target)
    r = engine.enum_run(t.EN_SEARCH, prog, spec="h = %d" % target)
    assert r["hits"] == len(want)
    got = sorted(_rev(x, bits) for x in r["hit"])
    assert set(got) <= set(int(x) for x in want)
    # x's bits are reversed against the hash input: the 16 smallest x, mapped back
    assert sorted(r["hit"]) == sorted(_rev(int(x), bits) for x in want)[:16] or len(want) > 16


def test_equiv(engine):
    x1 = t.net_parse("INPUT(a)\nINPUT(b)\nOUTPUT(out)\nout = XOR(a, b)\n")
    x2 = t.net_parse("INPUT(b)\nINPUT(a)\nOUTPUT(out)\nt1 = NAND(a, b)\nt2 = NAND(a, t1)\n"
                     "t3 = NAND(b, t1)\nout = NAND(t2, t3)\n")
    xn = t.net_parse("INPUT(a)\nINPUT(b)\nOUTPUT(out)\nout = XNOR(a, b)\n")
    assert engine.enum_run(t.EN_EQUIV, x1, x2)["hits"] == 0
    r = engine.enum_run(t.EN_EQUIV, x1, xn)
    assert r["hits"] == 4 and r["tried"] == 4
    bad = t.net_parse("INPUT(a)\nINPUT(b)\nOUTPUT(o)\no = XOR(a, b)\n")
    with pytest.raises(ValueError, match="no output called out"):
        engine.enum_run(t.EN_EQUIV, x1, bad)


def test_against_oracle_random(engine):
    rng = np.random.default_rng(5)
    for k in range(6):
        n = int(rng.integers(3, 14))
        prog = t.random_program(12, 200, 1000 + k)
        # make inputs/outputs: a netlist is simpler; use the mul generator shape instead
        mul = t.gen_mul(max(1, n // 2))
        for ticks in (1, 3):
            spec = t.Spec("p = a * b + 1", mul)
            want_tab = np.array([spec.value(x) for x in range(1 << len(mul.inputs))], np.uint64)
            cells = [spec.out_cell(b) for b in range(spec.width)]
            o = O.enum(t.EN_VERIFY, mul, out_cells=cells, expect=want_tab, ticks=ticks)
            g = engine.enum_run(t.EN_VERIFY, mul, spec="p = a * b + 1", ticks=ticks)
            assert (g["hits"], g["hit"], g["tried"]) == (o["hits"], o["hit"], o["tried"])


def test_mul16_exhaustive(engine):
    print("mul16 verify over 2^32", flush=True)
    r = engine.enum_run(t.EN_VERIFY, t.gen_mul(16), spec="p = a * b")
    assert r["tried"] == 1 << 32 and r["hits"] == 0
    assert r["ops"] == 6040 * (1 << 32)


def test_toyhash32_published_target(engine):
    print("toyhash32 search for 0x165c2a07 over 2^32", flush=True)
    r = engine.enum_run(t.EN_SEARCH, t.gen_hash(32, 8), spec="h = 0x165c2a07")
    assert r["tried"] == 1 << 32
    for x in r["hit"]:
        assert O.toyhash(_rev(x, 32), 32, 8) == 0x165c2a07
    assert r["hits"] == len(r["hit"]) and r["hits"] >= 1


def test_mul16_counts_every_assignment(engine):
    # a*b + (a & b & 1) differs from the circuit exactly when a and b are both odd: 2^30 of the
    # 2^32 assignments, so the count proves every assignment was evaluated and compared; a0 is
    # bit 31 of x and b0 bit 15, so the smallest are 2^31 + 2^15 + 0..15
    print("mul16 with a planted spec difference over 2^32", flush=True)
    r = engine.enum_run(t.EN_VERIFY, t.gen_mul(16), spec="p = a * b + (a & b & 1)")
    assert r["tried"] == 1 << 32 and r["hits"] == 1 << 30
    assert r["hit"] == [(1 << 31) + (1 << 15) + k for k in range(16)]


def test_toyhash32_search_random_target(engine):
    rng = np.random.default_rng(11)
    x = int(rng.integers(0, 1 << 32))
    target = O.toyhash(x, 32, 8)
    r = engine.enum_run(t.EN_SEARCH, t.gen_hash(32, 8), spec="h = %d" % target)
    assert _rev(x, 32) in r["hit"] or r["hits"] > 16
    for h in r["hit"]:
        assert O.toyhash(_rev(h, 32), 32, 8) == target


def test_empty_program_enumeration(engine):
    b = t.Builder()
    x = [b.cell("x%d" % i) for i in range(3)]
    for c in x:
        b.input(c)
        b.output(c)
    prog = b.finish()  # no instructions: outputs are the inputs
    r = engine.enum_run(t.EN_VERIFY, prog, spec="x = x")
    assert r["tried"] == 8 and r["hits"] == 0
    for tune in (None, {"tape": 1}):
        r = engine.enum_run(t.EN_SEARCH, prog, spec="x = 5", tune=tune)
        assert r["hits"] == 1 and r["hit"] == [5]
