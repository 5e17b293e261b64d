"""Parity of gpu_enum_run with the reference's enumeration: its known answers
(tests/test_work.c:239-318 in the reference), the oracle's enumeration on small programs, and
closed forms at full size (mul16 over 2^32 with 0 mismatches; toyhash32 preimages)."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tests.helpers import golden

pytestmark = pytest.mark.gpu


def test_adder8_known_answers(engine):
    add = golden("adder8")
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
    r = engine.enum_run(t.EN_VERIFY, golden("mul%d" % bits), spec="p = a * b")
    assert r["hits"] == 0 and r["tried"] == 1 << (2 * bits)


def _brute_preimages(bits, rounds, target):
    xs = np.arange(1 << bits, dtype=np.uint64)
    hs = np.array([O.toyhash(int(x), bits, rounds) for x in xs], dtype=np.uint64)
    return xs[hs == target]


def _rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


@pytest.mark.parametrize("bits", [3, 16])
def test_hash_search(engine, bits):
    prog = golden("hash%d_5" % bits)
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


def _random_io_program(rng, n_in, n_out, cells, length):
    """A random NAND/STO program with n_in declared inputs and n_out outputs: inputs are read and
    may be overwritten, outputs are plain cells written somewhere in the pass."""
    b = t.Builder()
    ins = [b.cell("x%d" % i) for i in range(n_in)]
    work = [b.cell("w%d" % i) for i in range(cells)]
    outs = [b.cell("o%d" % i) for i in range(n_out)]
    readable = [0, 1] + ins + work + outs
    writable = ins + work + outs
    for c in ins:
        b.input(c)
    for _ in range(length):
        if rng.random() < 0.55:
            b.nand(readable[int(rng.integers(0, len(readable)))])
        else:
            b.sto(writable[int(rng.integers(0, len(writable)))])
    for c in outs:
        b.output(c)
    return b.finish()


def test_random_programs_against_oracle(engine):
    # random programs with inputs and outputs, enumerated exhaustively on the device and by the
    # oracle's reference machine: verify against a spec, equiv of two programs, search for a
    # value; one and three ticks (state carried between ticks), whole and in shares
    rng = np.random.default_rng(5)
    for k in range(8):
        n = int(rng.integers(3, 14))
        nout = int(rng.integers(1, 6))
        a = _random_io_program(rng, n, nout, int(rng.integers(2, 30)), int(rng.integers(20, 400)))
        bb = _random_io_program(rng, n, nout, int(rng.integers(2, 30)), int(rng.integers(20, 400)))
        text = "o = x * 3 + 1"
        spec = t.Spec(text, a)
        want_tab = np.array([spec.value(x) for x in range(1 << n)], np.uint64)
        cells = [spec.out_cell(i) for i in range(spec.width)]
        for ticks in (1, 3):
            o = O.enum(t.EN_VERIFY, a, out_cells=cells, expect=want_tab, ticks=ticks)
            g = engine.enum_run(t.EN_VERIFY, a, spec=text, ticks=ticks)
            assert (g["hits"], g["hit"], g["tried"]) == (o["hits"], o["hit"], o["tried"]), (k, ticks)
            o = O.enum(t.EN_EQUIV, a, bb, b_in=bb.inputs, b_out=bb.outputs, ticks=ticks)
            g = engine.enum_run(t.EN_EQUIV, a, bb, ticks=ticks)
            assert (g["hits"], g["hit"], g["tried"]) == (o["hits"], o["hit"], o["tried"]), (k, ticks)
        target = int(rng.integers(0, 1 << nout))
        o = O.enum(t.EN_SEARCH, a, out_cells=cells, target=target)
        g = engine.enum_run(t.EN_SEARCH, a, spec="o = %d" % target)
        assert (g["hits"], g["hit"]) == (o["hits"], o["hit"]), k
        parts = [engine.enum_run(t.EN_VERIFY, a, spec=text, part=q, part_end=q, parts=3)
                 for q in range(3)]
        o = O.enum(t.EN_VERIFY, a, out_cells=cells, expect=want_tab)
        assert sum(p["hits"] for p in parts) == o["hits"]
        assert sorted(sum((p["hit"] for p in parts), []))[:16] == o["hit"]


def test_mul16_exhaustive(engine):
    print("mul16 verify over 2^32", flush=True)
    r = engine.enum_run(t.EN_VERIFY, golden("mul16"), spec="p = a * b")
    assert r["tried"] == 1 << 32 and r["hits"] == 0
    assert r["ops"] == 6040 * (1 << 32)


def test_toyhash32_published_target(engine, ref):
    # the published answer: exactly one preimage of 0x165c2a07 over 2^32 (README.md:67); the
    # reference's own enum_run on the same program gives the same count and hit list
    print("toyhash32 search for 0x165c2a07 over 2^32", flush=True)
    prog = golden("hash32_8")
    r = engine.enum_run(t.EN_SEARCH, prog, spec="h = 0x165c2a07")
    assert r["tried"] == 1 << 32
    assert r["hits"] == 1 and len(r["hit"]) == 1
    assert O.toyhash(_rev(r["hit"][0], 32), 32, 8) == 0x165c2a07
    print("the reference's enum_run over 2^32", flush=True)
    w = ref.enum_run(t.EN_SEARCH, prog.ptr, spec_text="h = 0x165c2a07", spec_inputs=False, threads=16)
    assert (r["tried"], r["hits"], r["hit"]) == (w["tried"], w["hits"], w["hit"])


def test_mul16_counts_every_assignment(engine):
    # a*b + (a & b & 1) differs from the circuit exactly when a and b are both odd: 2^30 of the
    # 2^32 assignments, so the count proves every assignment was evaluated and compared; a0 is
    # bit 31 of x and b0 bit 15, so the smallest are 2^31 + 2^15 + 0..15
    print("mul16 with a planted spec difference over 2^32", flush=True)
    r = engine.enum_run(t.EN_VERIFY, golden("mul16"), spec="p = a * b + (a & b & 1)")
    assert r["tried"] == 1 << 32 and r["hits"] == 1 << 30
    assert r["hit"] == [(1 << 31) + (1 << 15) + k for k in range(16)]


def test_enumerations_refuse_iterative_ilp(engine):
    # the compiler itself faults on some enumeration kernels scheduled iterative-ilp (this one
    # among them): the compile refuses the strategy with a message, and the timed choice of long
    # enumerations takes 1 or 3
    with pytest.raises((ValueError, RuntimeError), match="iterative-ilp"):
        engine.enum_run(t.EN_VERIFY, golden("mul16"), spec="p = a * b + (a & 1)", tune={"sched": 2})


def test_mul16_difference_in_group_compared_bits(engine):
    # a * b + (a & 1) keeps the linear form: the difference is in bit 0 (compared once per group
    # of 32 words, no lane and no word of a group reaching it) and above, wherever a is odd (x bit
    # 31): 2^31 assignments, the smallest 2^31 + 0..15
    print("mul16 against a * b + (a & 1) over 2^32", flush=True)
    r = engine.enum_run(t.EN_VERIFY, golden("mul16"), spec="p = a * b + (a & 1)")
    assert r["tried"] == 1 << 32 and r["hits"] == 1 << 31
    assert r["hit"] == [(1 << 31) + k for k in range(16)]


def test_toyhash32_search_random_target(engine, ref):
    # a seeded target with a known preimage: every hit maps back to the target, and the count
    # and the 16 smallest equal the reference's enum_run over all 2^32 assignments
    rng = np.random.default_rng(11)
    x = int(rng.integers(0, 1 << 32))
    target = O.toyhash(x, 32, 8)
    prog = golden("hash32_8")
    r = engine.enum_run(t.EN_SEARCH, prog, spec="h = %d" % target)
    assert _rev(x, 32) in r["hit"] or r["hits"] > 16
    for h in r["hit"]:
        assert O.toyhash(_rev(h, 32), 32, 8) == target
    w = ref.enum_run(t.EN_SEARCH, prog.ptr, spec_text="h = %d" % target, spec_inputs=False,
                     threads=16)
    assert (r["tried"], r["hits"], r["hit"]) == (w["tried"], w["hits"], w["hit"])
    assert r["hits"] >= 1


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
