"""Host side (CPU): generators against the closed-form instruction counts and against the
reference's own generators, specs against the reference's known values and error messages, Beneš
routing, random programs, netlists, and that the library exports every declared C entry point."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    lib = C.CDLL(t.LIB_PATH)
    names = []
    for h in ("tesseract_host.h", "tesseract_gpu.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b((?:ts|gpu)_[a-z0-9_]+)\s*\(", text)
    missing = [n for n in set(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(set(names)) > 50


def test_instruction_counts():
    # closed forms: mul16 = 16*AND6 + 16*BUF4 + 15*(16*AND6 + HA11 + 15*FA19); hash32 =
    # 8*32*17 + 3*popcount(round constants) + 32*4; Life 32x32 tile on 8+8 dims = 1024*123 +
    # 132*57 + 2*296
    assert t.gen_mul(16).len == 16 * 6 + 16 * 4 + 15 * (16 * 6 + 11 + 15 * 19) == 6040
    assert t.gen_hash(32, 8).len == 4861
    life, board = t.gen_life_cube(32, 32, 8, 8)
    assert life.len == 1024 * 123 + 132 * 57 + 2 * 296 == 134068
    assert life.dims == 16 and life.groups == 128
    assert t.gen_adder(8).len == 152
    p, _ = t.gen_life(8, 8)
    assert p.len == 64 * 123


@pytest.mark.parametrize("gen,args", [("gen_mul", (1,)), ("gen_mul", (7,)), ("gen_mul", (16,)),
                                      ("gen_hash", (32, 8)), ("gen_hash", (16, 5)),
                                      ("gen_rank", (9,)), ("gen_rank", (16,)), ("gen_wht", (12,)),
                                      ("random_program", (64, 2000, 5))])
def test_same_program_as_reference(ref, gen, args):
    ours = getattr(t, gen)(*args)
    theirs = ref.make(gen, *args)
    oa, aa = ours.code()
    ob, ab = theirs.code()
    assert ours.len == theirs.s.len and (oa == ob).all() and (aa == ab).all()
    assert [ours.name(i) for i in range(ours.ncells)] == theirs.names()


def test_life_cube_same_as_reference(ref):
    for args in ((32, 32, 8, 8), (4, 4, 3, 3), (8, 6, 5, 6)):
        ours, board = t.gen_life_cube(*args)
        bd = (C.c_uint32 * (args[0] * args[1]))()
        theirs = ref.Prog()
        ref.lib().gen_life_cube(theirs.ptr, *args, C.cast(bd, C.c_void_p))
        oa, aa = ours.code()
        ob, ab = theirs.code()
        assert (oa == ob).all() and (aa == ab).all()
        assert list(board) == list(bd)


def test_xorconv_same_as_reference(ref):
    ours = t.gen_xorconv()
    theirs = ref.make("gen_xorconv")
    assert (ours.code()[1] == theirs.code()[1]).all()


def test_toyhash_agrees(ref):
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 1 << 32, 200, dtype=np.uint64):
        x = int(x)
        assert t.toyhash(x, 32, 8) == O.toyhash(x, 32, 8) == ref.lib().toyhash(x, 32, 8)


def test_spec_values():
    # the reference's known answers (tests/test_work.c:191-210); x: a7..a0 b7..b0 c
    add = t.gen_adder(8)
    cases = [("c:s = a + b + c", (200 << 9) | (100 << 1) | 1, 301), ("s = a * 2 + b", (3 << 9) | (4 << 1), 10),
             ("s = (a + b) * 2", (3 << 9) | (4 << 1), 14), ("s = a - b", (3 << 9) | (4 << 1), 255),
             ("s = ~a & 0xf0 | b ^ 1", (0x0F << 9) | (0x02 << 1), 0xF3), ("s = 1 << 3 >> 1", 0, 4),
             ("c = -c", 1, 1)]
    for text, x, want in cases:
        assert t.Spec(text, add).value(x) == want, text


def test_spec_errors():
    add = t.gen_adder(8)
    bad = [("q = a", "no output called q", True), ("s = z", "no input called z", True),
           ("s = a +", "ends too early", True), ("s = (a", "expected )", True),
           ("s a", "expected =", True), ("s = a", "must be a number", False)]
    for text, msg, inputs in bad:
        with pytest.raises(ValueError, match=re.escape(msg)):
            t.Spec(text, add, inputs)


def test_spec_against_reference(ref):
    add = t.gen_adder(8)
    ra = ref.Prog()
    # the reference needs its own program object: rebuild the adder through its net parser
    net = "\n".join(["INPUT(a%d)" % i for i in range(7, -1, -1)] + ["INPUT(b%d)" % i for i in range(7, -1, -1)]
                    + ["INPUT(c)"] + ["OUTPUT(s%d)" % i for i in range(7, -1, -1)] + ["OUTPUT(cy)"]
                    + ["s%d = XOR(a%d, b%d)" % (i, i, i) for i in range(8)] + ["cy = AND(a0, b0)"])
    err = C.create_string_buffer(256)
    assert ref.lib().net_parse(net.encode(), b"x", ra.ptr, err, 256)
    ours = t.net_parse(net)
    rng = np.random.default_rng(2)
    for text in ("s = a * b + c", "s = (a - b) ^ ~c", "s = a << 3 >> 1 | b", "s = -a * 7 + (b & 0x5a)"):
        s = t.Spec(text, ours)
        rs = ref.lib().spec_parse(text.encode(), ra.ptr, 1, err, 256)
        for x in rng.integers(0, 1 << 17, 100):
            assert s.value(int(x)) == ref.lib().spec_value(rs, int(x))
        ref.lib().spec_free(rs)
    del add


def test_benes_routes_every_permutation():
    rng = np.random.default_rng(3)
    for _ in range(20):
        dest = rng.permutation(512).astype(np.uint16)
        ctrl = t.benes_route(dest)
        for s in range(17):  # both machines of a pair carry the same setting
            d = t.benes_dim(s)
            k = np.arange(512)
            assert (ctrl[s][k] == ctrl[s][k ^ (1 << d)]).all()
        v = rng.integers(0, 1 << 40, 512, dtype=np.uint64)
        out = O.benes_apply(ctrl, v)
        assert (out[dest] == v).all()


def test_random_program_properties():
    a = t.random_program(20, 500, 42)
    b = t.random_program(20, 500, 42)
    c = t.random_program(20, 500, 43)
    ops, addrs = a.code()
    assert a.ncells == 22
    assert (addrs[ops == 1] >= 2).all() and (addrs < 22).all()
    assert 200 < (ops == 0).sum() < 300
    assert (addrs[ops == 0] < 2).any()
    assert (a.code()[1] == b.code()[1]).all() and not (a.code()[1] == c.code()[1]).all()


def test_t1_start_matches_reference_stream():
    # group 0 is the reference's t1_start (src/tester.c:83-91): splitmix64 from seed ^ 0x5DEECE66D
    p = t.random_program(5, 10, 1)
    mem, rr = t.t1_start(p, 99, groups=2)
    s = C.c_uint64(99 ^ 0x5DEECE66D)
    words = [t.lib().ts_splitmix64(C.byref(s)) for _ in range(2 * 8 * 6)]
    assert list(mem[0, 2:, :].reshape(-1)) == words[:40]
    assert list(rr[0]) == words[40:48]
    assert list(mem[1, 2:, :].reshape(-1)) == words[48:88]


def test_netlist_errors_and_dff():
    with pytest.raises(ValueError, match="driven twice"):
        t.net_parse("INPUT(a)\nb = NOT(a)\nb = BUF(a)\n")
    with pytest.raises(ValueError, match="loop without a DFF"):
        t.net_parse("INPUT(a)\nb = AND(a, c)\nc = NOT(b)\n")
    p = t.net_parse("INPUT(a)\nOUTPUT(q)\nq = DFF(d)\nd = XOR(a, q)\n")
    assert p.len > 0


# ISCAS-85 c17: 5 inputs, 2 outputs, 6 NAND gates (the public benchmark circuit)
C17 = """INPUT(1)
INPUT(2)
INPUT(3)
INPUT(6)
INPUT(7)
OUTPUT(22)
OUTPUT(23)
10 = NAND(1, 3)
11 = NAND(3, 6)
16 = NAND(2, 11)
19 = NAND(11, 7)
22 = NAND(10, 16)
23 = NAND(16, 19)
"""


def _ref_net(ref, text):
    p = ref.Prog()
    e = C.create_string_buffer(512)
    assert ref.lib().net_parse(text.encode(), b"net", p.ptr, e, len(e)), e.value
    return p


def test_c17_program():
    # 6 NANDs of 5 instructions each (ROADMAP.md:299: c17.bench = 30 instructions)
    p = t.net_parse(C17)
    assert p.len == 30 and len(p.inputs) == 5 and len(p.outputs) == 2
    # the circuit's function: all 32 assignments in lanes 0..31 of one group on the oracle machine
    mem, rr = O.reset_lines(p, 1)
    for k, cell in enumerate(p.inputs):
        mem[0, cell, 0] = sum(((x >> (4 - k)) & 1) << x for x in range(32))
    mem, rr = O.run_lines(p, mem, rr, 1)
    for x in range(32):
        a = [(x >> (4 - k)) & 1 for k in range(5)]  # inputs 1, 2, 3, 6, 7
        n10, n11 = 1 - (a[0] & a[2]), 1 - (a[2] & a[3])
        n16, n19 = 1 - (a[1] & n11), 1 - (n11 & a[4])
        want = [1 - (n10 & n16), 1 - (n16 & n19)]
        assert [int(mem[0, c, 0]) >> x & 1 for c in p.outputs] == want


def test_netlists_same_program_as_reference(ref):
    from tesseract_amd import netlists as N
    texts = [C17, "INPUT(a)\nOUTPUT(q)\nq = DFF(d)\nd = XOR(a, q)\n",
             "INPUT(x)\nINPUT(y)\nOUTPUT(o)\nOUTPUT(p)\no = NOR(x, y)\np = XNOR(o, y)\n"]
    rng = np.random.default_rng(3)
    for k in range(4):
        c = N.circuit(int(rng.integers(4, 20)), int(rng.integers(1, 9)), int(rng.integers(20, 600)), rng)
        texts += [c.bench(), N.rewrite(c, rng).bench()]
    for text in texts:
        ours, theirs = t.net_parse(text), _ref_net(ref, text)
        oa, aa = ours.code()
        ob, ab = theirs.code()
        assert ours.len == theirs.s.len and (oa == ob).all() and (aa == ab).all()
        assert [ours.name(i) for i in range(ours.ncells)] == theirs.names()


def test_generated_kernels_compile_for_gfx950():
    # hipRTC needs no GPU: every kernel form the engine emits must compile for gfx950
    progs = [t.gen_adder(8), t.gen_rank(12), t.gen_perm8(), t.random_program(30, 300, 1)]
    life, _ = t.gen_life_cube(4, 4, 6, 6)
    progs.append(life)
    # cube programs whose links read cells no pass writes (loads inside cross-lane moves) and
    # cells written before any read (stored as the last pass makes them)
    rng = np.random.default_rng(11)
    for dims in (9, 12):
        b = t.Builder()
        plain = [b.cell("c%d" % i) for i in range(10)]
        ids = [b.cell("ID%d" % d) for d in range(dims)]
        links = [b.link("l%d" % k, (plain + ids)[int(rng.integers(0, 10 + dims))], k % dims)
                 for k in range(3 * dims)]
        for k in range(3 * dims):  # links of links: a move inside the word and across lanes
            links.append(b.link("m%d" % k, links[k], (k + 5) % dims))
        for _ in range(200):
            if rng.random() < 0.55:
                b.nand(([0, 1] + plain + ids + links)[int(rng.integers(0, 12 + 7 * dims))])
            else:
                b.sto(plain[int(rng.integers(0, 6))])
        progs.append(b.finish())
    for p in progs:
        src, info = t.source(p)
        t.rtc_check(src)
    t.rtc_check(t.enum_source(t.EN_VERIFY, t.gen_mul(8), spec="p = a * b"))
    t.rtc_check(t.enum_source(t.EN_VERIFY, t.gen_mul(8), spec="p = a * b", ticks=2))
    t.rtc_check(t.enum_source(t.EN_SEARCH, t.gen_hash(16, 5), spec="h = 77"))
    x = t.net_parse("INPUT(a)\nINPUT(b)\nOUTPUT(o)\no = XOR(a, b)\n")
    t.rtc_check(t.enum_source(t.EN_EQUIV, x, x))


def test_enum_validation_messages():
    big = t.net_parse("\n".join("INPUT(i%d)" % k for k in range(47)) + "\nOUTPUT(o)\no = AND(i0, i1)\n")
    with pytest.raises(ValueError, match="47 inputs is too many"):
        t.enum_source(t.EN_VERIFY, big, spec="o = 1")
    rank = t.gen_rank(9)
    with pytest.raises(ValueError, match="id and link cells connect the lanes"):
        t.enum_source(t.EN_VERIFY, rank, spec="p = x")
    with pytest.raises(ValueError, match="bad --part"):
        t.enum_source(t.EN_VERIFY, t.gen_mul(2), spec="p = a * b", part=2, part_end=1, parts=3)
