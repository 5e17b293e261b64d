"""The oracle, pinned: against the reference's own build (oracle/_ref), the reference's known
answers, exhaustive state enumeration, and closed forms chosen so a dropped term or a wrong index
fails."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tests.helpers import golden


def _cube_prog(cells, length, dims, seed):
    from tests.test_gpu_run import random_cube_program
    return random_cube_program(cells, length, dims, seed)


@pytest.mark.parametrize("dims,seed", [(9, 1), (9, 2), (10, 3)])
def test_run_lines_against_reference_cube(ref, dims, seed):
    prog = _cube_prog(12, 150, dims, seed)
    mem, rr = O.random_lines(prog, prog.groups, seed)
    want = ref.cube_run(ref_ptr(prog), prog.ncells, mem, rr, 3)
    got = O.run_lines(prog, mem, rr, 3)
    writable = [c for c in range(2, prog.ncells) if prog.kind(c) == t.CK_PLAIN]
    assert np.array_equal(got[1], want[1])
    assert np.array_equal(got[0][:, writable], want[0][:, writable])


@pytest.mark.parametrize("dims,seed", [(9, 4), (10, 5)])
def test_steps_lines_against_reference_cube(ref, dims, seed):
    # single instructions from pc 0 (ref_cube_step), into a second pass, against the reference
    prog = _cube_prog(12, 150, dims, seed)
    mem, rr = O.random_lines(prog, prog.groups, seed)
    writable = [c for c in range(2, prog.ncells) if prog.kind(c) == t.CK_PLAIN]
    for steps in (1, 37, 150 + 61):
        want = ref.cube_run(ref_ptr(prog), prog.ncells, mem, rr, 0, steps)
        gm, gr, pc = O.steps_lines(prog, mem, rr, 0, steps)
        assert pc == steps % 150
        assert np.array_equal(gr, want[1])
        assert np.array_equal(gm[:, writable], want[0][:, writable])


def test_steps_lines_compose():
    # steps from any pc compose: a + b steps = a steps then b steps; whole passes = run_lines
    prog = _cube_prog(10, 90, 9, 6)
    mem, rr = O.random_lines(prog, prog.groups, 6)
    m1, r1, pc1 = O.steps_lines(prog, mem, rr, 0, 53)
    m2, r2, pc2 = O.steps_lines(prog, m1, r1, pc1, 200)
    m3, r3, pc3 = O.steps_lines(prog, mem, rr, 0, 253)
    assert pc2 == pc3 == 253 % 90 and np.array_equal(m2, m3) and np.array_equal(r2, r3)
    m4, r4, pc4 = O.steps_lines(prog, mem, rr, 0, 2 * 90)
    w4, wr4 = O.run_lines(prog, mem, rr, 2)
    assert pc4 == 0 and np.array_equal(m4, w4) and np.array_equal(r4, wr4)


def ref_ptr(prog):
    import ctypes as C
    return C.cast(prog.ptr, C.c_void_p)


def test_run_lines_adder_closed_form():
    add = golden("adder8")
    mem, rr = O.reset_lines(add, 1)
    rng = np.random.default_rng(4)
    a = rng.integers(0, 256, 512)
    b = rng.integers(0, 256, 512)
    c = rng.integers(0, 2, 512)
    bits = lambda v: np.packbits(v.astype(np.uint8), bitorder="little").view(np.uint64)
    for i in range(8):
        mem[0, add.cell("a%d" % i)] = bits((a >> i) & 1)
        mem[0, add.cell("b%d" % i)] = bits((b >> i) & 1)
    mem[0, add.cell("c")] = bits(c)
    m2, _ = O.run_lines(add, mem, rr, 1)
    unbits = lambda line: np.unpackbits(line.view(np.uint8), bitorder="little")
    s = sum(unbits(m2[0, add.cell("s%d" % i)]).astype(np.int64) << i for i in range(8))
    s += unbits(m2[0, add.cell("c")]).astype(np.int64) << 8
    assert (s == a + b + c).all()


def test_enum_known_answers():
    add = golden("adder8")
    spec = t.Spec("c:s = a + b + c", add)
    cells = [spec.out_cell(i) for i in range(spec.width)]
    x = np.arange(1 << 17, dtype=np.uint64)
    a, b, c = (x >> 9) & 255, (x >> 1) & 255, x & 1
    r = O.enum(t.EN_VERIFY, add, out_cells=cells, expect=a + b + c)
    assert r["tried"] == 131072 and r["hits"] == 0
    r = O.enum(t.EN_VERIFY, add, out_cells=cells, expect=a + b)
    assert r["hits"] == 65536 and r["hit"] == [2 * h + 1 for h in range(16)]
    parts = [O.enum(t.EN_VERIFY, add, out_cells=cells, expect=a + b, part=k, part_end=k, parts=3)
             for k in range(3)]
    assert sum(p["tried"] for p in parts) == 131072 and sum(p["hits"] for p in parts) == 65536


def test_enum_against_reference(ref):
    mul = golden("mul6")
    theirs = ref.make("gen_mul", 6)
    for text in ("p = a * b", "p = a * b + 1", "p = a + b"):
        spec = t.Spec(text, mul)
        tab = np.array([spec.value(x) for x in range(1 << 12)], np.uint64)
        o = O.enum(t.EN_VERIFY, mul, out_cells=[spec.out_cell(i) for i in range(spec.width)], expect=tab)
        r = ref.enum_run(t.EN_VERIFY, theirs.ptr, spec_text=text, threads=2)
        assert (o["tried"], o["hits"], o["hit"]) == (r["tried"], r["hits"], r["hit"])
        r3 = ref.enum_run(t.EN_VERIFY, theirs.ptr, spec_text=text, part=1, part_end=1, parts=1, threads=1) \
            if False else None


def test_brent_equals_exact_cycle():
    for k in range(30):
        cells = 1 + k % 10
        prog = t.random_program(cells, 1 + (k * 7919) % 60, 1000 + k)
        mem, rr = t.t1_start(prog, 1000 + k)
        lam = O.brent_lines(prog, mem[0], rr[0], 1 << 16)
        for lane in range(0, 512, 37):
            assert lam[lane] == O.exact_cycle(prog, mem[0], rr[0], lane)


def test_brent_against_reference(ref):
    import ctypes as C
    for k in range(4):
        prog = t.random_program(40 + 10 * k, 120, 500 + k)
        theirs = ref.make("random_program", 40 + 10 * k, 120, 500 + k)
        lam = np.zeros(512, np.uint64)
        aborted = C.c_int()
        steps = ref.lib().t1_cycles(theirs.ptr, 500 + k, 4096, lam.ctypes.data, None,
                                    C.byref(aborted))
        mem, rr = t.t1_start(prog, 500 + k)
        ours, passes = O.brent_lines(prog, mem[0], rr[0], 4096, with_passes=True)
        assert (ours == lam).all()
        assert passes == steps  # the passes the group ran before every machine found its cycle


def test_life_scalar_patterns():
    b = np.zeros((8, 8), np.uint8)
    b[3, 2:5] = 1  # blinker: period 2
    b1 = O.life_scalar(b)
    assert b1[2:5, 3].all() and b1.sum() == 3
    assert (O.life_scalar(b, 2) == b).all()
    blk = np.zeros((6, 6), np.uint8)
    blk[1:3, 1:3] = 1
    assert (O.life_scalar(blk) == blk).all()
    g = np.zeros((10, 10), np.uint8)
    g[0, 1] = g[1, 2] = g[2, 0] = g[2, 1] = g[2, 2] = 1  # glider: moves (1, 1) every 4
    assert (O.life_scalar(g, 4) == np.roll(np.roll(g, 1, 0), 1, 1)).all()
    tor = np.zeros((5, 5), np.uint8)
    tor[0, 0] = tor[0, 4] = tor[4, 0] = 1  # three corners touch across the wrap: a birth at (4, 4)
    assert O.life_scalar(tor)[4, 4] == 1


def test_toyhash_published_target_form():
    # one round by hand for 3 bits: m = 5, permutation i -> (5 i + 3) mod 3
    x = 0b101
    y = 0
    for i in range(3):
        xi, x1, x2 = x >> i & 1, x >> (i + 1) % 3 & 1, x >> (i + 2) % 3 & 1
        y |= (xi ^ ((1 - x1) & x2)) << i
    z = sum(((y >> (5 * i + 3) % 3) & 1) << i for i in range(3))
    # toyhash(0) after one round = the round constant (all y bits are 0 for x = 0)
    assert O.toyhash(x, 3, 1) == z ^ O.toyhash(0, 3, 1)


@pytest.mark.parametrize("bits", range(1, 8))
def test_mul_programs_multiply(bits):
    # the reference's gen_mul programs (program files) on the oracle machine: p = a * b for every
    # assignment (a0 is input 0, the highest bit of x; b follows)
    mul = golden("mul%d" % bits)
    spec = t.Spec("p = a * b", mul)
    n = 2 * bits
    x = np.arange(1 << n, dtype=np.uint64)
    a = sum(((x >> np.uint64(n - 1 - i)) & np.uint64(1)) << np.uint64(i) for i in range(bits))
    b = sum(((x >> np.uint64(bits - 1 - i)) & np.uint64(1)) << np.uint64(i) for i in range(bits))
    assert all(spec.value(int(v)) == int(a[v] * b[v]) for v in range(0, 1 << n, max(1, (1 << n) // 300)))
    r = O.enum(t.EN_VERIFY, mul, out_cells=[spec.out_cell(i) for i in range(spec.width)], expect=a * b)
    assert r["hits"] == 0 and r["tried"] == 1 << n


@pytest.mark.parametrize("name,bits,rounds", [("hash12_4", 12, 4), ("hash16_5", 16, 5),
                                              ("hash32_8", 32, 8)])
def test_hash_programs_compute_toyhash(name, bits, rounds):
    # the reference's gen_hash programs on the oracle machine against the closed form, at random
    # inputs: input k (h's x bit k) is x's bit k, so machine value v hashes to toyhash(v)
    prog = golden(name)
    rng = np.random.default_rng(bits)
    v = rng.integers(0, 1 << bits, 512, dtype=np.uint64)
    mem, rr = O.reset_lines(prog, 1)
    for k, cell in enumerate(prog.inputs):
        mem[0, cell] = np.packbits(((v >> np.uint64(k)) & np.uint64(1)).astype(np.uint8),
                                   bitorder="little").view(np.uint64)
    mem, rr = O.run_lines(prog, mem, rr, 1)
    h = np.zeros(512, np.uint64)
    for k, cell in enumerate(prog.outputs):
        h |= np.unpackbits(mem[0, cell].view(np.uint8), bitorder="little").astype(np.uint64) << np.uint64(k)
    assert all(int(h[i]) == O.toyhash(int(v[i]), bits, rounds) for i in range(512))


@pytest.mark.parametrize("name", ["cube_any", "cube_rank", "cube_max", "rule110", "perm8"])
def test_run_lines_on_reference_examples(ref, name):
    # the oracle machine on the reference's own cube programs against its ref_cube_tick
    prog = golden(name)
    mem, rr = O.random_lines(prog, prog.groups, 3)
    want = ref.cube_run(ref_ptr(prog), prog.ncells, mem, rr, 2)
    got = O.run_lines(prog, mem, rr, 2)
    writable = [c for c in range(2, prog.ncells) if prog.kind(c) == t.CK_PLAIN]
    assert np.array_equal(got[1], want[1])
    assert np.array_equal(got[0][:, writable], want[0][:, writable])
