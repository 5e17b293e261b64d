"""The NTT polynomial product on the 512-machine cube (reference gen_polymul, src/gen.c:456-669):
the program file holds the reference's program instruction for instruction and its twiddle table,
the twiddles satisfy their closed form, and the oracle's product is the reference's polymul_scalar
and obeys the ring's identities. The program itself is run through the oracle's reference machine
on one cube, and on the GPU on many."""
import ctypes as C

import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tests.helpers import golden

Q = 12289


def test_polymul_program_is_the_references(ref):
    ours = golden("polymul")
    theirs = ref.make("gen_polymul")
    assert ours.len == theirs.s.len == 220758  # the count the reference quotes for this program
    oa, aa = ours.code()
    ob, ab = theirs.code()
    assert (oa == ob).all() and (aa == ab).all()
    assert [ours.name(i) for i in range(ours.ncells)] == theirs.names()
    assert ours.dims == 9


def _twiddles():
    p = golden("polymul")
    return p.extra["zf"], p.extra["zi"]


def test_twiddles_are_the_references(ref):
    zf, zi = _twiddles()
    rf = np.zeros((9, 512), np.uint16)
    ri = np.zeros((9, 512), np.uint16)
    ref.lib().ntt_twiddles(rf.ctypes.data, ri.ctypes.data)
    assert (zf == rf).all() and (zi == ri).all()


def test_twiddles_closed_form():
    # in Montgomery form: z * 2^14 mod q. The machines with bit d clear multiply by 1 on the way
    # back; psi^512 = -1 makes the round-8 forward twiddle of every machine psi^256 (a 4th root
    # of 1 squared to -1).
    zf, zi = _twiddles()
    mont = (1 << 14) % Q
    minv = pow(mont, Q - 2, Q)
    for d in range(9):
        for k in range(512):
            z = int(zf[d, k]) * minv % Q
            if (k >> d) & 1:
                assert int(zi[d, k]) * minv % Q * z % Q == 1
            else:
                assert zi[d, k] == mont
    psi256 = int(zf[8, 0]) * minv % Q
    assert psi256 * psi256 % Q == Q - 1 and (zf[8] == zf[8, 0]).all()


def test_oracle_product_is_polymul_scalar(ref):
    rng = np.random.default_rng(5)
    for _ in range(4):
        a = rng.integers(0, Q, 512).astype(np.uint16)
        b = rng.integers(0, Q, 512).astype(np.uint16)
        want = np.zeros(512, np.uint16)
        ref.lib().polymul_scalar(a.ctypes.data, b.ctypes.data, want.ctypes.data)
        assert (O.polymul(a, b) == want).all()


def test_oracle_product_identities():
    rng = np.random.default_rng(6)
    a = rng.integers(0, Q, (3, 512))
    b = rng.integers(0, Q, (3, 512))
    one = np.zeros(512, np.int64)
    one[0] = 1
    x = np.zeros(512, np.int64)
    x[1] = 1
    assert (O.polymul(a, one) == a).all()
    shifted = np.concatenate([(Q - a[:, -1:]) % Q, a[:, :-1]], axis=1)  # x * a, x^512 = -1
    assert (O.polymul(a, x) == shifted).all()
    assert (O.polymul(a, b) == O.polymul(b, a)).all()
    s = O.polymul((a + b) % Q, a)
    assert (s == (O.polymul(a, a).astype(np.int64) + O.polymul(b, a)) % Q).all()


def test_program_computes_the_product_on_the_reference_machine():
    # one cube on the oracle's machine: inputs per machine, one pass, c against the product
    from tesseract_amd.cube import PolyMul
    rng = np.random.default_rng(7)
    a = rng.integers(0, Q, (1, 512)).astype(np.uint16)
    b = rng.integers(0, Q, (1, 512)).astype(np.uint16)
    pm = PolyMul.__new__(PolyMul)
    pm.prog = golden("polymul")
    pm._cells()
    mem, rr = pm.lines(a, b)
    out, _ = O.run_lines(pm.prog, mem, rr, 1)
    assert (pm.read(out) == O.polymul(a, b)).all()


def test_reference_jit_computes_the_product(ref):
    # the CPU baseline's path: the reference's own compiled code on one cube gives the product
    from tesseract_amd.cube import PolyMul
    rng = np.random.default_rng(8)
    a = rng.integers(0, Q, (1, 512)).astype(np.uint16)
    b = rng.integers(0, Q, (1, 512)).astype(np.uint16)
    pm = PolyMul.__new__(PolyMul)
    pm.prog = golden("polymul")
    pm._cells()
    mem, _ = pm.lines(a, b)
    theirs = ref.make("gen_polymul")
    inputs = pm.a + pm.b + list(pm.twiddles)
    got, secs = ref.jit_run_cells(theirs.ptr, {c: mem[0, c] for c in inputs}, 2, pm.c)
    out = np.zeros((1, pm.prog.ncells, 8), np.uint64)
    for c, line in got.items():
        out[0, c] = line
    assert (pm.read(out) == O.polymul(a, b)).all() and secs > 0


@pytest.fixture(scope="module")
def polymul(engine):
    from tesseract_amd.cube import PolyMul
    print("compiling gen_polymul (220,758 instructions)", flush=True)
    return PolyMul(engine)


@pytest.mark.gpu
def test_polymul_gpu(polymul):
    rng = np.random.default_rng(9)
    cubes = 64
    a = rng.integers(0, Q, (cubes, 512)).astype(np.uint16)
    b = rng.integers(0, Q, (cubes, 512)).astype(np.uint16)
    a[0] = 0                      # the zero polynomial
    a[1] = 0
    a[1, 0] = 1                   # 1: c = b
    a[2] = 0
    a[2, 1] = 1                   # x: b shifted, x^512 = -1
    a[3] = Q - 1                  # every coefficient -1
    b[4] = Q - 1
    print("polymul on %d cubes" % cubes, flush=True)
    c, ms = polymul.run(a, b)
    assert (c == O.polymul(a, b)).all()
    assert (c[0] == 0).all() and (c[1] == b[1]).all()
    assert (c[2, 1:] == b[2, :-1]).all() and c[2, 0] == (Q - b[2, -1]) % Q


@pytest.mark.gpu
def test_polymul_gpu_repeated_passes(polymul):
    # the inputs are only read: three passes leave the same product; one cube and an odd count
    rng = np.random.default_rng(10)
    for cubes in (1, 3):
        a = rng.integers(0, Q, (cubes, 512)).astype(np.uint16)
        b = rng.integers(0, Q, (cubes, 512)).astype(np.uint16)
        c, _ = polymul.run(a, b, passes=3)
        assert (c == O.polymul(a, b)).all()


@pytest.mark.gpu
def test_polymul_gpu_calls_on_one_state(polymul):
    # one-pass calls one after another on the same state: each computes the product in full
    rng = np.random.default_rng(11)
    a = rng.integers(0, Q, (5, 512)).astype(np.uint16)
    b = rng.integers(0, Q, (5, 512)).astype(np.uint16)
    c, ms = polymul.run_each(a, b, 4)
    assert (c == O.polymul(a, b)).all() and len(ms) == 4
