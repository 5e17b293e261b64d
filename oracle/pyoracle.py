"""ORACLE bindings — test infrastructure only.

ctypes access to liboracle.so (the C restatement of the reference machine and closed forms, see
oracle.h for how each function is pinned) and, when it was built, to _ref/libalpha_ref.so (the
reference's own sources compiled by oracle/Makefile). Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libalpha_ref.so")

_o = None


class OracleEnumResult(C.Structure):
    _fields_ = [("space", C.c_uint64), ("tried", C.c_uint64), ("hits", C.c_uint64),
                ("hit", C.c_uint64 * 16), ("nhit", C.c_int)]


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(ORACLE_LIB)
        vp, u64 = C.c_void_p, C.c_uint64
        L.oracle_run_lines.argtypes = [vp, C.c_uint, vp, vp, u64]
        L.oracle_run_lines.restype = C.c_int
        L.oracle_reset_lines.argtypes = [vp, C.c_uint, vp, vp]
        L.oracle_random_lines.argtypes = [vp, C.c_uint, u64, vp, vp]
        L.oracle_enum.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, u64, C.c_uint, C.c_int,
                                  C.c_int, C.c_int, C.POINTER(OracleEnumResult)]
        L.oracle_enum.restype = C.c_int
        L.oracle_steps_lines.argtypes = [vp, C.c_uint, vp, vp, u64, u64]
        L.oracle_steps_lines.restype = C.c_int64
        L.oracle_brent_lines.argtypes = [vp, vp, vp, u64, vp]
        L.oracle_brent_lines.restype = u64
        L.oracle_exact_cycle.argtypes = [vp, vp, vp, C.c_uint]
        L.oracle_exact_cycle.restype = u64
        L.oracle_life_scalar.argtypes = [vp, C.c_int, C.c_int]
        L.oracle_benes_apply.argtypes = [vp, vp]
        L.oracle_life_words.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.oracle_toyhash.argtypes = [u64, C.c_int, C.c_int]
        L.oracle_toyhash.restype = u64
        _o = L
    return _o


def _addr(p):
    return C.cast(p.ptr, C.c_void_p)


def reset_lines(prog, groups):
    mem = np.zeros((groups, prog.ncells, 8), np.uint64)
    rr = np.zeros((groups, 8), np.uint64)
    lib().oracle_reset_lines(_addr(prog), groups, mem.ctypes.data, rr.ctypes.data)
    return mem, rr


def random_lines(prog, groups, seed):
    mem = np.zeros((groups, prog.ncells, 8), np.uint64)
    rr = np.zeros((groups, 8), np.uint64)
    lib().oracle_random_lines(_addr(prog), groups, seed, mem.ctypes.data, rr.ctypes.data)
    return mem, rr


def run_lines(prog, mem, rr, passes):
    """Returns new (mem, rr) after `passes` passes on the reference machine."""
    mem = np.array(mem, dtype=np.uint64, order="C")
    rr = np.array(rr, dtype=np.uint64, order="C")
    if not lib().oracle_run_lines(_addr(prog), rr.shape[0], mem.ctypes.data, rr.ctypes.data, passes):
        raise ValueError("groups not a multiple of the cube")
    return mem, rr


def enum(mode, a, b=None, b_in=None, b_out=None, out_cells=None, expect=None, target=0, ticks=1,
         part=0, part_end=0, parts=1):
    r = OracleEnumResult()
    arr = lambda x, t: None if x is None else np.ascontiguousarray(x, dtype=t)
    b_in, b_out = arr(b_in, np.uint32), arr(b_out, np.uint32)
    out_cells, expect = arr(out_cells, np.uint32), arr(expect, np.uint64)
    ptr = lambda x: None if x is None else x.ctypes.data
    ok = lib().oracle_enum(mode, _addr(a), _addr(b) if b is not None else None, ptr(b_in), ptr(b_out),
                           ptr(out_cells), 0 if out_cells is None else len(out_cells), ptr(expect),
                           target, ticks, part, part_end, parts, C.byref(r))
    if not ok:
        raise ValueError("oracle_enum refused the program")
    return {"space": r.space, "tried": r.tried, "hits": r.hits, "hit": list(r.hit[:r.nhit]), "nhit": r.nhit}


def steps_lines(prog, mem, rr, pc, steps):
    """(mem, rr, pc) after `steps` instructions from instruction pc on the reference machine."""
    mem = np.array(mem, dtype=np.uint64, order="C")
    rr = np.array(rr, dtype=np.uint64, order="C")
    pc = lib().oracle_steps_lines(_addr(prog), rr.shape[0], mem.ctypes.data, rr.ctypes.data, pc, steps)
    if pc < 0:
        raise ValueError("groups not a multiple of the cube, or pc past the program")
    return mem, rr, int(pc)


def brent_lines(prog, mem, rr, max_passes, with_passes=False):
    """Cycle length per lane of one 512-machine group (0 = none within max_passes); with
    with_passes also the passes the group ran (the reference's t1_cycles return value)."""
    lam = np.zeros(512, np.uint64)
    mem = np.ascontiguousarray(mem, dtype=np.uint64)
    rr = np.ascontiguousarray(rr, dtype=np.uint64)
    passes = lib().oracle_brent_lines(_addr(prog), mem.ctypes.data, rr.ctypes.data, max_passes,
                                      lam.ctypes.data)
    return (lam, int(passes)) if with_passes else lam


def exact_cycle(prog, mem, rr, lane):
    mem = np.ascontiguousarray(mem, dtype=np.uint64)
    rr = np.ascontiguousarray(rr, dtype=np.uint64)
    return int(lib().oracle_exact_cycle(_addr(prog), mem.ctypes.data, rr.ctypes.data, lane))


def life_scalar(board, gens=1):
    b = np.ascontiguousarray(board, dtype=np.uint8).copy()
    h, w = b.shape
    for _ in range(gens):
        lib().oracle_life_scalar(b.ctypes.data, w, h)
    return b


def benes_apply(ctrl, v):
    c = np.ascontiguousarray(ctrl, dtype=np.uint8)
    x = np.ascontiguousarray(v, dtype=np.uint64).copy()
    lib().oracle_benes_apply(c.ctypes.data, x.ctypes.data)
    return x


NTT_Q = 12289


def polymul(a, b):
    """c = a * b in Z_12289[x] / (x^512 + 1), coefficient k at index k of the last axis: the
    schoolbook product folded with x^512 = -1, as the reference's polymul_scalar
    (src/gen.c:671-684). a, b: (..., 512) integers in 0 .. 12288. Pinned in tests against that
    function and against the ring's identities (1 is the unit, x shifts negacyclically)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    lead = np.broadcast_shapes(a.shape[:-1], b.shape[:-1])
    a = np.broadcast_to(a, lead + (512,)).reshape(-1, 512)
    b = np.broadcast_to(b, lead + (512,)).reshape(-1, 512)
    out = np.empty(a.shape, np.int64)
    for i in range(a.shape[0]):
        full = np.convolve(a[i], b[i])  # 1,023 terms, each below 512 q^2 < 2^63
        low = full[:512].copy()
        low[:511] -= full[512:]
        out[i] = low % NTT_Q
    return out.reshape(lead + (512,)).astype(np.uint16)


def toyhash(x, bits, rounds):
    return int(lib().oracle_toyhash(x, bits, rounds))


def life_words(board_bits, gens):
    """B3/S23 on a torus, bit-parallel: board_bits is (h, w) of 0/1 with w a multiple of 64."""
    h, w = board_bits.shape
    words = np.packbits(np.ascontiguousarray(board_bits, dtype=np.uint8), axis=1,
                        bitorder="little").view(np.uint64).copy()
    lib().oracle_life_words(words.ctypes.data, w // 64, h, gens)
    return np.unpackbits(words.view(np.uint8), axis=1, bitorder="little")[:, :w]
