"""Programs on the 512-machine cube, the front end of the reference's cube demos
(src/cubedemo.c:253-377): a Benes permutation (PERM8 of lib/cube.inc) and a bitonic sort (SORT8)
of one byte per machine, checked against benes_apply and qsort there, and the polynomial product
by the number-theoretic transform (gen_polymul), checked against polymul_scalar. The programs are
the reference's, from program files (default: tests/golden/programs). Many independent cubes run
in one state: cube i is group i.
"""
import numpy as np

from . import BENES_ROUNDS, CUBE_DIMS, benes_route
from .programs import golden


def _bits_to_lines(bits):
    """(cubes, 512) of 0/1 -> (cubes, 8) u64 lines (machine k = bit k % 64 of word k / 64)."""
    return np.packbits(np.asarray(bits, np.uint8), axis=1, bitorder="little").view(np.uint64)


def _lines_to_bits(lines):
    return np.unpackbits(np.ascontiguousarray(lines).view(np.uint8), axis=-1, bitorder="little")


class CubeByteProgram:
    def __init__(self, engine, prog):
        self.prog = prog
        self.code = engine.compile(prog)
        self.v = [prog.cell("v%d" % i) for i in range(8)]

    def _run(self, values, extra=None, repeats=0):
        """One pass over every cube: (bytes afterwards, device ms). With repeats, that many more
        one-pass calls follow on the same state (moving the bytes on again) and the ms is their
        mean, a first call's cold start left out."""
        values = np.asarray(values, np.uint8)
        cubes = values.shape[0]
        mem = np.zeros((cubes, self.prog.ncells, 8), np.uint64)
        rr = np.zeros((cubes, 8), np.uint64)
        for i, c in enumerate(self.v):
            mem[:, c, :] = _bits_to_lines((values >> i) & 1)
        for c, bits in (extra or {}).items():
            mem[:, c, :] = _bits_to_lines(bits)
        st = self.code.state(cubes)
        st.put(mem, rr)
        ms = st.run(1)
        out, _ = st.get()
        if repeats:
            ms = sum(st.run(1) for _ in range(repeats)) / repeats
        res = np.zeros(values.shape, np.uint8)
        for i, c in enumerate(self.v):
            res |= (_lines_to_bits(out[:, c, :]) << i).astype(np.uint8)
        return res, ms


class Perm8(CubeByteProgram):
    """Machine k's byte moves to machine dest[k] through 17 Benes rounds."""

    def __init__(self, engine, prog=None):
        super().__init__(engine, prog if prog is not None else golden("perm8"))
        self.c = [self.prog.cell("c%d" % s) for s in range(BENES_ROUNDS)]

    def run(self, values, dests, repeats=0):
        values = np.atleast_2d(values)
        dests = np.atleast_2d(dests)
        ctrl = np.stack([benes_route(d) for d in dests])  # cubes, 17, 512
        extra = {self.c[s]: ctrl[:, s, :] for s in range(BENES_ROUNDS)}
        return self._run(values, extra, repeats)


class Sort8(CubeByteProgram):
    """Afterwards machine k holds the k-th smallest of the cube's 512 bytes."""

    def __init__(self, engine, prog=None):
        super().__init__(engine, prog if prog is not None else golden("sort8"))

    def run(self, values, repeats=0):
        return self._run(np.atleast_2d(values), repeats=repeats)


class PolyMul:
    """c = a * b in Z_12289[x] / (x^512 + 1) (gen_polymul, 220,758 instructions): machine k of
    cube i holds coefficient k of a[i] and b[i] and its twiddles; one pass leaves c[i] in it
    (the reference's demo runs one pass, src/cubedemo.c:253-281). Further passes give the same
    product: the inputs are only read."""

    BITS = 14

    def __init__(self, engine, tune=None, prog=None):
        self.prog = prog if prog is not None else golden("polymul")
        self._cells()
        self.code = engine.compile(self.prog, tune)

    def _cells(self):
        p = self.prog
        self.a = [p.cell("a%d" % i) for i in range(self.BITS)]
        self.b = [p.cell("b%d" % i) for i in range(self.BITS)]
        self.c = [p.cell("c%d" % i) for i in range(self.BITS)]
        zf, zi = p.extra["zf"], p.extra["zi"]  # the reference's ntt_twiddles, per round
        self.twiddles = {}  # cell -> its line, the same in every cube
        for d in range(CUBE_DIMS):
            for i in range(self.BITS):
                self.twiddles[p.cell("zf%d_%d" % (d, i))] = _bits_to_lines(((zf[d] >> i) & 1)[None])[0]
                self.twiddles[p.cell("zi%d_%d" % (d, i))] = _bits_to_lines(((zi[d] >> i) & 1)[None])[0]

    def lines(self, a, b):
        """The state of len(a) cubes: (cubes, ncells, 8) u64 lines and (cubes, 8) RR lines."""
        a = np.atleast_2d(np.asarray(a, np.uint16))
        b = np.atleast_2d(np.asarray(b, np.uint16))
        cubes = a.shape[0]
        mem = np.zeros((cubes, self.prog.ncells, 8), np.uint64)
        for i in range(self.BITS):
            mem[:, self.a[i], :] = _bits_to_lines((a >> i) & 1)
            mem[:, self.b[i], :] = _bits_to_lines((b >> i) & 1)
        for cell, line in self.twiddles.items():
            mem[:, cell, :] = line
        return mem, np.zeros((cubes, 8), np.uint64)

    def read(self, mem):
        out = np.zeros((mem.shape[0], 512), np.uint16)
        for i in range(self.BITS):
            out |= (_lines_to_bits(mem[:, self.c[i], :]).astype(np.uint16) << i)
        return out

    def run(self, a, b, passes=1):
        """(c, device ms for `passes` passes of every cube)."""
        mem, rr = self.lines(a, b)
        st = self.code.state(mem.shape[0])
        st.put(mem, rr)
        ms = st.run(passes)
        out, _ = st.get()
        return self.read(out), ms

    def run_each(self, a, b, runs):
        """`runs` calls of one pass each on one state (every call computes every product in
        full; within one call of several passes the engine computes the product only in the
        last pass, since nothing reads it before). (c, device ms of each call)."""
        mem, rr = self.lines(a, b)
        st = self.code.state(mem.shape[0])
        st.put(mem, rr)
        ms = [st.run(1) for _ in range(runs)]
        out, _ = st.get()
        return self.read(out), ms
