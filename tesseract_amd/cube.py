"""Data movement on the 512-machine cube: a Benes permutation (PERM8) and a bitonic sort (SORT8) of
one byte per machine, the front end of the reference's cube demos (src/cubedemo.c:352-377 check
them against benes_apply and qsort). Many independent cubes run in one state: cube i is group i.
"""
import numpy as np

from . import BENES_ROUNDS, benes_route, gen_perm8, gen_sort8


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

    def _run(self, values, extra=None):
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
        res = np.zeros(values.shape, np.uint8)
        for i, c in enumerate(self.v):
            res |= (_lines_to_bits(out[:, c, :]) << i).astype(np.uint8)
        return res, ms


class Perm8(CubeByteProgram):
    """Machine k's byte moves to machine dest[k] through 17 Benes rounds."""

    def __init__(self, engine):
        super().__init__(engine, gen_perm8())
        self.c = [self.prog.cell("c%d" % s) for s in range(BENES_ROUNDS)]

    def run(self, values, dests):
        values = np.atleast_2d(values)
        dests = np.atleast_2d(dests)
        ctrl = np.stack([benes_route(d) for d in dests])  # cubes, 17, 512
        extra = {self.c[s]: ctrl[:, s, :] for s in range(BENES_ROUNDS)}
        return self._run(values, extra)


class Sort8(CubeByteProgram):
    """Afterwards machine k holds the k-th smallest of the cube's 512 bytes."""

    def __init__(self, engine):
        super().__init__(engine, gen_sort8())

    def run(self, values):
        return self._run(np.atleast_2d(values))
