"""Game of Life on the hypercube: the front end of the reference's `alpha life --dims D`
(src/work.c:490-701) on the GPU engine.

One torus of (2^rbits * h) x (2^cbits * w) cells, a w x h tile in each of the 2^(rbits+cbits)
machines, tile (R, C) in machine gray(R) << cbits | gray(C) (the reference's grid_machine,
src/gen.c:223-226), one generation per pass of the reference's gen_life_cube program (a program
file carrying its board map and geometry). The start board is the reference's: splitmix64 from
seed 42, 8 words per board cell per group of 512 machines, group by group (src/work.c:588-592).
"""
import numpy as np

from .programs import golden

GAMMA = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB


def splitmix64_stream(seed, n):
    """The first n outputs of splitmix64 started at `seed`, vectorised."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.uint64(GAMMA) * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
        return z ^ (z >> np.uint64(31))


def inverse_gray(g):
    g = np.asarray(g, dtype=np.int64)
    b = g.copy()
    shift = 1
    while shift < 32:
        b ^= b >> shift
        shift <<= 1
    return b


# Slice sizes measured best for a board (AIG nodes per slice). The run time is not smooth in the
# slice size (it depends on which outputs share a slice): 8192 x 8192 on gfx950, us per
# generation at 400 / 425 / 450 / 475 / 500 / 550 / 600 nodes: 6.5 / 6.5 / 5.9 / 6.6 / 6.1 /
# 7.1 / 6.1; the engine's own choice (272 slices) 6.6.
TUNED_SLICE_NODES = {(32, 32, 8, 8): 450}


class LifeCube:
    """prog: a gen_life_cube program file (default: the 8192 x 8192 board, 32 x 32 tiles on
    2^16 machines)."""

    def __init__(self, engine, prog=None, seed=42, tune=None):
        self.prog = prog if prog is not None else golden("life_cube_32_32_8_8")
        w, h, rbits, cbits = (int(v) for v in self.prog.extra["geometry"])
        self.w, self.h, self.rbits, self.cbits = w, h, rbits, cbits
        self.board = np.asarray(self.prog.extra["board"], np.uint32)
        if tune is None and (w, h, rbits, cbits) in TUNED_SLICE_NODES:
            tune = {"slice_nodes": TUNED_SLICE_NODES[(w, h, rbits, cbits)]}
        self.code = engine.compile(self.prog, tune)
        self.groups = self.prog.groups
        self.state = self.code.state(self.groups)
        n = self.prog.ncells
        mem = np.zeros((self.groups, n, 8), np.uint64)
        rr = np.zeros((self.groups, 8), np.uint64)
        cells = w * h
        draws = splitmix64_stream(seed, self.groups * cells * 8).reshape(self.groups, cells, 8)
        mem[:, self.board.astype(np.int64), :] = draws
        self.start_mem = mem
        self.state.put(mem, rr)
        self.generations = 0

    def run(self, gens):
        ms = self.state.run(gens)
        self.generations += gens
        return ms

    def board_bits(self, mem=None):
        """The whole torus as a (2^rbits h, 2^cbits w) array of 0/1."""
        if mem is None:
            mem, _ = self.state.get()
        w, h, rb, cb = self.w, self.h, self.rbits, self.cbits
        tiles = mem[:, self.board.astype(np.int64), :]             # groups, cells, 8 words
        bits = np.unpackbits(tiles.view(np.uint8), axis=2, bitorder="little")  # groups, cells, 512
        per_machine = bits.transpose(0, 2, 1).reshape(-1, h, w)      # machine, r, c
        nm = 1 << (rb + cb)  # machines beyond the grid (a 9-dim cube over a smaller grid) repeat it
        per_machine = per_machine[:nm]
        m = np.arange(nm)
        R = inverse_gray(m >> cb)
        Cc = inverse_gray(m & ((1 << cb) - 1))
        big = np.zeros((1 << rb, h, 1 << cb, w), np.uint8)
        big[R, :, Cc, :] = per_machine
        return big.reshape((1 << rb) * h, (1 << cb) * w)

    def start_bits(self):
        return self.board_bits(self.start_mem)
