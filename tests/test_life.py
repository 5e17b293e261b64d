"""Game of Life on the cube (the reference's `alpha life --dims`): the whole torus after G
generations equals plain-C B3/S23 on the same start board (src/work.c:601-622)."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tesseract_amd.life import LifeCube, splitmix64_stream
from tests.helpers import life_cube


class _CpuLife(LifeCube):
    """The same front end with the oracle machine in place of the GPU state (CPU test only)."""

    def __init__(self, w, h, rbits, cbits, seed=42):
        self.w, self.h, self.rbits, self.cbits = w, h, rbits, cbits
        self.prog, self.board = life_cube(w, h, rbits, cbits)
        self.groups = self.prog.groups
        mem, self.rr = O.reset_lines(self.prog, self.groups)
        draws = splitmix64_stream(seed, self.groups * w * h * 8).reshape(self.groups, w * h, 8)
        mem[:, self.board.astype(np.int64), :] = draws
        self.mem = self.start_mem = mem

    def run(self, gens):
        self.mem, self.rr = O.run_lines(self.prog, self.mem, self.rr, gens)

    def board_bits(self, mem=None):
        return LifeCube.board_bits(self, self.mem if mem is None else mem)


def test_front_end_on_the_oracle_machine():
    life = _CpuLife(8, 8, 4, 5)  # a 128 x 256 torus on the 512 machines of one group
    start = life.start_bits()
    assert start.shape == (128, 256) and 0.4 < start.mean() < 0.6
    life.run(3)
    assert np.array_equal(life.board_bits(), O.life_words(start, 3))


@pytest.mark.gpu
def test_life_dims12(engine):
    print("life 4x4 tiles on 6+6 dims", flush=True)
    life = LifeCube(engine, life_cube(4, 4, 6, 6)[0])
    start = life.start_bits()
    life.run(7)
    assert np.array_equal(life.board_bits(), O.life_words(start, 7))


@pytest.mark.gpu
def test_life_8192_dims16(engine):
    # the configuration: 1,000 generations of the 8192 x 8192 torus, in two calls of uneven
    # length (graph replays of both parities), against the oracle's bit-parallel C Life
    print("life 8192 x 8192 on 65,536 machines, 1,000 generations", flush=True)
    life = LifeCube(engine)  # 32 x 32 tiles, rbits = cbits = 8
    start = life.start_bits()
    assert start.shape == (8192, 8192)
    life.run(333)
    life.run(667)
    print("oracle: 1,000 generations of C Life", flush=True)
    assert np.array_equal(life.board_bits(), O.life_words(start, 1000))
