"""The dims-9 cube config: a Benes permutation of 512 seeded random bytes and a bitonic sort, bit-
exact against benes_apply and a plain sort (the reference's cube demos, src/cubedemo.c:352-377),
and the programs' state after one pass against the oracle machine."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tesseract_amd.cube import Perm8, Sort8
from tests.helpers import golden


def test_perm8_on_the_oracle_machine():
    # the generated PERM8 program itself, run by the oracle, moves byte k to dest[k]
    rng = np.random.default_rng(7)
    prog = golden("perm8")
    vals = rng.integers(0, 256, 512).astype(np.uint8)
    dest = rng.permutation(512)
    ctrl = t.benes_route(dest.astype(np.uint16))
    mem, rr = O.reset_lines(prog, 1)
    pack = lambda bits: np.packbits(bits.astype(np.uint8), bitorder="little").view(np.uint64)
    for i in range(8):
        mem[0, prog.cell("v%d" % i)] = pack((vals >> i) & 1)
    for s in range(17):
        mem[0, prog.cell("c%d" % s)] = pack(ctrl[s])
    mem, rr = O.run_lines(prog, mem, rr, 1)
    out = np.zeros(512, np.uint8)
    for i in range(8):
        out |= (np.unpackbits(mem[0, prog.cell("v%d" % i)].view(np.uint8), bitorder="little") << i).astype(np.uint8)
    assert (out[dest] == vals).all()
    assert (out == O.benes_apply(ctrl, vals.astype(np.uint64))).all()


@pytest.mark.gpu
def test_perm8_gpu(engine):
    rng = np.random.default_rng(42)
    cubes = 64
    vals = rng.integers(0, 256, (cubes, 512)).astype(np.uint8)
    dests = np.stack([rng.permutation(512) for _ in range(cubes)])
    out, _ = Perm8(engine).run(vals, dests)
    for i in range(cubes):
        assert (out[i, dests[i]] == vals[i]).all()
        ctrl = t.benes_route(dests[i].astype(np.uint16))
        assert (out[i] == O.benes_apply(ctrl, vals[i].astype(np.uint64))).all()


@pytest.mark.gpu
def test_sort8_gpu(engine):
    rng = np.random.default_rng(43)
    vals = rng.integers(0, 256, (64, 512)).astype(np.uint8)
    vals[0] = 7            # all equal
    vals[1] = np.arange(512) % 256  # two of each
    out, _ = Sort8(engine).run(vals)
    assert (out == np.sort(vals, axis=1)).all()
