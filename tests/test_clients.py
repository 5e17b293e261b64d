"""The large cube clients beta and gamma (reference src/beta.c, src/gamma.c) as program files: their
programs on the oracle's reference machine do what the reference's plain-C definitions do (the
beta emulator's beta_step, the world's gamma_step), with the states moved in and out by the
reference's own packers (beta_to_lane / beta_from_lane, gamma_to_cube / gamma_from_cube).
The device runs the same programs in tests/test_gpu_clients.py."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O
from tesseract_amd.programs import golden


def random_betas(ref, n, seed):
    """n powered-on betas (the monitor in ROM) with random RAM, registers, carry and an even
    PC, no input waiting: each runs whatever its random RAM holds."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        b = ref.Beta()
        ref.lib().beta_reset(C.byref(b))
        ram = rng.integers(0, 256, 0xA0).astype(np.uint8)
        C.memmove(b.mem, ram.ctypes.data, 0xA0)
        b.a, b.x = (int(v) for v in rng.integers(0, 256, 2))
        b.c = int(rng.integers(0, 2))
        b.pc = int(rng.integers(0, 128)) * 2
        out.append(b)
    return out


def beta_fields(b):
    return (bytes(b.mem), b.a, b.x, b.pc, b.c, b.out, b.out_strobe, b.halted, b.waiting)


def test_beta_program_is_the_emulator(ref):
    # 512 random betas, 24 instructions each: the program on the oracle machine, read back
    # through beta_from_lane, equals beta_step run 24 times on each
    prog = golden("beta")
    cells = prog.extra["cells"]
    betas = random_betas(ref, 512, 1)
    mem, rr = ref.beta_lines(prog.ptr, prog.ncells, cells.ctypes.data, betas)
    out, out_rr = O.run_lines(prog, mem, rr, 24)
    got = ref.beta_states(prog.ptr, prog.ncells, cells.ctypes.data, out, out_rr, len(betas))
    for b in betas:
        for _ in range(24):
            ref.lib().beta_step(C.byref(b))
    assert [beta_fields(g) for g in got] == [beta_fields(b) for b in betas]
    assert len({b.pc for b in betas}) > 50  # the betas went their own ways


@pytest.mark.parametrize("name", ["gamma_3_3_3"])
def test_gamma_program_is_the_world(ref, name):
    # a soup of cells with predators on the 8 x 8 x 8 torus: 6 ticks of the program on the oracle
    # machine, read back through gamma_from_cube, equal 6 gamma_step ticks of the world
    prog = golden(name)
    cells = prog.extra["cells"]
    geo = [int(v) for v in prog.extra["geometry"]]
    L = ref.lib()
    world, want, got = (ref.GammaWorld(*geo) for _ in range(3))
    assert L.gamma_soup(world.buf, 0.3, 60, 7) > 0
    L.gamma_predators(world.buf, 20, 8)
    mem, rr = ref.gamma_lines(prog.ptr, prog.ncells, cells.ctypes.data, world)
    out, out_rr = O.run_lines(prog, mem, rr, 6)
    ref.gamma_world_of(prog.ptr, prog.ncells, cells.ctypes.data, out, out_rr, got)
    ref.gamma_world_of(prog.ptr, prog.ncells, cells.ctypes.data, mem, rr, want)
    for _ in range(6):
        L.gamma_step(want.buf)
    assert ref.gamma_compare(got, want) == (-1, "")
    # and the world changed on the way
    start = ref.GammaWorld(*geo)
    ref.gamma_world_of(prog.ptr, prog.ncells, cells.ctypes.data, mem, rr, start)
    assert ref.gamma_compare(start, want)[0] != -1
