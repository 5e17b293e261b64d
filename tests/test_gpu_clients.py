"""The large cube clients on the device: beta (36,346 instructions, 512 byte computers per group)
and gamma (a 3-D world of cells, one voxel per machine, links across dims 0-11), compiled and run
by the engine from the reference's own program files, the states moved in and out by the
reference's own packers, against the reference's plain-C definitions (beta_step, gamma_step)."""
import ctypes as C

import numpy as np
import pytest

from tesseract_amd.programs import golden
from tests.test_clients import beta_fields, random_betas

pytestmark = pytest.mark.gpu


def test_beta_on_the_device(engine, ref):
    # 4,096 random betas (8 groups), 40 instructions in two calls (17 + 23 passes)
    prog = golden("beta")
    cells = prog.extra["cells"]
    betas = random_betas(ref, 4096, 2)
    mem, rr = ref.beta_lines(prog.ptr, prog.ncells, cells.ctypes.data, betas)
    print("compiling beta (36,346 instructions)", flush=True)
    code = engine.compile(prog)
    st = code.state(8)
    st.put(mem, rr)
    st.run(17)
    st.run(23)
    out, out_rr = st.get()
    got = ref.beta_states(prog.ptr, prog.ncells, cells.ctypes.data, out, out_rr, len(betas))
    for b in betas:
        for _ in range(40):
            ref.lib().beta_step(C.byref(b))
    assert [beta_fields(g) for g in got] == [beta_fields(b) for b in betas]


@pytest.mark.parametrize("name,ticks,tune", [("gamma_3_3_3", 9, None), ("gamma_4_4_4", 12, None),
                                             ("gamma_4_4_4", 12, {"lds_links": 1})])
def test_gamma_on_the_device(engine, ref, name, ticks, tune):
    # a soup with predators on the torus; ticks in two calls against as many gamma_step ticks.
    # gamma_4_4_4 is a cube of 128 words: segmented by default, and once in one workgroup per
    # cube (its links across the two waves through LDS, ~1,850 carried cells spilled)
    prog = golden(name)
    cells = prog.extra["cells"]
    geo = [int(v) for v in prog.extra["geometry"]]
    L = ref.lib()
    world = ref.GammaWorld(*geo)
    assert L.gamma_soup(world.buf, 0.3, 60, 11) > 0
    L.gamma_predators(world.buf, 1 << (sum(geo) - 4), 12)
    mem, rr = ref.gamma_lines(prog.ptr, prog.ncells, cells.ctypes.data, world)
    print("compiling %s (%d instructions, %d dims), tune %s" % (name, prog.len, prog.dims, tune),
          flush=True)
    code = engine.compile(prog, tune)
    if tune:
        assert code.stats()["resident"] == 1
    st = code.state()
    st.put(mem, rr)
    st.run(ticks // 3)
    st.run(ticks - ticks // 3)
    out, out_rr = st.get()
    got, want = ref.GammaWorld(*geo), ref.GammaWorld(*geo)
    ref.gamma_world_of(prog.ptr, prog.ncells, cells.ctypes.data, out, out_rr, got)
    ref.gamma_world_of(prog.ptr, prog.ncells, cells.ctypes.data, mem, rr, want)
    for _ in range(ticks):
        L.gamma_step(want.buf)
    assert ref.gamma_compare(got, want) == (-1, "")
