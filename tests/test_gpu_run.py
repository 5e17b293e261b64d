"""Parity of gpu_run with the reference machine: the reference's check protocol (random state for
every writable cell and RR of every machine, several passes, every machine compared), on link-free
programs and on cube programs whose links cross in-word, in-wave and cross-wave dimensions."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tests.helpers import golden, life_cube

pytestmark = pytest.mark.gpu


def check(engine, prog, groups=None, ticks=8, seed=1, tune=None, runs=(1,)):
    code = engine.compile(prog, tune)
    groups = groups or prog.groups
    mem, rr = O.random_lines(prog, groups, seed)
    st = code.state(groups)
    st.put(mem, rr)
    done = 0
    for n in runs:  # several gpu_run calls continue the same machines
        st.run(n)
        done += n
    want_mem, want_rr = O.run_lines(prog, mem, rr, done)
    got_mem, got_rr = st.get()
    assert np.array_equal(got_rr, want_rr), "RR differs"
    bad = np.argwhere(got_mem != want_mem)
    assert bad.size == 0, "cells differ, first (group, cell, word) %s of %d" % (bad[0], len(bad))
    return code


def random_cube_program(cells, length, dims, seed, links=24):
    """A random program with id and link cells reaching every dimension below dims."""
    rng = np.random.default_rng(seed)
    b = t.Builder()
    plain = [b.cell("c%d" % i) for i in range(cells)]
    ids = [b.cell("ID%d" % d) for d in range(dims)]
    lk = []
    for k in range(links):
        d = int(rng.integers(0, dims)) if k else dims - 1
        src = plain[int(rng.integers(0, cells))] if k % 5 else ids[int(rng.integers(0, dims))]
        lk.append(b.link("l%d" % k, src, d))
    for k in range(4):  # chains of links add up
        lk.append(b.link("m%d" % k, lk[int(rng.integers(0, len(lk)))], int(rng.integers(0, dims))))
    readable = [0, 1] + plain + ids + lk
    for _ in range(length):
        if rng.random() < 0.55:
            b.nand(readable[int(rng.integers(0, len(readable)))])
        else:
            b.sto(plain[int(rng.integers(0, cells))])
    return b.finish()


@pytest.mark.parametrize("name", ["adder8", "mul4", "hash16", "rand20", "rand400"])
def test_linkfree(engine, name):
    prog = {"adder8": lambda: golden("adder8"), "mul4": lambda: golden("mul4"),
            "hash16": lambda: golden("hash16_5"), "rand20": lambda: t.random_program(20, 500, 42),
            "rand400": lambda: t.random_program(400, 3000, 7)}[name]()
    print("check", name, flush=True)
    check(engine, prog, groups=4, ticks=8, runs=(3, 5))


def test_linkfree_many_groups(engine):
    prog = t.random_program(50, 300, 9)
    print("check rand50 x 64 groups", flush=True)
    check(engine, prog, groups=64, runs=(4,))


@pytest.mark.parametrize("dims", [9, 10, 11, 12, 13, 14, 15, 16])
def test_random_cube(engine, dims):
    prog = random_cube_program(24, 400, dims, seed=100 + dims)
    print("check random cube dims", dims, flush=True)
    check(engine, prog, runs=(2, 3))


@pytest.mark.parametrize("lds_links", [1, 2])
@pytest.mark.parametrize("dims", [12, 13, 14, 15])
def test_cube_in_one_workgroup(engine, dims, lds_links):
    # cubes of 128-1,024 words: each cube in one workgroup, reads across its waves through LDS and
    # the pass loop in-kernel (lds_links 1), or segments through HBM (2); three cubes per state,
    # several calls, every machine against the oracle
    prog = random_cube_program(24, 600, dims, seed=300 + dims, links=40)
    print("check cube dims %d lds_links %d" % (dims, lds_links), flush=True)
    code = check(engine, prog, groups=3 * prog.groups, tune={"lds_links": lds_links},
                 runs=(1, 4, 2))
    info = code.stats()
    assert info["resident"] == (lds_links == 1)
    if lds_links == 1:
        assert info["segments"] == 1


def wide_cross_program(dims=14, cells=48):
    """Every cell read across the top dim at the start of the pass and again after it was
    written: more values cross the waves at one level than one chunk of LDS rows holds."""
    b = t.Builder()
    plain = [b.cell("c%d" % i) for i in range(cells)]
    far = [b.link("f%d" % i, plain[i], dims - 1) for i in range(cells)]
    near = [b.link("n%d" % i, plain[i], dims - 2) for i in range(cells)]
    for i in range(cells):  # start-of-pass values from the other half of the cube
        b.nand(far[i])
        b.nand(plain[(i + 1) % cells])
        b.sto(plain[i])
    for i in range(cells):  # values computed in the pass, read across the waves
        b.nand(near[(i * 7) % cells])
        b.nand(far[(i * 5 + 3) % cells])
        b.sto(plain[(i + 11) % cells])
    return b.finish()


def test_cube_in_one_workgroup_many_rows(engine):
    # 48 start-of-pass values cross the waves at one level: with 512-word cubes a chunk holds 32
    # LDS rows, so the level takes two chunks (write, barrier, reads, barrier each)
    prog = wide_cross_program()
    print("check wide cross-wave program, dims 14", flush=True)
    code = check(engine, prog, groups=2 * prog.groups, tune={"lds_links": 1}, runs=(1, 3))
    assert code.stats()["resident"] == 1


@pytest.mark.parametrize("dims", [9, 12])
def test_rank(engine, dims):
    print("check rank", dims, flush=True)
    check(engine, golden("rank%d" % dims), runs=(1, 1))


def test_perm8_sort8(engine):
    print("check perm8/sort8", flush=True)
    check(engine, golden("perm8"), runs=(2,))
    check(engine, golden("sort8"), groups=2, runs=(1,))


def test_life_cube_small(engine):
    prog, board = life_cube(4, 4, 3, 3)  # 6 dims of board, run on dims max(9, 6) = 9: one group
    print("check life cube 4x4 on 3+3 dims", flush=True)
    check(engine, prog, runs=(2,))
    prog, board = life_cube(4, 4, 6, 6)  # dims 12: links across the cube's two waves
    print("check life cube 4x4 on 6+6 dims, one workgroup per cube", flush=True)
    assert check(engine, prog, runs=(1, 2)).stats()["resident"] == 1
    print("check life cube 4x4 on 6+6 dims, streamed (cross-wave values recomputed)", flush=True)
    assert check(engine, prog, tune={"lds_links": 2}, runs=(1, 2)).stats()["resident"] == 0


@pytest.mark.parametrize("gp", [4, 3, 5])
def test_graph_replays_match_eager(engine, gp):
    # segmented program, passes replayed from captured graphs of gp passes: several replays with
    # eager launches in between, each checked against the oracle. An odd gp flips the parity at
    # every replay, so both parities' graphs are captured and replayed within one call.
    prog = random_cube_program(20, 300, 12, seed=77)
    code = engine.compile(prog, {"graph_passes": gp, "lds_links": 2})
    assert code.stats()["segments"] > 1
    mem, rr = O.random_lines(prog, prog.groups, 5)
    st = code.state()
    st.put(mem, rr)
    done = 0
    for n in (9, 8, 3, 4, 10, 20):
        st.run(n)
        done += n
        gm, gr = st.get()
        wm, wr = O.run_lines(prog, mem, rr, done)
        assert np.array_equal(gm, wm) and np.array_equal(gr, wr), "after %d passes" % done
    code1 = engine.compile(prog, {"graph_passes": 1, "lds_links": 2})
    st1 = code1.state()
    st1.put(mem, rr)
    st1.run(done)
    assert all(np.array_equal(a, b) for a, b in zip(st1.get(), st.get()))


@pytest.mark.parametrize("dims,seed", [(9, 41), (12, 42), (14, 43), (16, 44)])
def test_run_steps_against_oracle(engine, dims, seed):
    # partial passes: single instructions, a pass finished from the middle, whole passes from
    # pc 0 and a run wrapping from the middle of one pass into the middle of another, each
    # compared with the oracle machine stepped the same way (every cell of every machine)
    prog = random_cube_program(12, 120, dims, seed)
    groups = prog.groups
    mem, rr = O.random_lines(prog, groups, seed)
    code = engine.compile(prog)
    st = code.state(groups)
    st.put(mem, rr)
    writable = [c for c in range(2, prog.ncells) if prog.kind(c) == t.CK_PLAIN]
    pc = 0
    print("steps, dims", dims, flush=True)
    for steps in (1, 4, 115, 240, 3 * 120, 50, 130, 59):
        st.run_steps(steps)
        mem, rr, pc = O.steps_lines(prog, mem, rr, pc, steps)
        gm, gr = st.get()
        assert st.pc == pc
        assert np.array_equal(gr, rr), steps
        assert np.array_equal(gm[:, writable], mem[:, writable]), steps
    with pytest.raises(RuntimeError, match="pass is in progress"):
        st.run(1)


def test_peek_poke(engine):
    # one machine's cells after a pass, links and ids included, as the oracle machine has them;
    # poke changes one machine's cell and leaves read-only cells alone
    prog = random_cube_program(10, 80, 10, 43)
    groups = prog.groups
    mem, rr = O.random_lines(prog, groups, 43)
    code = engine.compile(prog)
    st = code.state(groups)
    st.put(mem, rr)
    st.run(1)
    want_mem, want_rr = O.run_lines(prog, mem, rr, 1)
    bit = lambda line, k: int(line[k // 64]) >> (k % 64) & 1
    cube = 1 << prog.dims
    n = prog.ncells
    machines = [0, 1, 511, 512, 700, groups * 512 - 1]
    for m in machines:
        assert st.peek(m, n) == bit(want_rr[m // 512], m % 512)
        for cell in range(n):
            kind = prog.kind(cell)
            if kind == t.CK_PLAIN:
                want = bit(want_mem[m // 512, cell], m % 512)
            elif kind == t.CK_ID:
                want = (m % cube) >> prog.cell_dim(cell) & 1
            else:
                continue  # links: checked through their base cell below
            assert st.peek(m, cell) == want, (m, cell)
    for cell in range(n):
        before = [st.peek(m, cell) for m in machines]
        st.poke(700, cell, 1 - before[4])
        after = [st.peek(m, cell) for m in machines]
        want = list(before)
        if prog.kind(cell) == t.CK_PLAIN and cell >= 2:
            want[4] = 1 - before[4]
        assert after == want, cell


def test_empty_program(engine):
    # a program of no instructions still runs passes: every cell and RR keep their values
    # (reference src/ref.c:52-55)
    b = t.Builder()
    x = b.cell("x")
    b.input(x)
    b.output(x)
    prog = b.finish()
    mem, rr = O.random_lines(prog, 1, 9)
    st = engine.compile(prog).state(1)
    st.put(mem, rr)
    st.run(5)
    st.run_steps(3)
    gm, gr = st.get()
    assert np.array_equal(gr, rr) and np.array_equal(gm[:, 2:], mem[:, 2:])
    assert st.pc == 0


@pytest.mark.parametrize("dims", [9, 12])
def test_power_on_state(engine, dims):
    # a new state and a reset one hold the reference's power-on state (src/ref.c:24-32,
    # src/lanes.c:110-192): every cell 0, ONE all ones, RR 0, id cells their machine-number
    # patterns; then one pass from it equals the oracle's pass from its own power-on lines
    prog = random_cube_program(16, 200, dims, seed=90 + dims)
    code = engine.compile(prog)
    groups = 2 * prog.groups
    want_mem, want_rr = O.reset_lines(prog, groups)
    st = code.state(groups)
    got_mem, got_rr = st.get()
    assert np.array_equal(got_mem, want_mem) and np.array_equal(got_rr, want_rr)
    st.run(2)
    run_mem, run_rr = O.run_lines(prog, want_mem, want_rr, 2)
    got_mem, got_rr = st.get()
    assert np.array_equal(got_mem, run_mem) and np.array_equal(got_rr, run_rr)
    st.reset()
    got_mem, got_rr = st.get()
    assert np.array_equal(got_mem, want_mem) and np.array_equal(got_rr, want_rr)


@pytest.mark.parametrize("name", ["cube_any", "cube_rank", "cube_max", "rule110", "adder8",
                                  "counter4", "mux", "xor", "full_adder", "c17", "perm8", "sort8"])
def test_reference_examples(engine, name):
    # the reference's own example and cube-library programs (examples/*.asm, lib/cube.inc through
    # its assembler): the check protocol against the oracle, whole passes over several calls
    prog = golden(name)
    print("check", name, prog.len, "instructions, dims", prog.dims, flush=True)
    check(engine, prog, groups=2 * prog.groups, runs=(1, 2))
