"""Batched cycle search (tester 1 on the GPU): per-program cycle-length statistics equal the
oracle's Brent schedule run machine by machine on the reference machine, whether the programs run
batched in persistent launches or one launch per program, in every residency class."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def bucket(lam):
    b = 0
    while b < 62 and (1 << b) < lam:
        b += 1
    return b


def expected(prog, seed, groups, max_passes):
    mem, rr = t.t1_start(prog, seed, groups)
    hist = [0] * 64
    total = longest = passes = 0
    for g in range(groups):
        lams, ran = O.brent_lines(prog, mem[g], rr[g], max_passes, with_passes=True)
        passes += ran
        for lam in lams:
            lam = int(lam)
            if lam == 0:
                hist[63] += 1
            else:
                hist[bucket(lam)] += 1
                total += lam
                longest = max(longest, lam)
    return {"hist": hist, "lam_sum": total, "longest": longest, "passes": passes, "unfinished": 0}


def programs():
    rng = np.random.default_rng(2024)
    out = []
    for k, cells in enumerate([1, 2, 3, 5, 8, 16, 30, 64, 100, 256, 600]):
        length = int(rng.integers(4, 400)) if cells < 16 else int(rng.integers(50, 1500))
        seed = t.t1_program_seed(99, cells, k)
        out.append((t.random_program(cells, length, seed), seed))
    return out


@pytest.mark.parametrize("plain", [False, True])
def test_batch_matches_oracle(engine, plain):
    progs = programs()
    print("batch of %d programs, plain=%s" % (len(progs), plain), flush=True)
    got, info = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=4,
                                 max_passes=64, plain=plain)
    for (p, s), g in zip(progs, got):
        want = expected(p, s, 4, 64)
        assert g == want, (p.ncells, p.len)
    # executed work as the reference counts it: passes x len x 512 (src/tester.c:305)
    assert info["ops"] == sum(p.len * g["passes"] * 512 for (p, _), g in zip(progs, got))


@pytest.mark.parametrize("groups", [1, 2, 3])
@pytest.mark.parametrize("lds_cells", [0, 20])
def test_batch_narrow_items(engine, groups, lds_cells):
    # one or two groups run as items of 16 or 32 words (the wave's other lanes idle), three as a
    # 64-word item with a throwaway group; lds_cells 20 sends every program of more than 18 cells
    # to HBM, 0 lets few items keep even the largest in LDS
    progs = programs()
    print("batch groups=%d lds_cells=%d" % (groups, lds_cells), flush=True)
    for plain in (False, True):
        got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=groups,
                                  max_passes=96, plain=plain, lds_cells=lds_cells)
        for (p, s), g in zip(progs, got):
            assert g == expected(p, s, groups, 96), (p.ncells, p.len, plain)


@pytest.mark.parametrize("groups", [1, 4])
def test_batch_machine_cycle_lengths(engine, groups):
    # the optional per-machine output: machine 512 g + k's cycle length is the oracle's Brent
    # result for lane k of group g (the reference's t1_cycles lam_out), 0 where none was found;
    # asking for it changes no statistic
    progs = programs()
    print("batch lams groups=%d" % groups, flush=True)
    for plain in (False, True):
        got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=groups,
                                  max_passes=64, plain=plain, lams=True)
        for (p, s), g in zip(progs, got):
            mem, rr = t.t1_start(p, s, groups)
            want = np.concatenate([O.brent_lines(p, mem[k], rr[k], 64) for k in range(groups)])
            assert np.array_equal(g.pop("lam"), want), (p.ncells, p.len, plain)
            assert g == expected(p, s, groups, 64), (p.ncells, p.len, plain)


@pytest.mark.parametrize("groups", [1, 3])
def test_batch_given_start_states(engine, groups):
    # a caller's start state (the reference's lines) instead of t1_start: the oracle's Brent from
    # the same state, per machine and per program
    progs = programs()
    starts = [O.random_lines(p, groups, 500 + k) for k, (p, _) in enumerate(progs)]
    print("batch given states groups=%d" % groups, flush=True)
    for plain in (False, True):
        got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=groups,
                                  max_passes=64, plain=plain, lams=True, starts=starts)
        for (p, _), (mem, rr), g in zip(progs, starts, got):
            lams, hist, total, longest, passes = [], [0] * 64, 0, 0, 0
            for k in range(groups):
                lam, ran = O.brent_lines(p, mem[k], rr[k], 64, with_passes=True)
                lams.append(lam)
                passes += ran
                for x in (int(v) for v in lam):
                    hist[63 if x == 0 else bucket(x)] += 1
                    total += x
                    longest = max(longest, x)
            assert np.array_equal(g.pop("lam"), np.concatenate(lams)), (p.ncells, plain)
            assert g == {"hist": hist, "lam_sum": total, "longest": longest, "passes": passes,
                         "unfinished": 0}


@pytest.mark.parametrize("lds_large_cells", [0, 1])
def test_batch_large_lds_class(engine, lds_large_cells):
    # programs too large for lds_cells: the 300-cell one by default in the large LDS class (two
    # items per CU), with lds_large_cells 1 in HBM like the larger ones; the same statistics either
    # way, batched and one launch each
    rng = np.random.default_rng(77)
    progs = []
    for k, cells in enumerate([300, 700, 1100]):
        seed = t.t1_program_seed(5, cells, k)
        progs.append((t.random_program(cells, int(rng.integers(400, 1500)), seed), seed))
    print("batch lds_large_cells=%d" % lds_large_cells, flush=True)
    for plain in (False, True):
        got, info = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=4,
                                     max_passes=48, plain=plain, lds_cells=20,
                                     lds_large_cells=lds_large_cells)
        for (p, s), g in zip(progs, got):
            assert g == expected(p, s, 4, 48), (p.ncells, p.len, plain)
        if not plain:
            assert (info["class_ops"][2] > 0) == (lds_large_cells == 0) and info["class_ops"][3] > 0


@pytest.mark.parametrize("waves_per_cu", [1, 16])
def test_batch_tortoise_placement(engine, waves_per_cu):
    # 128 groups = 32 items per program, 352 in all: at one wave per CU they outnumber the wave
    # slots, so the LDS classes keep the tortoise in HBM; at 16 both fit in LDS
    progs = programs()
    print("batch groups=128 waves_per_cu=%d" % waves_per_cu, flush=True)
    got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=128,
                              max_passes=24, waves_per_cu=waves_per_cu)
    for (p, s), g in zip(progs, got):
        assert g == expected(p, s, 128, 24), (p.ncells, p.len)


def test_batch_long_cycles(engine):
    # more passes than the default: buckets beyond 64 and "none" both occur
    progs = [(t.random_program(4, 30, 5), 5), (t.random_program(12, 40, 6), 6)]
    got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=4,
                              max_passes=300)
    for (p, s), g in zip(progs, got):
        assert g == expected(p, s, 4, 300)


def counter(bits):
    """A bits-bit binary counter: every pass adds one, so every machine cycles with length
    2^bits once the carry cells have been rewritten (a known answer for the cycle search)."""
    b = t.Builder()
    x = [b.cell("x%d" % i) for i in range(bits)]
    c, tmp = b.cell("c"), b.cell("t")
    b.gate(t.G_BUF, c, [t.CELL_ONE])
    for xi in x:
        b.gate(t.G_AND, tmp, [xi, c])
        b.gate(t.G_XOR, xi, [xi, c])
        b.gate(t.G_BUF, c, [tmp])
    return b.finish()


def test_batch_2_24_passes(engine):
    # the reference tester's largest search, 2^24 passes: a 2^21 counter's cycle is found after
    # 2^22 - 1 passes, a 2^23 counter's not within 2^24 (the whole search runs)
    progs = [(counter(21), 41), (counter(23), 42)]
    print("2^24 passes", flush=True)
    got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=1,
                              max_passes=1 << 24)
    for g, bits in zip(got, (21, 23)):
        found = (1 << (bits + 1)) - 1 <= 1 << 24
        hist = [0] * 64
        hist[bits if found else 63] = 512
        assert g == {"hist": hist, "lam_sum": 512 << bits if found else 0,
                     "longest": 1 << bits if found else 0,
                     "passes": (1 << (bits + 1)) - 1 if found else 1 << 24, "unfinished": 0}


def test_batch_deadline(engine):
    # a deadline of 0.5 s: the small programs finish, the long counter search is cut at a 64th
    # pass and reports unfinished with the passes it ran; the same small programs alone agree
    small = [(t.random_program(8, 32, 5), 5), (t.random_program(12, 48, 6), 6)]
    progs = small + [(counter(24), 7)]
    print("batch with a 0.5 s deadline", flush=True)
    got, info = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=1,
                                 max_passes=1 << 30, seconds=0.5)
    assert info["ms"] < 5000
    long = got[2]
    assert long["unfinished"] == 1 and long["passes"] > 0 and long["passes"] % 64 == 0
    alone, _ = engine.batch_run([p for p, _ in small], [s for _, s in small], groups=1,
                                max_passes=1 << 30)
    assert got[:2] == alone


def test_batch_reference_scale_passes(engine):
    # the reference's tester searches cycles over up to 2^24 passes: 2^20 here, with counters
    # whose cycle (2^12, 2^18) is found and one (2^21) that is not. The 2^12 counter is checked
    # against the oracle's Brent; the others against the closed form (every machine: lambda =
    # 2^bits, found once Brent's window reaches it, within 2^(bits+1) + 1 passes).
    progs = [(counter(12), 31), (counter(18), 32), (counter(21), 33)]
    print("2^20 passes", flush=True)
    got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=1,
                              max_passes=1 << 20)
    assert got[0] == expected(progs[0][0], progs[0][1], 1, 1 << 20)
    for g, bits in zip(got, (12, 18, 21)):
        # tail 1 (the carry cells), so Brent's tortoise waits at pass 2^k - 1 >= 1 with window
        # 2^k >= 2^bits: found at pass 2^bits - 1 + 2^bits
        found = (1 << (bits + 1)) - 1 <= 1 << 20
        hist = [0] * 64
        hist[bits if found else 63] = 512
        assert g == {"hist": hist, "lam_sum": 512 << bits if found else 0,
                     "longest": 1 << bits if found else 0,
                     "passes": (1 << (bits + 1)) - 1 if found else 1 << 20, "unfinished": 0}


def test_batch_program_without_instructions(engine):
    # nothing changes, so every machine is back at its start after one pass: cycle length 1
    b = t.Builder()
    b.cell("x")
    b.cell("y")
    prog = b.finish()
    got, info = engine.batch_run([prog], [7], groups=4, max_passes=64)
    assert got[0] == expected(prog, 7, 4, 64)
    assert got[0]["hist"][0] == 4 * 512 and got[0]["passes"] == 4


def stores_program(kind):
    b = t.Builder()
    x, y, z = b.cell("x"), b.cell("y"), b.cell("z")
    b.cell("never")  # neither read nor stored
    if kind == "dead":  # the first STO of x is overwritten before any read
        ops = [("n", x), ("n", y), ("s", x), ("n", z), ("s", x), ("s", y), ("n", x), ("s", z)]
    elif kind == "no_sto":  # RR alone changes
        ops = [("n", x), ("n", y), ("n", 1)]
    else:  # y is stored and never read in the pass: its end-of-pass value is all that counts
        ops = [("s", x), ("s", y), ("n", x), ("n", z), ("s", z), ("s", y), ("s", y), ("n", 0),
               ("n", z)]
    for op, c in ops:
        (b.nand if op == "n" else b.sto)(c)
    return b.finish()


@pytest.mark.parametrize("kind", ["dead", "no_sto", "unread"])
def test_batch_store_rules(engine, kind):
    # dead STOs are dropped and Brent compares, copies and starts only the cells the program
    # stores or reads: the cycle lengths stay the oracle's, machine by machine
    prog = stores_program(kind)
    print("batch store rules: %s" % kind, flush=True)
    for plain in (False, True):
        got, _ = engine.batch_run([prog], [11], groups=4, max_passes=64, plain=plain, lams=True)
        mem, rr = t.t1_start(prog, 11, 4)
        want = np.concatenate([O.brent_lines(prog, mem[k], rr[k], 64) for k in range(4)])
        assert np.array_equal(got[0].pop("lam"), want), plain
        assert got[0] == expected(prog, 11, 4, 64), plain


@pytest.mark.parametrize("residency", [1, 2])
def test_batch_sink_branch(engine, residency):
    # slots without a NAND or a STO branch around the sink's load and store in every class: the
    # same cycle lengths as the oracle's
    progs = programs()
    print("batch sink branch, residency %d" % residency, flush=True)
    for plain in (False, True):
        got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=4,
                                  max_passes=64, plain=plain, lams=True,
                                  residency=[residency] * len(progs), sink_branch=15)
        for (p, s), g in zip(progs, got):
            mem, rr = t.t1_start(p, s, 4)
            want = np.concatenate([O.brent_lines(p, mem[k], rr[k], 64) for k in range(4)])
            assert np.array_equal(g.pop("lam"), want), (p.ncells, plain)
            assert g == expected(p, s, 4, 64), (p.ncells, plain)


@pytest.mark.parametrize("residency", [0, 1, 2])
def test_batch_item_widths(engine, residency):
    # a caller's words per work item (16, 32, 64 for programs of 4 groups), mixed in one batch
    # and in every residency: per-machine cycle lengths and statistics equal the oracle's
    progs = programs()
    words = [(16, 32, 64)[k % 3] for k in range(len(progs))]
    print("batch item widths, residency %d" % residency, flush=True)
    for plain in (False, True):
        got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=4,
                                  max_passes=64, plain=plain, lams=True,
                                  residency=[residency] * len(progs), words=words)
        for (p, s), g, w in zip(progs, got, words):
            mem, rr = t.t1_start(p, s, 4)
            want = np.concatenate([O.brent_lines(p, mem[k], rr[k], 64) for k in range(4)])
            assert np.array_equal(g.pop("lam"), want), (p.ncells, w, plain)
            assert g == expected(p, s, 4, 64), (p.ncells, w, plain)


def test_batch_matches_reference_t1_cycles(engine, ref):
    # group 0 of each program is exactly the reference's t1_cycles run (its own jit-compiled
    # code): same cycle lengths, same passes
    from tesseract_amd.workloads import batch_mix
    progs, seeds = batch_mix(40, 5)
    keep = [k for k, p in enumerate(progs) if p.len <= 3000]
    progs, seeds = [progs[k] for k in keep], [seeds[k] for k in keep]
    got, _ = engine.batch_run(progs, seeds, groups=1, max_passes=64)
    for p, s, g in zip(progs, seeds, got):
        theirs = ref.make("random_program", p.ncells - 2, p.len, s)
        lam, ran = ref.t1_cycles(theirs, s, 64)
        hist = [0] * 64
        for x in lam:
            x = int(x)
            hist[63 if x == 0 else bucket(x)] += 1
        found = [int(x) for x in lam if x]
        assert g == {"hist": hist, "lam_sum": sum(found), "longest": max(found, default=0),
                     "passes": ran, "unfinished": 0}, (p.ncells, p.len)


def test_batch_mix_10000_matches_reference(engine, ref):
    # the batching configuration's 10,000 programs (cells 16..4096, 10^2..10^5 instructions), one
    # 512-machine group each, 64 passes: every machine's cycle length and every program's passes
    # equal the reference's own t1_cycles (its jit-compiled code, src/tester.c:96-164) on the same
    # program and seed, the programs run by 16 host threads
    import concurrent.futures as F
    from tesseract_amd.workloads import batch_mix
    progs, seeds = batch_mix(10000, 1)
    print("batch: 10,000 programs, %d instructions" % sum(p.len for p in progs), flush=True)
    got, info = engine.batch_run(progs, seeds, groups=1, max_passes=64, lams=True)
    print("device %.1f ms; the reference's t1_cycles on 16 threads" % info["ms"], flush=True)

    def theirs(k):
        p = progs[k]
        rp = ref.make("random_program", p.ncells - 2, p.len, seeds[k])
        assert rp.s.len == p.len and rp.s.ncells == p.ncells
        return ref.t1_cycles(rp, seeds[k], 64)

    bad = []
    with F.ThreadPoolExecutor(16) as pool:
        for k, (lam, ran) in enumerate(pool.map(theirs, range(len(progs)))):
            if not (np.array_equal(got[k]["lam"], lam) and got[k]["passes"] == ran):
                bad.append(k)
            if k % 2000 == 1999:
                print("%d programs compared" % (k + 1), flush=True)
    assert not bad, "programs differing: %s" % bad[:10]
    assert info["ops"] == sum(p.len * g["passes"] * 512 for p, g in zip(progs, got))


def test_batch_autotuned_mix_same_statistics(engine):
    # the autotuner's placements (HBM or an LDS class, item widths, windows) change where and how
    # programs run, never what they compute: the tuned batch, the tuned plain run and the untuned
    # batch give the same statistics per program
    from tesseract_amd.workloads import autotune_batch, batch_mix
    progs, seeds = batch_mix(60, 3)
    print("autotuned mix of 60 programs, 8 groups", flush=True)
    residency, win, words, report = autotune_batch(engine, progs, seeds, 8, sample=4,
                                                   sample_passes=2, sweeps=1)
    assert all(r in (2, 3, 4, 5) for r in residency) and set(words) <= {16, 32, 64}
    base, _ = engine.batch_run(progs, seeds, groups=8, max_passes=32)
    for plain in (False, True):
        got, _ = engine.batch_run(progs, seeds, groups=8, max_passes=32, plain=plain,
                                  residency=residency, win=win, words=words)
        assert got == base, plain
