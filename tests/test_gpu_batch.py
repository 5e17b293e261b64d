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
    total = longest = 0
    for g in range(groups):
        for lam in O.brent_lines(prog, mem[g], rr[g], max_passes):
            lam = int(lam)
            if lam == 0:
                hist[63] += 1
            else:
                hist[bucket(lam)] += 1
                total += lam
                longest = max(longest, lam)
    return {"hist": hist, "lam_sum": total, "longest": longest}


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
    assert info["ops"] == sum(p.len for p, _ in progs) * 64 * 2048


def test_batch_long_cycles(engine):
    # more passes than the default: buckets beyond 64 and "none" both occur
    progs = [(t.random_program(4, 30, 5), 5), (t.random_program(12, 40, 6), 6)]
    got, _ = engine.batch_run([p for p, _ in progs], [s for _, s in progs], groups=4,
                              max_passes=300)
    for (p, s), g in zip(progs, got):
        assert g == expected(p, s, 4, 300)
