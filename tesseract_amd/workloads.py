"""Secondary workloads of the engine, each printing one JSON line:

  python -m tesseract_amd.workloads batch  [--programs N] [--groups G] [--passes P] [--seed S]
  python -m tesseract_amd.workloads life   [--gens G]
  python -m tesseract_amd.workloads hash   (toyhash32 preimage search over 2^32)
  python -m tesseract_amd.workloads benes  (Benes permutation and bitonic sort of 512 bytes)

batch: the program mix of the batching configuration. Program k draws two splitmix64 values from
the stream seeded with S: cells = (16, 64, 256, 1024, 4096)[d0 % 5] and
len = floor(10^(2 + 3 u)) with u = (d1 >> 11) / 2^53 (log-uniform in 10^2..10^5); its instructions
come from random_program(cells, len, t1_program_seed(S, cells, k)) and its machines start from the
t1_start stream of that seed. Every machine runs `passes` passes; the batch (persistent launches fed
by a device queue) and the plain form (one launch per program) must give identical statistics.
"""
import argparse
import ctypes as C
import json
import math
import time

import numpy as np

import tesseract_amd as t

SIZES = (16, 64, 256, 1024, 4096)


def batch_mix(n, seed):
    s = C.c_uint64(seed)
    progs, seeds = [], []
    for k in range(n):
        d0 = t.lib().ts_splitmix64(C.byref(s))
        d1 = t.lib().ts_splitmix64(C.byref(s))
        cells = SIZES[d0 % 5]
        u = (d1 >> 11) / float(1 << 53)
        length = int(math.floor(10 ** (2 + 3 * u)))
        ps = t.t1_program_seed(seed, cells, k)
        progs.append(t.random_program(cells, length, ps))
        seeds.append(ps)
    return progs, seeds


def run_batch(engine, n, groups, passes, seed, plain_too=True, lds_cells=0, waves_per_cu=0):
    progs, seeds = batch_mix(n, seed)
    out = {"workload": "batch", "programs": n, "machines_per_program": groups * 512,
           "passes": passes, "instructions": sum(p.len for p in progs)}
    t0 = time.perf_counter()
    stats, info = engine.batch_run(progs, seeds, groups=groups, max_passes=passes,
                                   lds_cells=lds_cells, waves_per_cu=waves_per_cu)
    out["batched_wall_s"] = time.perf_counter() - t0
    out["batched_ms"] = info["ms"]
    out["ops"] = info["ops"]
    out["batched_ops_per_s"] = info["ops"] / (info["ms"] * 1e-3)
    out["launches_batched"] = info["launches"]
    out["class_ms"] = info["class_ms"]
    out["class_ops_per_s"] = [o / (m * 1e-3) if m else 0 for o, m in zip(info["class_ops"], info["class_ms"])]
    if plain_too:
        stats_p, info_p = engine.batch_run(progs, seeds, groups=groups, max_passes=passes, plain=True,
                                           lds_cells=lds_cells, waves_per_cu=waves_per_cu)
        out["plain_ms"] = info_p["ms"]
        out["plain_ops_per_s"] = info_p["ops"] / (info_p["ms"] * 1e-3)
        out["launches_plain"] = info_p["launches"]
        out["batched_over_plain"] = info_p["ms"] / info["ms"]
        out["same_results_alone_and_batched"] = stats == stats_p
    hist = np.zeros(64, np.uint64)
    for st in stats:
        hist += np.array(st["hist"], np.uint64)
    out["cycle_hist_nonzero"] = {str(b): int(v) for b, v in enumerate(hist) if v}
    out["lam_sum"] = int(sum(st["lam_sum"] for st in stats))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["batch", "life", "hash", "benes"])
    ap.add_argument("--programs", type=int, default=10000)
    ap.add_argument("--groups", type=int, default=2048)
    ap.add_argument("--passes", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--gens", type=int, default=1000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--lds-cells", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    a = ap.parse_args()
    eng = t.Engine(0)
    if a.workload == "batch":
        out = run_batch(eng, a.programs, a.groups, a.passes, a.seed, not a.no_plain, a.lds_cells,
                        a.waves_per_cu)
    else:
        raise SystemExit("not built yet: " + a.workload)
    out["device"] = eng.describe()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
