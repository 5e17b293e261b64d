"""Secondary workloads of the engine, each printing one JSON line of device measurements:

  python -m tesseract_amd.workloads batch  [--programs N] [--groups G] [--passes P] [--seed S]
  python -m tesseract_amd.workloads life   [--gens G]
  python -m tesseract_amd.workloads hash   (toyhash32 preimage search over 2^32)
  python -m tesseract_amd.workloads benes  (Benes permutation and bitonic sort of 512 bytes)
  python -m tesseract_amd.workloads serve  [--requests N] [--rate R]
  python -m tesseract_amd.workloads polymul [--cubes C] (NTT product mod x^512 + 1, 12289)
  python -m tesseract_amd.workloads tape   (one equiv on the tape interpreter and compiled)

`python bench.py --workload life|batch|serve|polymul` runs the same with their CPU baselines and
the exactness checks against plain C beside them.

batch: the program mix of the batching configuration. Program k draws two splitmix64 values from
the stream seeded with S: cells = (16, 64, 256, 1024, 4096)[d0 % 5] and
len = floor(10^(2 + 3 u)) with u = (d1 >> 11) / 2^53 (log-uniform in 10^2..10^5); its instructions
come from random_program(cells, len, t1_program_seed(S, cells, k)) and its machines start from the
t1_start stream of that seed. A group of 512 machines runs until every machine found its cycle, at
most `passes` passes; the batch (persistent launches fed by a device queue) and the plain form (one
launch per program) must give identical statistics.
"""
import argparse
import ctypes as C
import json
import math
import sys
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


def t1_start_group(prog, seed, g):
    """Group g's start state in the t1_start stream of `seed` (group 0 is the reference's
    t1_start, src/tester.c:83-91; later groups continue the stream, cells then RR, group after
    group): (mem (ncells, 8), rr (8,)) u64 lines, computed directly at the stream's offset."""
    per = (prog.ncells - 2) * 8 + 8
    k = np.arange(g * per + 1, (g + 1) * per + 1, dtype=np.uint64)
    base = np.uint64((seed ^ 0x5DEECE66D) & (2 ** 64 - 1))
    with np.errstate(over="ignore"):
        z = base + np.uint64(0x9E3779B97F4A7C15) * k
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    mem = np.zeros((prog.ncells, 8), np.uint64)
    mem[1] = ~np.uint64(0)
    mem[2:] = z[:-8].reshape(prog.ncells - 2, 8)
    return mem, z[-8:].copy()


LDS_BYTES = 160 * 1024  # a CU's LDS on gfx950
WAVES_PER_CU = 16       # the batch's default persistent waves per CU per class


def lds_classes(needs):
    """Sizes placed in LDS -> one of the three LDS classes (0, 1, 2), given each size's item
    bytes. A class's launch sizes every workgroup's LDS for its largest item, so a size in a class
    runs as many waves per CU as that item allows: sizes are sorted by need and, while there are
    more than three, the two neighbours whose wave counts differ least share a class."""
    slots = lambda n: min(WAVES_PER_CU, LDS_BYTES // n)
    groups = [[c] for c in sorted(needs, key=lambda c: needs[c])]
    while len(groups) > 3:
        ratio = [slots(needs[groups[k][-1]]) / slots(needs[groups[k + 1][-1]])
                 for k in range(len(groups) - 1)]
        k = min(range(len(ratio)), key=lambda k: ratio[k])
        groups[k:k + 2] = [groups[k] + groups[k + 1]]
    return {c: k for k, g in enumerate(groups) for c in g}


def autotune_batch(engine, progs, seeds, groups, sample=24, sample_passes=4, sweeps=2):
    """The batch autotuner: per machine size of the mix, where its machines live (LDS or HBM),
    the words of its work items (lanes per wave: 16, 32 or 64; narrower items need less LDS, so
    more waves share a CU) and the slots per window (8, 16, or the host cost model's choice per
    program: the unroll depth of the interpreter's load rounds), picked by timing samples (the
    first `sample` programs of each size, `sample_passes` passes each) at their 1-bit ops/s.
    Windows are a local choice: each size's sample runs alone under each of them for every
    placement (HBM, or LDS at each width two items of which fit a CU). Placements are not: the
    classes run side by side, share HBM, and sizes in one LDS class share its LDS size
    (lds_classes), so placements are chosen on the mixed sample of all sizes together, starting
    from each size's best alone and trying every other placement of each size in turn (largest
    total instruction count first) for `sweeps` rounds, keeping what runs the whole sample faster.
    Returns per-program residency (2 = HBM, 3-5 = an LDS class), window and width lists for
    gpu_batch_run, and the measured tables."""
    default_words = 16 if groups == 1 else 32 if groups == 2 else 64
    widths = [w for w in (16, 32, 64) if w <= default_words]
    need = lambda c, words: (c + 1) * 4 * words
    by_cells = {}
    for i, p in enumerate(progs):
        by_cells.setdefault(p.ncells, []).append(i)
    sizes = sorted(by_cells)
    samples = {c: by_cells[c][:sample] for c in sizes}
    name = lambda place, w: "%s%s/%s" % ("HBM" if place is None else "LDS", "" if place is None
                                         else "/%dw" % place, w or "model")
    rate = lambda info: info["ops"] / (info["ms"] * 1e-3) if info["ms"] else 0.0
    # placements: None = HBM (default width), or an LDS width
    places = {c: [None] + [w for w in widths if need(c, w) <= LDS_BYTES // 2] for c in sizes}
    best_win, alone, table = {}, {}, []
    for c in sizes:
        idx = samples[c]
        sp, ss = [progs[i] for i in idx], [seeds[i] for i in idx]
        rates = {}
        for place in places[c]:
            for w in (0, 8, 16):
                _, info = engine.batch_run(sp, ss, groups=groups, max_passes=sample_passes,
                                           residency=[2 if place is None else 1] * len(sp),
                                           win=[w] * len(sp),
                                           words=[place or default_words] * len(sp))
                rates[(place, w)] = rate(info)
            best_win[(c, place)] = max((0, 8, 16), key=lambda w: rates[(place, w)])
            alone[(c, place)] = rates[(place, best_win[(c, place)])]
        table.append({"cells": c - 2, "sample_programs": len(idx),
                      "alone": {name(*k): v for k, v in rates.items()}})
    mixed = [i for c in sizes for i in samples[c]]

    def assignment(place_of):
        lds = {c: need(c, place_of[c]) for c in sizes if place_of[c] is not None}
        cls = lds_classes(lds)
        res = {c: 2 if place_of[c] is None else 3 + cls[c] for c in sizes}
        return res, cls

    def joint(place_of):
        res, _ = assignment(place_of)
        _, info = engine.batch_run([progs[i] for i in mixed], [seeds[i] for i in mixed],
                                   groups=groups, max_passes=sample_passes,
                                   residency=[res[progs[i].ncells] for i in mixed],
                                   win=[best_win[(progs[i].ncells, place_of[progs[i].ncells])]
                                        for i in mixed],
                                   words=[place_of[progs[i].ncells] or default_words
                                          for i in mixed])
        return info["ms"]

    place_of = {c: max(places[c], key=lambda pl: alone[(c, pl)]) for c in sizes}
    tried = []
    best_ms = joint(place_of)
    tried.append({"placement": {c - 2: name(place_of[c], best_win[(c, place_of[c])]) for c in sizes},
                  "mixed_sample_ms": best_ms})
    order = sorted(sizes, key=lambda c: -sum(progs[i].len for i in by_cells[c]))
    for _ in range(sweeps):
        changed = False
        for c in order:
            for pl in places[c]:
                if pl == place_of[c]:
                    continue
                trial = dict(place_of)
                trial[c] = pl
                ms = joint(trial)
                tried.append({"placement": {x - 2: name(trial[x], best_win[(x, trial[x])])
                                            for x in sizes}, "mixed_sample_ms": ms})
                if ms < best_ms:
                    best_ms, place_of, changed = ms, trial, True
        if not changed:
            break
    res, cls = assignment(place_of)
    residency, win, words = [0] * len(progs), [0] * len(progs), [0] * len(progs)
    for c in sizes:
        for i in by_cells[c]:
            residency[i] = res[c]
            win[i] = best_win[(c, place_of[c])]
            words[i] = place_of[c] or default_words
    for row in table:
        c = row["cells"] + 2
        row["choice"] = name(place_of[c], best_win[(c, place_of[c])])
        if place_of[c] is not None:
            row["lds_class"] = cls[c]
    return residency, win, words, {"sizes": table, "placements_tried": tried,
                                   "mixed_sample_ms": best_ms}


def run_batch(engine, n, groups, passes, seed, plain_too=True, lds_cells=0, waves_per_cu=0,
              lds_large_cells=0, lam_groups=None, keep=None, tuned=True):
    """The mix batched (and, with plain_too, one launch per program), with the autotuner's
    per-size choices when tuned (both forms run with them). lam_groups: per program the group
    whose 512 cycle lengths the batched run returns (for checks); keep, a dict, then receives
    the programs, seeds and per-program results."""
    progs, seeds = batch_mix(n, seed)
    if callable(lam_groups):
        lam_groups = lam_groups(progs)
    out = {"workload": "batch", "programs": n, "machines_per_program": groups * 512,
           "passes": passes, "instructions": sum(p.len for p in progs)}
    residency = win = words = None
    if tuned:
        t0 = time.perf_counter()
        residency, win, words, out["autotuner"] = autotune_batch(engine, progs, seeds, groups)
        out["autotune_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    stats, info = engine.batch_run(progs, seeds, groups=groups, max_passes=passes,
                                   lds_cells=lds_cells, waves_per_cu=waves_per_cu,
                                   lds_large_cells=lds_large_cells,
                                   lams=lam_groups is not None, lam_groups=lam_groups,
                                   residency=residency, win=win, words=words)
    if keep is not None:
        keep.update(progs=progs, seeds=seeds, stats=stats, lam_groups=lam_groups)
    out["batched_wall_s"] = time.perf_counter() - t0
    out["batched_ms"] = info["ms"]
    out["ops"] = info["ops"]
    out["batched_ops_per_s"] = info["ops"] / (info["ms"] * 1e-3)
    out["launches_batched"] = info["launches"]
    out["class_ms"] = info["class_ms"]
    out["class_ops_per_s"] = [o / (m * 1e-3) if m else 0 for o, m in zip(info["class_ops"], info["class_ms"])]
    if plain_too:
        stats_p, info_p = engine.batch_run(progs, seeds, groups=groups, max_passes=passes, plain=True,
                                           lds_cells=lds_cells, waves_per_cu=waves_per_cu,
                                           lds_large_cells=lds_large_cells,
                                           residency=residency, win=win, words=words)
        out["plain_ms"] = info_p["ms"]
        out["plain_ops_per_s"] = info_p["ops"] / (info_p["ms"] * 1e-3)
        out["launches_plain"] = info_p["launches"]
        out["batched_over_plain"] = info_p["ms"] / info["ms"]
        strip = lambda L: [{k: v for k, v in d.items() if k != "lam"} for d in L]
        out["same_results_alone_and_batched"] = strip(stats) == strip(stats_p)
    hist = np.zeros(64, np.uint64)
    for st in stats:
        hist += np.array(st["hist"], np.uint64)
    out["cycle_hist_nonzero"] = {str(b): int(v) for b, v in enumerate(hist) if v}
    out["lam_sum"] = int(sum(st["lam_sum"] for st in stats))
    return out


def run_life(engine, gens, tune=None, keep=None):
    """Life on the 8192 x 8192 torus: 64 warm-up generations, then `gens` timed. keep, a dict,
    receives the LifeCube (its start board and final state) for a caller's own checks."""
    from .life import LifeCube
    t0 = time.perf_counter()
    life = LifeCube(engine, tune=tune)  # 8192 x 8192 torus on 65,536 machines, 32 x 32 tiles
    setup = time.perf_counter() - t0
    info = life.code.stats()
    life.run(64)  # warm-up: builds the captured graphs of both parities
    ms = life.run(gens)
    out = {"workload": "life", "board": "8192x8192", "machines": 65536, "generations": gens,
           "instructions_per_generation": life.prog.len, "segments": info["segments"],
           "slices": info["slices"], "bitop3_per_generation_word": info["bitop3"],
           "compile_and_setup_s": setup, "ms": ms, "ms_per_generation": ms / gens,
           "cell_updates_per_s": 8192.0 * 8192 * gens / (ms * 1e-3),
           "ops_per_s": life.prog.len * 65536.0 * gens / (ms * 1e-3)}
    if keep is not None:
        keep["life"] = life
    return out


def run_benes(engine, cubes, seed):
    from .cube import Perm8, Sort8
    rng = np.random.default_rng(seed)
    vals = rng.integers(0, 256, (cubes, 512)).astype(np.uint8)
    dests = np.stack([rng.permutation(512) for _ in range(cubes)])
    perm, sort = Perm8(engine), Sort8(engine)
    perm.run(vals[:1], dests[:1])
    sort.run(vals[:1])
    got_p, ms_p = perm.run(vals, dests, repeats=20)
    got_s, ms_s = sort.run(vals, repeats=20)
    ok_p = all((got_p[i, dests[i]] == vals[i]).all() for i in range(cubes))
    ok_s = bool((got_s == np.sort(vals, axis=1)).all())
    return {"workload": "benes", "cubes": cubes, "perm8_instructions": perm.prog.len,
            "sort8_instructions": sort.prog.len, "perm_ms": ms_p, "sort_ms": ms_s,
            "permutations_per_s": cubes / (ms_p * 1e-3), "sorts_per_s": cubes / (ms_s * 1e-3),
            "perm_ops_per_s": perm.prog.len * 512.0 * cubes / (ms_p * 1e-3),
            "sort_ops_per_s": sort.prog.len * 512.0 * cubes / (ms_s * 1e-3),
            "perm_bit_exact": ok_p, "sort_bit_exact": ok_s}


def run_polymul(engine, cubes, seed, runs=20, keep=None):
    """The NTT polynomial product (gen_polymul, 220,758 instructions) on `cubes` independent
    512-machine cubes of seeded random polynomials: one warm-up call, then `runs` calls of one
    pass each, timed (each call computes every cube's product in full). keep, a dict, receives
    a, b and c for a caller's checks."""
    from .cube import PolyMul
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 12289, (cubes, 512)).astype(np.uint16)
    b = rng.integers(0, 12289, (cubes, 512)).astype(np.uint16)
    t0 = time.perf_counter()
    pm = PolyMul(engine)
    setup = time.perf_counter() - t0
    info = pm.code.stats()
    c1, _ = pm.run(a, b, 1)
    c, each = pm.run_each(a, b, runs)
    ms = sum(each)
    out = {"workload": "polymul", "cubes": cubes, "instructions": pm.prog.len,
           "bitop3_per_word": info["bitop3"], "resident": bool(info["resident"]),
           "compile_s": setup, "runs": runs, "ms": ms, "us_per_product_batch": ms * 1e3 / runs,
           "products_per_s": cubes * runs / (ms * 1e-3),
           "ops_per_s": pm.prog.len * 512.0 * cubes * runs / (ms * 1e-3),
           "same_product_every_call": bool((c == c1).all())}
    if keep is not None:
        keep.update(a=a, b=b, c=c, polymul=pm)
    return out


def run_tape(engine, inputs=26, gates=2000, seed=3, runs=5):
    """The tape tier on its own: an equiv of a seeded random netlist of `gates` gates and `inputs`
    inputs against its NAND-level rewrite (netlists.circuit / rewrite), enumerated on the
    precompiled tape interpreter (no hipRTC) and, for comparison, as a compiled kernel."""
    from . import netlists as N
    rng = np.random.default_rng(seed)
    c = N.circuit(inputs, 8, gates, rng)
    r = N.rewrite(c, rng)
    a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
    tape = t.enum_tape(t.EN_EQUIV, a, b)
    out = {"workload": "tape", "inputs": inputs, "gates": gates, "and_nodes": tape["and_nodes"],
           "groups": int(tape["ops"].shape[0]), "slots": tape["slots"]}
    for tier, tune in (("tape", {"tape": 1}), ("compiled", None)):
        code = t.EnumCode(engine, t.EN_EQUIV, a, b, tune=tune)
        code.run()
        ms = []
        for _ in range(runs):
            res = code.run()
            ms.append(engine.last_ms)
        out[tier + "_ms"] = min(ms)
        out[tier + "_and_per_s"] = tape["and_nodes"] * float(res["tried"]) / (min(ms) * 1e-3)
        out[tier + "_hits"] = res["hits"]
    return out


def run_hash(engine, target=0x165C2A07):
    from .programs import golden
    prog = golden("hash32_8")  # the reference's gen_hash(32, 8)
    code = t.EnumCode(engine, t.EN_SEARCH, prog, spec="h = %d" % target)
    code.run()
    r = code.run()
    return {"workload": "hash", "program": "gen_hash(32, 8)", "instructions": prog.len,
            "target": target, "assignments": r["tried"], "hits": r["hits"], "hit": r["hit"],
            "ms": engine.last_ms, "ops_per_s": r["ops"] / (engine.last_ms * 1e-3)}


def serve_requests(seed, n):
    """The stream's requests: dicts of mode, a, b (equiv) or spec (verify), n_in, faulty, and for
    equiv the circuits c, r themselves. Every VERIFY_EVERY-th request is a verify."""
    from . import netlists as N
    from .programs import golden
    reqs = []
    t_gen = time.perf_counter()
    for k in range(n):
        if k % N.VERIFY_EVERY == N.VERIFY_EVERY - 1:
            name, spec, n_in, planted = N.verify_request(seed, k)
            reqs.append({"mode": t.EN_VERIFY, "a": golden(name), "b": None, "spec": spec,
                         "program": name, "n_in": n_in, "faulty": planted})
        else:
            c, r, n_in, faulty = N.request(seed, k)
            reqs.append({"mode": t.EN_EQUIV, "a": t.net_parse(c.bench()),
                         "b": t.net_parse(r.bench()), "spec": None, "c": c, "r": r,
                         "n_in": n_in, "faulty": faulty})
        if (k + 1) % 50 == 0:
            print("serve: %d requests generated (%.1f s)" % (k + 1, time.perf_counter() - t_gen),
                  file=sys.stderr, flush=True)
    return reqs


def run_serve(engine, n, rate, seed, streams=4, chunk_words=1 << 24, check=True, keep=None,
              tape_race_ms=0.0):
    """A seeded Poisson stream of equiv and verify requests (serve_requests) submitted at their
    arrival times; latency = done - arrival. Offered load is the stream's; served load is what
    finished."""
    import threading
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = serve_requests(seed, n)
    srv = t.Server(engine, streams=streams, chunk_words=chunk_words, tape_race_ms=tape_race_ms)
    tickets = [None] * n
    t0 = time.perf_counter() + 0.05

    def feeder():
        for k in range(n):
            delay = t0 + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            q = reqs[k]
            tickets[k] = srv.submit(q["mode"], q["a"], q["b"], spec=q["spec"])

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    results = [srv.wait(tk) for tk in tickets]
    times = [srv.times(tk) for tk in tickets]
    end = max(tm["done_s"] for tm in times)
    first = min(tm["submitted_s"] for tm in times)
    srv.close()
    lat = np.array([tm["done_s"] - tm["submitted_s"] for tm in times])
    n_in = np.array([q["n_in"] for q in reqs])
    out = {"workload": "serve", "requests": n, "offered_rate_per_s": rate,
           "served_rate_per_s": n / (end - first), "wall_s": end - first,
           "ops": int(sum(r["ops"] for r in results)),
           "ops_per_s": sum(r["ops"] for r in results) / (end - first),
           "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "compile_p50_ms": float(np.median([tm["compiled_s"] - tm["submitted_s"] for tm in times]) * 1e3),
           "verify_requests": int(sum(q["mode"] == t.EN_VERIFY for q in reqs)),
           "faulty": int(sum(q["faulty"] for q in reqs)), "mismatching": int(sum(r["hits"] > 0 for r in results)),
           "tape_tier_requests": int(sum(tm["tier"] == 1 for tm in times)),
           "tape_then_compiled_requests": int(sum(tm["tier"] == 2 for tm in times))}
    classes = {}
    for lo, hi in ((8, 16), (16, 24), (24, 30), (30, 37)):
        m = (n_in >= lo) & (n_in < hi)
        if m.any():
            classes["2^%d-2^%d" % (lo, hi - 1)] = {
                "n": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50) * 1e3),
                "p99_ms": float(np.percentile(lat[m], 99) * 1e3)}
    out["latency_by_input_space"] = classes
    compiled = [tm for tm in times if tm["tier"] != 1]
    out["footprint"] = {  # what admission counted (gpu_request_times)
        "vgprs_max": int(max((tm["vgprs"] for tm in compiled), default=0)),
        "waves_per_cu_min": int(min((tm["waves_per_cu"] for tm in times), default=0)),
        "lds_bytes_max": int(max(tm["lds_bytes"] for tm in times)),
        "hbm_bytes_max": int(max(tm["hbm_bytes"] for tm in times)),
        "hbm_bytes_sum": int(sum(tm["hbm_bytes"] for tm in times))}
    if keep is not None:
        keep["requests"] = reqs
        keep["results"] = results
    if check:
        # fault-free pairs and specs have no mismatch; every counterexample separates the two
        # netlists; a planted spec difference hits exactly where a and b are both odd
        bad = 0
        for q, res in zip(reqs, results):
            if not q["faulty"] and res["hits"]:
                bad += 1
            if q["mode"] == t.EN_EQUIV:
                for x in res["hit"]:
                    bad += q["c"].evaluate(x) == q["r"].evaluate(x)
            elif q["faulty"]:
                bad += res["hits"] != 1 << (q["n_in"] - 2)
        out["verdicts_checked"] = bad == 0
    return out


def run_cube_forms(engine, names=("rank12", "gamma_4_4_4"), counts=(1, 16, 256, 1024, 4096),
                   seed=5):
    """Cubes of 2^12 machines in both forms: one workgroup per cube (links across its waves
    through LDS, the pass loop in-kernel; gpu_tune.lds_links 1) and segmented (2), and what the
    cost model picks (0): microseconds per pass for each number of cubes, from the same random
    states (put ignores the read-only cells), and whether the two forms' states agree."""
    from .programs import golden
    rng = np.random.default_rng(seed)
    out = {"workload": "cubeforms", "programs": []}
    for name in names:
        p = golden(name)
        codes = {ll: engine.compile(p, {"lds_links": ll}) for ll in (1, 2)}
        row = {"program": name, "instructions": p.len, "dims": p.dims,
               "default": "workgroup" if engine.compile(p).stats()["resident"] else "segmented",
               "chain_luts": {k: codes[k].stats()["chain"] for k in (1, 2)},
               "segments": codes[2].stats()["segments"], "cubes": []}
        for cubes in counts:
            groups = cubes * p.groups
            mem = rng.integers(0, 1 << 64, size=(groups, p.ncells, 8), dtype=np.uint64)
            mem[:, 0, :] = 0
            mem[:, 1, :] = np.uint64(0xFFFFFFFFFFFFFFFF)
            rr = rng.integers(0, 1 << 64, size=(groups, 8), dtype=np.uint64)
            us, states = {}, {}
            for ll in (1, 2):
                st = codes[ll].state(groups)
                st.put(mem, rr)
                st.run(4)
                n = 256 if cubes <= 16 else 64
                us[ll] = st.run(n) * 1e3 / n
                states[ll] = st.get()
            row["cubes"].append({"cubes": cubes, "workgroup_us_per_pass": us[1],
                                 "segmented_us_per_pass": us[2],
                                 "same_states": all(np.array_equal(a, b) for a, b in
                                                    zip(states[1], states[2]))})
        out["programs"].append(row)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["batch", "life", "hash", "benes", "serve", "polymul",
                                         "tape", "cubeforms"])
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--programs", type=int, default=10000)
    ap.add_argument("--groups", type=int, default=2048)
    ap.add_argument("--passes", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--gens", type=int, default=1000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cubes", type=int, default=4096)
    ap.add_argument("--lds-cells", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--lds-large-cells", type=int, default=0,
                    help="batch: largest program (cells + 1) of the large LDS class "
                         "(0 = what two items per CU allow, 1 = none)")
    ap.add_argument("--untuned", action="store_true",
                    help="batch: the fixed cell thresholds and the window cost model instead of "
                         "the autotuner's measured per-size choices")
    ap.add_argument("--tape-race-ms", type=float, default=0.0,
                    help="serve: tape time up to which a compiled request races on the tape "
                         "(0 = the server's default, negative = never)")
    a = ap.parse_args()
    eng = t.Engine(0)
    if a.workload == "batch":
        out = run_batch(eng, a.programs, a.groups, a.passes, a.seed, not a.no_plain, a.lds_cells,
                        a.waves_per_cu, a.lds_large_cells, tuned=not a.untuned)
    elif a.workload == "life":
        out = run_life(eng, a.gens)
    elif a.workload == "serve":
        out = run_serve(eng, a.requests, a.rate, a.seed, a.streams, tape_race_ms=a.tape_race_ms)
    elif a.workload == "benes":
        out = run_benes(eng, a.cubes, a.seed)
    elif a.workload == "polymul":
        out = run_polymul(eng, a.cubes, a.seed)
    elif a.workload == "tape":
        out = run_tape(eng)
    elif a.workload == "cubeforms":
        out = run_cube_forms(eng)
    else:
        out = run_hash(eng)
    out["device"] = eng.describe()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
