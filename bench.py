"""Headline benchmark: exhaustive verify of the 16x16 array multiplier over all 2^32 inputs.

One step = one gpu_enum_exec of mul16 (gen_mul(16): 6,040 NAND/STO instructions, 72 cells) against
the spec "p = a * b": every one of the 2^32 assignments generated in registers from the word index,
the program evaluated 32 machines per 32-bit word, every output compared with a * b, mismatches
counted (must be 0). The kernel is compiled once through hipRTC before the timed region.

value = 1-bit NAND/STO ops per second over the timed steps (ops = instructions x assignments, the
reference's convention, src/enum.c:164). The roofline is the 32-bit integer VALU: a wave64
bitwise instruction issues in 2 cycles on a SIMD (the rate behind the published 157.3 TFLOP/s
FP32 vector figure), so 256 CUs x 4 SIMDs x 32 lanes x 2.4 GHz = 7.864e13 32-bit lane-ops/s,
x 32 bits = 2.516e15 1-bit slots/s. roofline.achieved is the kernel's executed VALU lane-ops
per second: its VALU instructions per launch (SQ_INSTS_VALU, from the committed counter profile
of this exact kernel source) x 64 lanes / the launch's live HIP-event time; frac = achieved /
peak is the VALU issue utilisation. The algorithmic 1-bit ops per second are reported beside it
(algorithmic_1bit_ops_per_s): they exceed the 1-bit slot peak because one v_bitop3_b32 LUT
evaluates several NANDs and a STO costs no instruction.

The CPU baseline is the reference's own AVX2 JIT engine (enum_run built from its sources into
oracle/_ref by oracle/Makefile) on a bounded share of the same enumeration on this host's cores.

--workload life|batch|serve runs one of the secondary workloads of tesseract_amd.workloads instead
and adds beside its device figures a CPU baseline and the check of its results:
  life   Life B3/S23 on the 8192 x 8192 torus; the final board compared bit for bit with the
         oracle's bit-parallel C Life from the same start; that C Life timed as the baseline
  batch  the random-program mix; every program's group-0 cycle lengths against the reference's
         own t1_cycles (timed: the CPU baseline), sampled groups above 0 (>= 50 programs of 4,096
         cells among them) against the oracle's Brent
  serve  the seeded request stream; the reference's enum_run on the first requests (at most
         2^28 assignments each), every result compared with the device's
  tester1  tester 1 (gpu_t1_run) at the reference's defaults; the reference's own t1_run on the
         same programs on host threads, its summary compared with the device's
  polymul  the NTT polynomial product on many cubes; every cube's product checked against the
         oracle's schoolbook product; the reference's jit on one cube, one host thread
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_LANE_OPS = 256 * 4 * 32 * 2.4e9  # MI355X 32-bit VALU lane-ops/s (wave64 issues in 2 cycles)
PEAK_1BIT = PEAK_LANE_OPS * 32        # the same in 1-bit slots


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(threads, parts):
    """The reference's jit-compiled enum_run on share 0 of `parts` of mul16's 2^32 assignments."""
    from oracle import pyref
    if not pyref.available():
        return None
    if not pyref.lib().cpu_has_avx2():
        return None
    prog = pyref.make("gen_mul", 16)
    t0 = time.perf_counter()
    r = pyref.enum_run(0, prog.ptr, spec_text="p = a * b", part=0, part_end=0, parts=parts,
                       threads=threads)
    wall = time.perf_counter() - t0
    assert r["hits"] == 0
    return {"value": r["ops"] / r["seconds"], "unit": "1-bit ops/s", "cores": r["threads"],
            "cpu_model": cpu_model(), "kind": "reference",
            "sample": "mul16 verify p = a * b, share 0 of %d (%d assignments, %.3g ops), "
                      "reference jit enum_run, %.1f s (%.1f s wall)"
                      % (parts, r["tried"], r["ops"], r["seconds"], wall),
            "seconds_full_2p32_extrapolated": r["seconds"] * (1 << 32) / r["tried"]}


def batch_sample_groups(progs, groups, seed=7, big=60, other=40):
    """The group whose cycle lengths the batched run returns, per program: a random group above
    0 for the first `big` programs of 4,096 cells of at most 20,000 instructions and `other`
    further programs of other sizes (checked against the oracle's Brent from the continued
    start stream), group 0 for the rest (checked against the reference's own t1_cycles)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    choice = [0] * len(progs)
    if groups < 2:
        return choice
    big_k = [k for k, p in enumerate(progs) if p.ncells == 4098 and p.len <= 20000][:big]
    rest = [k for k, p in enumerate(progs) if p.ncells != 4098 and p.len <= 20000]
    other_k = list(rng.choice(rest, size=min(other, len(rest)), replace=False)) if rest else []
    for k in big_k + other_k:
        choice[int(k)] = int(rng.integers(1, groups))
    return choice


def batch_checks(keep, passes, threads):
    """The batched run's cycle lengths against the reference and the oracle:
      group 0 of a program: the reference's own t1_cycles (jit-compiled, oracle/_ref) on its own
        random_program of the same seed, on `threads` host threads; those calls, timed, are the
        CPU baseline (like the reference's tester, each compiles its program);
      a sampled group g > 0: the oracle's Brent on that group's start state (the t1_start stream
        continued), on `threads` host threads.
    Returns (check fields, cpu_baseline)."""
    import concurrent.futures as F
    import numpy as np
    from oracle import pyref
    from oracle import pyoracle as O
    from tesseract_amd.workloads import t1_start_group
    progs, seeds, stats, lg = keep["progs"], keep["seeds"], keep["stats"], keep["lam_groups"]
    out = {}
    ref_k = [k for k in range(len(progs)) if lg[k] == 0]
    ora_k = [k for k in range(len(progs)) if lg[k] != 0]
    if pyref.available():
        def theirs(k):
            p = progs[k]
            rp = pyref.make("random_program", p.ncells - 2, p.len, seeds[k])
            t0 = time.perf_counter()
            lam, ran = pyref.t1_cycles(rp, seeds[k], passes)
            return lam, ran, time.perf_counter() - t0
        t0 = time.perf_counter()
        bad, ops, busy = [], 0, 0.0
        with F.ThreadPoolExecutor(threads) as pool:
            for k, (lam, ran, secs) in zip(ref_k, pool.map(theirs, ref_k)):
                if not np.array_equal(stats[k]["lam"], lam):
                    bad.append(k)
                ops += ran * progs[k].len * 512  # src/tester.c:305
                busy += secs
        wall = time.perf_counter() - t0
        out["bit_exact_vs_reference"] = not bad
        out["reference_checked"] = {"programs": len(ref_k), "machines": 512 * len(ref_k),
                                    "differing_programs": bad[:10]}
        base = {"value": ops / wall, "unit": "1-bit ops/s", "cores": threads, "kind": "reference",
                "cpu_model": cpu_model(),
                "sample": "the reference's t1_cycles (compile + cycle search) on group 0 (512 "
                          "machines) of %d programs of the mix, %.1f s on %d threads"
                          % (len(ref_k), wall, threads)}
    else:
        base = None

    def ours(k):
        mem, rr = t1_start_group(progs[k], seeds[k], lg[k])
        return O.brent_lines(progs[k], mem, rr, passes)
    bad = []
    with F.ThreadPoolExecutor(threads) as pool:
        for k, lam in zip(ora_k, pool.map(ours, ora_k)):
            if not np.array_equal(stats[k]["lam"], lam):
                bad.append(k)
    out["bit_exact_vs_oracle"] = not bad
    out["oracle_checked"] = {"programs": len(ora_k),
                             "of_4096_cells": sum(progs[k].ncells == 4098 for k in ora_k),
                             "groups": [int(lg[k]) for k in ora_k[:8]],
                             "differing_programs": bad[:10]}
    return out, base


def serve_checks(eng, reqs, results, seconds, threads):
    """The reference's own enum_run (its netlist parser and jit, oracle/_ref) on the requests in
    arrival order until `seconds` have passed. A request of more than 2^28 assignments runs share 0
    of 2^(inputs - 28) (its --part), so that no single request overruns the bound. Each reference
    result is compared with the device's: the served result itself when the share is the whole
    request, else gpu_enum_run of the same share. Timed, as the CPU baseline: enum_run (its jit
    compile and enumeration), as the device side is timed from a parsed request on; the
    reference's netlist parsing is left out of the rate. Returns (check fields, cpu_baseline)."""
    import ctypes as C
    from oracle import pyref
    import tesseract_amd as t
    if not pyref.available():
        return {}, None
    t0 = time.perf_counter()
    ops = done = tried = whole = 0
    secs = 0.0
    bad = []
    for k, q in enumerate(reqs):
        if time.perf_counter() - t0 > seconds:
            break
        parts = 1 << max(0, q["n_in"] - 28)
        if q["mode"] == t.EN_EQUIV:  # the reference's own netlist loader on the same text
            ra, rb = pyref.Prog(), pyref.Prog()
            e = C.create_string_buffer(512)
            for prog, circ in ((ra, q["c"]), (rb, q["r"])):
                if not pyref.lib().net_parse(circ.bench().encode(), b"net", prog.ptr, e, len(e)):
                    raise ValueError(e.value.decode())
            t1 = time.perf_counter()
            res = pyref.enum_run(t.EN_EQUIV, ra.ptr, rb.ptr, parts=parts, threads=threads)
        else:  # the reference's own multiplier, the same program as the request's
            ra = pyref.make("gen_mul", q["n_in"] // 2)
            t1 = time.perf_counter()
            res = pyref.enum_run(t.EN_VERIFY, ra.ptr, spec_text=q["spec"], parts=parts,
                                 threads=threads)
        secs += time.perf_counter() - t1
        if parts == 1:
            dev = results[k]
            whole += 1
        else:
            dev = eng.enum_run(q["mode"], q["a"], q["b"], spec=q["spec"], part=0, part_end=0,
                               parts=parts)
        if (dev["tried"], dev["hits"], dev["hit"]) != (res["tried"], res["hits"], res["hit"]):
            bad.append(k)
        ops += res["ops"]
        tried += res["tried"]
        done += 1
    checks = {"bit_exact_vs_reference": not bad,
              "reference_checked": {"requests": done, "whole_requests": whole,
                                    "share_0_of_larger": done - whole, "assignments": tried,
                                    "hits": int(sum(results[k]["hits"] for k in range(done))),
                                    "differing_requests": bad[:10]}}
    base = {"value": ops / secs, "unit": "1-bit ops/s", "cores": threads, "kind": "reference",
            "cpu_model": cpu_model(),
            "sample": "reference enum_run (jit compile + enumeration) on the first %d requests in "
                      "order (at most 2^28 assignments of each: %d in all), %.1f s of enum_run in "
                      "%.1f s" % (done, tried, secs, time.perf_counter() - t0)}
    return checks, base


def life_check(life, gens):
    """The final board against the oracle's bit-parallel C Life (64 cells per word, one thread)
    run from the same start for the same generations; that run timed as the baseline."""
    import numpy as np
    from oracle import pyoracle as O
    start = life.start_bits()
    t0 = time.perf_counter()
    want = O.life_words(start, gens)
    cpu_s = time.perf_counter() - t0
    exact = bool(np.array_equal(life.board_bits(), want))
    base = {"value": 8192.0 * 8192 * gens / cpu_s, "unit": "cell updates/s", "cores": 1,
            "kind": "port",
            "sample": "the oracle's bit-parallel C Life, %d generations of the same board, %.1f s"
                      % (gens, cpu_s)}
    return exact, base


def run_tester1(args):
    """Tester 1 on the device and the reference's t1_run on host threads over the same programs:
    the per-size summaries must agree (the longest cycle's seed aside, which depends on the order
    the host threads finish in). With --t1-seconds both run for that long (the reference's
    --seconds mode) and only their rates compare."""
    import ctypes as C
    import tesseract_amd as t
    eng = t.Engine(0)
    cells = [int(c) for c in args.cells.split(",")]
    timed = args.t1_seconds > 0  # the reference's --seconds mode: as many programs as time allows
    programs = 0 if timed else args.programs
    cfg = t.t1_config(cells, args.ratio, programs, args.max_passes, args.seed,
                      seconds=args.t1_seconds)
    # warm-up, untimed: the same programs with cycle searches cut at 2^10 passes (the device
    # buffers and kernels of the timed run, little of its work)
    eng.t1_run(t.t1_config(cells, args.ratio, 64 if timed else args.programs, 1 << 10, args.seed))
    got, gsec = eng.t1_run(cfg)
    ops = sum(d["passes"] * d["len"] * 512 for d in got)  # src/tester.c:302-305
    out = {"workload": "tester1", "cells": cells, "ratio": args.ratio,
           "programs_per_size": [d["programs"] for d in got] if timed else args.programs,
           "unfinished": [d["unfinished"] for d in got],
           "max_passes": args.max_passes, "seed": args.seed,
           "seconds": gsec, "ops": ops, "ops_per_s": ops / gsec,
           "residency": eng.t1_choices(),
           "longest": [d["longest"] for d in got],
           "none_found": [d["hist"][63] for d in got], "cpu_baseline": None}
    if not args.no_cpu:
        from oracle import pyref
        if pyref.available():
            cfg.threads = args.cpu_threads
            ref = (t.T1SizeStruct * len(cells))()
            rsec, th = pyref.t1_run(C.byref(cfg), ref)
            ref = [t.t1_size_dict(x) for x in ref]
            rops = sum(d["passes"] * d["len"] * 512 for d in ref)
            strip = lambda L: [{k: v for k, v in d.items() if k != "longest_seed"} for d in L]
            sample = ("the reference's t1_run for %.1f s (%s programs finished, %s cut off)"
                      % (rsec, [d["programs"] for d in ref], [d["unfinished"] for d in ref])
                      if timed else "the reference's t1_run on the same %d programs, %.2f s"
                      % (len(cells) * args.programs, rsec))
            out["cpu_baseline"] = {"value": rops / rsec, "unit": "1-bit ops/s", "cores": th,
                                   "kind": "reference", "cpu": cpu_model(), "sample": sample}
            if not timed:  # timed runs finish different programs: rates only
                out["same_summary_as_reference"] = strip(got) == strip(ref)
            if not timed and not out["same_summary_as_reference"]:
                print(json.dumps(out), flush=True)
                sys.exit("tester1: the device summary differs from the reference's")
    out["device"] = eng.describe()
    print(json.dumps(out), flush=True)


def cpu_polymul_baseline(keep, seconds):
    """The reference's jit (oracle/_ref) running gen_polymul on cube 0, one host thread, passes
    for about `seconds`: products per second (one product per pass)."""
    from oracle import pyref
    if not pyref.available():
        return None
    pm = keep["polymul"]
    mem, _ = pm.lines(keep["a"][:1], keep["b"][:1])
    theirs = pyref.make("gen_polymul")
    cells = {c: mem[0, c] for c in pm.a + pm.b + list(pm.twiddles)}
    _, t = pyref.jit_run_cells(theirs.ptr, cells, 100, [])
    passes = max(100, int(seconds / (t / 100)))
    got, t = pyref.jit_run_cells(theirs.ptr, cells, passes, pm.c)
    return {"value": passes / t, "unit": "products/s", "cores": 1, "kind": "reference",
            "ops_per_s": theirs.s.len * 512.0 * passes / t,
            "sample": "the reference's jit_run of gen_polymul on one cube, %d passes, %.1f s"
                      % (passes, t)}


def run_workload(args):
    """One secondary workload, its checks and its CPU baseline, as one JSON line."""
    import tesseract_amd as t
    from tesseract_amd import workloads as W
    eng = t.Engine(0)
    keep = {}
    if args.workload == "life":
        out = W.run_life(eng, args.gens, keep=keep)
        if not args.no_cpu:  # warm-up generations count: the board ran 64 + gens
            exact, base = life_check(keep["life"], args.gens + 64)
            out["bit_exact_vs_oracle"] = exact
            out["cpu_baseline"] = base
            if not exact:
                print(json.dumps(out), flush=True)
                sys.exit("life: the final board differs from the oracle's")
    elif args.workload == "batch":
        out = W.run_batch(eng, args.programs, args.groups, args.passes, args.seed,
                          lam_groups=lambda progs: batch_sample_groups(progs, args.groups),
                          keep=keep)
        if not args.no_cpu:
            checks, base = batch_checks(keep, args.passes, args.cpu_threads)
            out.update(checks)
            out["cpu_baseline"] = base
            if not (checks.get("bit_exact_vs_reference", True) and checks["bit_exact_vs_oracle"]):
                print(json.dumps(out), flush=True)
                sys.exit("batch: cycle lengths differ from the reference's or the oracle's")
    elif args.workload == "polymul":
        out = W.run_polymul(eng, args.cubes, args.seed, keep=keep)
        if not args.no_cpu:
            from oracle import pyoracle as O
            exact = bool((keep["c"] == O.polymul(keep["a"], keep["b"])).all())
            out["bit_exact_vs_oracle"] = exact
            out["cpu_baseline"] = cpu_polymul_baseline(keep, args.cpu_seconds)
            if not exact:
                print(json.dumps(out), flush=True)
                sys.exit("polymul: a product differs from the oracle's")
    else:
        out = W.run_serve(eng, args.requests, args.rate, args.seed, keep=keep)
        if not args.no_cpu:
            checks, base = serve_checks(eng, keep["requests"], keep["results"], args.cpu_seconds,
                                        args.cpu_threads)
            out.update(checks)
            out["cpu_baseline"] = base
            if not checks.get("bit_exact_vs_reference", True):
                print(json.dumps(out), flush=True)
                sys.exit("serve: results differ from the reference's")
    out["device"] = eng.describe()
    print(json.dumps(out), flush=True)


def profiled(source_sha16):
    """From the committed counter summary of the ts_enum kernel: HBM bytes per launch
    (FETCH_SIZE + WRITE_SIZE runs scaled by the copy-kernel factor) and VALU instructions per
    launch (SQ_INSTS_VALU; the same under each scheduling strategy the timed choice takes). Each
    only if the summary was measured on the kernel source this run compiles (its sha256 prefix),
    else None."""
    path = os.path.join(ROOT, "profiles", "ts_enum_mul16_verify_counters.txt")
    got = {}
    try:
        with open(path) as f:
            for ln in f:
                if ":" in ln and not ln.startswith(" "):
                    k, v = ln.split(":", 1)
                    got[k.strip()] = v.strip()
    except OSError:
        return None, None
    if got.get("kernel_source_sha16") != source_sha16:
        return None, None
    tr = got.get("traffic_bytes_per_launch")
    va = got.get("valu_insts_per_launch")
    return (float(tr) if tr else None), (float(va) if va else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-parts", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sched", type=int, default=0, choices=[0, 1, 3, 4],
                    help="mul16: the kernel's scheduling strategy (gpu_tune.sched; 0 = timed "
                         "choice of 1 or 3)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="afterwards copy a MIB-MiB device buffer (copy_probe): a kernel of known "
                         "bytes whose FETCH_SIZE/WRITE_SIZE readings scale the kernel's")
    ap.add_argument("--workload", choices=["mul16", "life", "batch", "serve", "tester1", "polymul"],
                    default="mul16")
    ap.add_argument("--cubes", type=int, default=4096, help="polymul: independent products")
    ap.add_argument("--gens", type=int, default=1000, help="life: timed generations")
    ap.add_argument("--programs", type=int, default=None,
                    help="batch: programs of the mix (1000); tester1: programs per size (1024)")
    ap.add_argument("--cells", default="8,16,32,64,128,256", help="tester1: machine sizes")
    ap.add_argument("--ratio", type=float, default=4.0, help="tester1: instructions per cell")
    ap.add_argument("--t1-seconds", type=float, default=0.0,
                    help="tester1: run for this long (the reference's --seconds) instead of a "
                         "fixed number of programs")
    ap.add_argument("--max-passes", type=int, default=1 << 20, help="tester1: cycle search limit")
    ap.add_argument("--groups", type=int, default=2048, help="batch: 512-machine groups each")
    ap.add_argument("--passes", type=int, default=64, help="batch: most passes per group")
    ap.add_argument("--requests", type=int, default=100, help="serve: requests of the stream")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: offered requests/s")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="batch, serve: bound on the CPU baseline's sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU and has no multi-GPU path (each enumeration share would "
              "run on its own GPU with a host-side merge); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "tester1":
        if args.programs is None:
            args.programs = 1024
        run_tester1(args)
        return
    if args.workload != "mul16":
        if args.programs is None:
            args.programs = 1000
        run_workload(args)
        return

    import tesseract_amd as t
    eng = t.Engine(0)
    from tesseract_amd.programs import golden
    prog = golden("mul16")  # the reference's gen_mul(16), a program file
    t0 = time.perf_counter()
    tune = {"sched": args.sched} if args.sched else None
    code = t.EnumCode(eng, t.EN_VERIFY, prog, spec="p = a * b", tune=tune)
    compile_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        r = code.run()
        assert r["hits"] == 0 and r["tried"] == 1 << 32
    kernel_ms = []
    hits = 0
    import ctypes
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = code.run()
        hits += r["hits"]
        kernel_ms.append(eng.last_ms)
    wall = time.perf_counter() - t0
    assert hits == 0
    ops_per_step = prog.len * (1 << 32)
    ms_per_step = wall * 1e3 / args.steps
    value = ops_per_step * args.steps / wall
    kavg = sum(kernel_ms) / len(kernel_ms)
    algorithmic = ops_per_step / (kavg * 1e-3)
    valu_rate = eng.valu_rate()  # measured v_bitop3_b32 lane-ops/s on this device
    import hashlib
    sha16 = hashlib.sha256(t.enum_source(t.EN_VERIFY, prog, spec="p = a * b").encode()).hexdigest()[:16]
    traffic, valu_insts = profiled(sha16)
    achieved = valu_insts * 64 / (kavg * 1e-3) if valu_insts else None
    # the same kernel shape with a planted spec difference (a * b + (a & b & 1)): 2^30 of the 2^32
    # assignments mismatch, every one counted and listed through the wave-level hit reports
    planted = t.EnumCode(eng, t.EN_VERIFY, prog, spec="p = a * b + (a & b & 1)")
    planted.run()
    rp = planted.run()
    assert rp["hits"] == 1 << 30 and rp["hit"][0] == (1 << 31) + (1 << 15)
    planted_ms = eng.last_ms
    line = {
        "metric": "1-bit NAND/STO ops/s on 1 GPU (% int-VALU peak); mul16 2^32 verify sec",
        "value": value,
        "unit": "1-bit ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (every 2^32 input assignment, generated in registers)",
        "config": {"workload": "mul16 exhaustive verify: gen_mul(16), 6040 NAND/STO instructions, "
                               "spec p = a * b, all 2^32 assignments",
                   "program_instructions": prog.len, "assignments": 1 << 32,
                   "verify_seconds": ms_per_step / 1e3, "hits": 0,
                   "planted_spec_difference": {"spec": "p = a * b + (a & b & 1)",
                                               "hits": rp["hits"], "kernel_ms": planted_ms},
                   "compile_seconds": compile_s, "device": eng.describe(),
                   "scheduling": code.stats()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_LANE_OPS,
                     "unit": "32-bit VALU lane-op/s",
                     "frac": achieved / PEAK_LANE_OPS if achieved else None, "traffic": traffic,
                     "kernel": "ts_enum", "kernel_ms": kavg, "kernel_source_sha16": sha16,
                     "valu_insts_per_launch": valu_insts,
                     "algorithmic_1bit_ops_per_s": algorithmic,
                     "peak_1bit_slots_per_s": PEAK_1BIT,
                     "algorithmic_1bit_ops_per_1bit_slot": algorithmic / PEAK_1BIT,
                     "probe_valu_lane_ops_per_s": valu_rate},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_threads, args.cpu_parts)
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        n = args.copy_calibration << 20
        rate = eng.hbm_rate(n)  # copy_probe: 4 launches, each reads n bytes and writes n bytes
        print(json.dumps({"copy_probe_bytes_read_per_launch": n,
                          "copy_probe_bytes_written_per_launch": n, "copy_bytes_per_s": rate}),
              flush=True)


if __name__ == "__main__":
    main()
