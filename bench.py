"""Headline benchmark: exhaustive verify of the 16x16 array multiplier over all 2^32 inputs.

One step = one gpu_enum_exec of mul16 (gen_mul(16): 6,040 NAND/STO instructions, 72 cells) against
the spec "p = a * b": every one of the 2^32 assignments generated in registers from the word index,
the program evaluated 32 machines per 32-bit word, every output compared with a * b, mismatches
counted (must be 0). The kernel is compiled once through hipRTC before the timed region.

value = 1-bit NAND/STO ops per second over the timed steps (ops = instructions x assignments, the
reference's convention, src/enum.c:164). The roofline is the 32-bit integer VALU:
256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz x 32 bits = 1.258e15 1-bit ops/s (the ratio can pass 1:
v_bitop3_b32 evaluates several NANDs at once).

The CPU baseline is the reference's own AVX2 JIT engine (enum_run built from its sources into
oracle/_ref by oracle/Makefile) on a bounded share of the same enumeration on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_1BIT = 256 * 4 * 16 * 2.4e9 * 32  # MI355X 32-bit integer VALU lane-ops/s x 32 bits


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
            "kind": "reference",
            "sample": "mul16 verify p = a * b, share 0 of %d (%d assignments, %.3g ops), "
                      "reference jit enum_run, %.1f s (%.1f s wall)"
                      % (parts, r["tried"], r["ops"], r["seconds"], wall),
            "seconds_full_2p32_extrapolated": r["seconds"] * (1 << 32) / r["tried"]}


def profiled_traffic():
    """HBM bytes per ts_enum launch from the committed counter summary (FETCH_SIZE + WRITE_SIZE
    runs, scaled by the copy-kernel factor), or None if it is absent."""
    path = os.path.join(ROOT, "profiles", "ts_enum_mul16_verify_counters.txt")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-parts", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="afterwards copy a MIB-MiB device buffer (copy_probe): a kernel of known "
                         "bytes whose FETCH_SIZE/WRITE_SIZE readings scale the kernel's")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each enumeration share "
              "would run on its own GPU with a host-side merge); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import tesseract_amd as t
    eng = t.Engine(0)
    prog = t.gen_mul(16)
    t0 = time.perf_counter()
    code = t.EnumCode(eng, t.EN_VERIFY, prog, spec="p = a * b")
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
    achieved = ops_per_step / (kavg * 1e-3)
    valu_rate = eng.valu_rate()  # measured v_bitop3_b32 lane-ops/s on this device
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
                   "compile_seconds": compile_s, "device": eng.describe()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_1BIT,
                     "unit": "1-bit op/s", "frac": achieved / PEAK_1BIT, "traffic": None,
                     "kernel": "ts_enum", "kernel_ms": kavg,
                     "measured_valu_lane_ops_per_s": valu_rate,
                     "frac_of_measured": achieved / (valu_rate * 32)},
        "cpu_baseline": None,
    }
    traffic = profiled_traffic()
    if traffic is not None:
        line["roofline"]["traffic"] = traffic
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
