"""Benchmark of the short-range nonbonded force path on one MI355X.

Workload: 32,000 rigid-geometry waters (96,000 atoms) in a periodic 98.6 A
box, LJ (O-O) + Ewald real-space erfc(0.32 r)/r, C2 switching 7.5 -> 9 A,
cell list rebuilt from the positions on every step.  One step = one full
force evaluation (cell hashing, counting sort, pair kernel, reductions) on
positions already resident in HBM.  ns/day = steps/s * 1 fs * 86400 / 1e6;
pair evals/s = P_alg * steps/s with P_alg the in-cutoff pair count.

Prints one JSON line.  The mixed-precision path gives `value`; the FP64 GPU
path runs the same workload beside it.  The CPU baseline is the FP64 C
restatement of the reference path (oracle/pmnb_oracle.c), one host thread,
on a bounded sample of steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("ns/day + pair evals/s, 96k-atom LJ+Ewald-real; "
          "mixed-prec vs FP64 on GPU")
FLOP_PER_PAIR = 90        # LJ + Ewald real (A&S) + shared C2 + atomic virial
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6   # MI355X published FP64 vector peak
DT_FS = 1.0


def ns_per_day(steps_per_s):
    return steps_per_s * DT_FS * 86400.0 / 1.0e6


def time_gpu(w, precision, steps, warmup, flags, torch):
    force = w.make_force(precision)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(np.ascontiguousarray(w.positions)).to(dev)
    frc = torch.zeros_like(pos)
    out = torch.zeros(20, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    for _ in range(warmup):
        force.evaluate_device(w.topology, pos, frc, out, flags, stream=sp)
    torch.cuda.synchronize()
    force.kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        force.evaluate_device(w.topology, pos, frc, out, flags, stream=sp)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    bms, kms, launches = force.kernel_timing(False)
    # one more evaluation for the forces and the pair count of one step
    frc.zero_()
    out.zero_()
    force.evaluate_device(w.topology, pos, frc, out, flags, stream=sp)
    torch.cuda.synchronize()
    npairs = force.last_pair_count()
    f = frc.cpu().numpy()
    force.close()
    return {"ms_per_step": 1e3 * dt / steps, "kernel_ms": kms / max(launches, 1),
            "build_ms": bms / max(launches, 1),
            "launches": launches, "pairs": npairs, "forces": f,
            "energies": out[:2].cpu().numpy()}


def cpu_baseline(w, budget_s):
    from oracle import oracle as O
    O.build()
    s = O.System.from_workload(w, flags=1)
    times = []
    npairs = 0
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        r = O.evaluate(s)
        times.append(time.perf_counter() - t0)
        npairs = r.n_pairs
        if time.perf_counter() - t_start > budget_s or len(times) >= 20:
            break
    t = float(np.median(times))
    return {"value": ns_per_day(1.0 / t), "unit": "ns/day", "cores": 1,
            "kind": "port",
            "pair_evals_per_s": npairs / t,
            "sample": "%d full force evaluations of the 96k-atom workload "
                      "(median %.3f s each), FP64 C restatement of "
                      "NonbondedCutoffSystemForce, cellsize rc/2, 1 thread, "
                      "gcc -O2" % (len(times), t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=0,
                    help="steps of the FP64 path (default: steps/4)")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU baseline sampling (0 = skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB "
                         "(3 times): the reference kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of "
              "this build (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from protomoladdon_amd import systems

    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("copied %d MiB x3" % args.copy_mib, file=sys.stderr, flush=True)
        del src, dst

    print("generating workload ...", file=sys.stderr, flush=True)
    w = systems.config1()
    flags = 1  # atomic virial on, as in the per-pair FLOP convention

    print("mixed precision ...", file=sys.stderr, flush=True)
    mx = time_gpu(w, "mixed", args.steps, args.warmup, flags, torch)
    fp64_steps = args.fp64_steps or max(10, args.steps // 4)
    print("fp64 ...", file=sys.stderr, flush=True)
    f64 = time_gpu(w, "fp64", fp64_steps, max(2, args.warmup // 4), flags,
                   torch)

    sps = 1e3 / mx["ms_per_step"]
    sps64 = 1e3 / f64["ms_per_step"]
    P = mx["pairs"]
    ref = f64["forces"]
    rms = float(np.sqrt(np.mean(np.sum((mx["forces"] - ref) ** 2, 1)) /
                        np.mean(np.sum(ref ** 2, 1))))
    achieved = P * FLOP_PER_PAIR / (mx["kernel_ms"] * 1e-3) / 1e12
    achieved64 = P * FLOP_PER_PAIR / (f64["kernel_ms"] * 1e-3) / 1e12

    cpu = None
    if args.cpu_budget > 0:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, args.cpu_budget)

    line = {
        "metric": METRIC,
        "value": ns_per_day(sps),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "water96k_ewald: 32,000 waters (96,000 atoms), 98.6 A "
                        "periodic box, LJ O-O + Ewald real space alpha 0.32, "
                        "C2 7.5->9 A, cell list rebuilt every step, virial on",
            "n_atoms": w.n_atoms,
            "pairs_per_step": P,
        },
        "pair_evals_per_s": P * sps,
        "fp64_gpu": {"value": ns_per_day(sps64), "unit": "ns/day",
                     "ms_per_step": f64["ms_per_step"],
                     "pair_evals_per_s": f64["pairs"] * sps64,
                     "pair_kernel_ms": f64["kernel_ms"],
                     "list_build_ms": f64["build_ms"],
                     "pair_kernel_tflops": achieved64,
                     "pairs_per_step": f64["pairs"]},
        "speedup_vs_fp64": sps / sps64,
        "rms_rel_force_err_vs_fp64": rms,
        "roofline": {"bound": "valu", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_tq<mixed, Ewald-real, C2>",
                     "kernel_ms": mx["kernel_ms"],
                     "list_build_ms": mx["build_ms"],
                     "achieved_incl_list_build": P * FLOP_PER_PAIR / (
                         (mx["kernel_ms"] + mx["build_ms"]) * 1e-3) / 1e12,
                     "flop_per_pair": FLOP_PER_PAIR},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
