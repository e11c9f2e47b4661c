"""Benchmark of the short-range nonbonded force path on one MI355X.

Workload: 32,000 rigid-geometry waters (96,000 atoms) in a periodic 98.6 A
box, LJ (O-O) + Ewald real-space erfc(0.32 r)/r, C2 switching 7.5 -> 9 A.
One step = one whole 1 fs leapfrog step on the device: half kicks, drift,
SHAKE and RATTLE, the cell list rebuilt from the positions and the force
evaluated (pair kernel and reductions), with positions, velocities and
forces resident in HBM.  ns/day = steps/s * 1 fs * 86400 / 1e6 of that MD
step is `value`.  pair evals/s = P_alg * force evaluations/s, P_alg the
in-cutoff pair count, comes from force-only evaluations at fixed positions.

Prints one JSON line.  The mixed-precision path gives `value`; the FP64 GPU
path runs the same workload beside it.  Extra keys hold the batched-replica
and force-matching workloads in compact form.  The CPU baseline is the FP64
C restatement of the reference path (oracle/pmnb_oracle.c), one host thread
and all host threads, on a bounded sample of steps; the same leg checks the
GPU forces against that FP64 oracle (RMS relative error per workload).
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
FLOP_PER_PAIR_FORCES = 73  # the same without the energy sums (2) and virial (15)
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6   # MI355X published FP64 vector peak
DT_FS = 1.0
GRAPH_CHUNK = 20          # MD steps per captured run (replayed)
MD_KERNEL_STEPS = 20      # steps of the run whose pair kernel is timed


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


def time_md(w, precision, steps, warmup, flags, torch):
    """LeapfrogIntegrator::run on the device: 1 fs steps, SHAKE and RATTLE
    at their defaults (epsilon 1e-5, 30 iterations), the cutoff force as the
    force group; velocities Maxwell-Boltzmann at 300 K (seeded)."""
    from protomoladdon_amd.integrator import (LeapfrogIntegrator,
                                              topology_constraints)
    t = w.topology
    cons = topology_constraints(t)
    rng = np.random.default_rng(1234)
    v0 = rng.standard_normal((w.n_atoms, 3)) * np.sqrt(
        0.001987191 * 300.0 / t.masses)[:, None]
    force = w.make_force(precision)
    lf = LeapfrogIntegrator(DT_FS, force, shake=(1e-5, 30), rattle=(1e-5, 30))
    dev = torch.device("cuda:0")
    x = torch.from_numpy(np.ascontiguousarray(w.positions)).to(dev)
    v = torch.from_numpy(v0).to(dev)
    f = torch.zeros_like(x)
    out = torch.zeros(20, dtype=torch.float64, device=dev)
    sp = torch.cuda.current_stream(dev).cuda_stream
    lf.initialize(t, x, f, out, cons, flags=flags, stream=sp)
    lf.run(t, x, v, f, out, warmup, cons, flags=flags, stream=sp)
    torch.cuda.synchronize()
    lf.record_iterations(1 << 16)
    t0 = time.perf_counter()
    lf.run(t, x, v, f, out, steps, cons, flags=flags, stream=sp)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its, n_calls, unconverged = lf.record_iterations(0)
    # call order: SHAKE, then (RATTLE, SHAKE) per later step, then RATTLE
    shake_its, rattle_its = its[0::2], its[1::2]
    # the pair kernel of the dynamics' evaluations (HIP events on the launch
    # stream): run(1), whose one evaluation forms energies, then run(K),
    # K - 1 evaluations without energies and the last with them
    force.kernel_timing(True)
    lf.run(t, x, v, f, out, 1, cons, flags=flags, stream=sp)
    _, k_one, n_one = force.kernel_timing(False)
    force.kernel_timing(True)
    lf.run(t, x, v, f, out, MD_KERNEL_STEPS, cons, flags=flags, stream=sp)
    _, k_run, n_run = force.kernel_timing(False)
    k_energy = k_one / max(n_one, 1)
    k_forces = (k_run - k_energy) / max(n_run - 1, 1)
    # the same steps as replays of one captured run(chunk) (pmnb_lf_run_graph)
    chunk = GRAPH_CHUNK if steps >= GRAPH_CHUNK else steps
    reps = max(1, steps // chunk)
    lf.run_graph(t, x, v, f, out, chunk, 1, cons, flags=flags, stream=sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lf.run_graph(t, x, v, f, out, chunk, reps, cons, flags=flags, stream=sp)
    torch.cuda.synchronize()
    dtg = time.perf_counter() - t0
    _, _, unconverged_g = lf.record_iterations(0)
    lf.close()
    force.close()
    ms_eager = 1e3 * dt / steps
    ms_graph = 1e3 * dtg / (chunk * reps)
    # the headline is the faster of the two ways of running the same steps
    graph_faster = ms_graph < ms_eager
    ms = ms_graph if graph_faster else ms_eager
    return {"value": ns_per_day(1e3 / ms), "unit": "ns/day",
            "ms_per_step": ms, "steps": chunk * reps if graph_faster else steps,
            "mode": ("graph: %d replays of one captured run(%d)" % (reps, chunk)
                     if graph_faster else "eager: one run(%d) call" % steps),
            "eager_ms_per_step": ms_eager,
            "graph_ms_per_step": ms_graph,
            "graph": "%d replays of one captured run(%d)" % (reps, chunk),
            "pair_kernel_ms_forces_only": k_forces,
            "pair_kernel_ms_with_energies": k_energy,
            "mean_shake_sweeps": float(shake_its.mean()),
            "mean_rattle_sweeps": float(rattle_its.mean()),
            "unconverged_calls": int(unconverged + unconverged_g)}


def time_ewald_kspace(w, alpha, accuracy, reps, torch):
    """The k-space terms of NonbondedFullEwaldSystemForce (reciprocal sum,
    exclusion correction, self and charged-system energies) on the
    device; returns ms per evaluation and the k-vector count."""
    from protomoladdon_amd.ewald import NonbondedFullEwaldKSpace
    force = w.make_force("fp64")
    ks = NonbondedFullEwaldKSpace(force, alpha, accuracy)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(np.ascontiguousarray(w.positions)).to(dev)
    f = torch.zeros_like(x)
    out = torch.zeros(20, dtype=torch.float64, device=dev)
    sp = torch.cuda.current_stream(dev).cuda_stream
    ks.evaluate_device(w.topology, x, f, out, 1, stream=sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ks.evaluate_device(w.topology, x, f, out, 1, stream=sp)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    a, rc, kc, nk = ks.info(w.topology)
    ks.close()
    force.close()
    return {"ms": ms, "alpha": a, "accuracy": accuracy, "kc": kc,
            "n_kvectors": nk, "n_atoms": w.n_atoms,
            "atom_kvector_terms_per_s": w.n_atoms * nk / (ms * 1e-3)}


def time_pme_kspace(w, grid, order, alpha, reps, torch, ewald_alpha=None):
    """The k-space terms of NonbondedPMEwaldSystemForce (B-spline
    spreading, FFTs, scalar sum, gather, correction, self and
    charged-system terms) on the device: ms per evaluation, and the RMS
    relative difference of its k-space forces from the direct Ewald sum's
    at the same alpha (device, accuracy 1e-8)."""
    from protomoladdon_amd.ewald import NonbondedFullEwaldKSpace
    from protomoladdon_amd.pme import NonbondedPMEwaldKSpace
    force = w.make_force("fp64")
    ks = NonbondedPMEwaldKSpace(force, grid, 9.0, order=order, alpha=alpha)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(np.ascontiguousarray(w.positions)).to(dev)
    f = torch.zeros_like(x)
    out = torch.zeros(20, dtype=torch.float64, device=dev)
    sp = torch.cuda.current_stream(dev).cuda_stream
    ks.evaluate_device(w.topology, x, f, out, 1, stream=sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ks.evaluate_device(w.topology, x, f, out, 1, stream=sp)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    f.zero_()
    out.zero_()
    ks.evaluate_device(w.topology, x, f, out, 1, stream=sp)
    torch.cuda.synchronize()
    fp = f.cpu().numpy()
    a, kc = ks.info(w.topology)
    ks.close()
    ew = NonbondedFullEwaldKSpace(force, alpha, 1e-8)
    fe = torch.zeros_like(x)
    oe = torch.zeros(20, dtype=torch.float64, device=dev)
    ew.evaluate_device(w.topology, x, fe, oe, 1, stream=sp)
    torch.cuda.synchronize()
    fr = fe.cpu().numpy()
    ew.close()
    force.close()
    return {"ms": ms, "grid": list(grid), "order": order, "alpha": a,
            "n_atoms": w.n_atoms,
            "rms_rel_force_diff_vs_direct_ewald": rms_rel(fp, fr),
            "energy_diff_vs_direct_ewald": float(out[1].item() - oe[1].item())}


def host_cpu():
    """(model name, logical CPUs visible to this process)."""
    import platform
    import subprocess
    model = platform.processor()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def rms_rel(f, ref):
    return float(np.sqrt(np.mean(np.sum((f - ref) ** 2, 1)) /
                         np.mean(np.sum(ref ** 2, 1))))


def cpu_baseline(w, budget_s, checks, batch=None):
    """The FP64 C restatement timed on the host (one thread, then all
    threads) on a bounded sample of force evaluations of `w`; then, as the
    checker, the RMS relative force error of each GPU result in `checks`
    ({name: (workload, {precision: forces})}) and of each replica of
    `batch` ([(workload, mixed forces)]) against the FP64 oracle."""
    from oracle import oracle as O
    O.build()
    s = O.System.from_workload(w, flags=1)
    times = []
    ref = None
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        r = O.evaluate(s)
        times.append(time.perf_counter() - t0)
        ref = r if ref is None else ref
        if time.perf_counter() - t_start > budget_s or len(times) >= 20:
            break
    t1 = float(np.median(times))
    model, ncpu = host_cpu()
    threads = min(ncpu, 16)
    tt = []
    for _ in range(max(3, min(10, len(times)))):
        t0 = time.perf_counter()
        _, _, _, n_thr, used = O.evaluate_threads(s, threads)
        tt.append(time.perf_counter() - t0)
    tall = float(np.median(tt))
    base = {"value": ns_per_day(1.0 / t1), "unit": "ns/day", "cores": 1,
            "kind": "port",
            "pair_evals_per_s": ref.n_pairs / t1,
            "sample": "%d force evaluations of the 96k-atom workload (median "
                      "%.3f s each; one force evaluation per 1 fs step), FP64 "
                      "C restatement of NonbondedCutoffSystemForce, cellsize "
                      "rc/2, 1 thread, gcc -O2" % (len(times), t1),
            "all_threads": {"value": ns_per_day(1.0 / tall), "unit": "ns/day",
                            "cores": used, "s_per_eval": tall,
                            "pair_evals_per_s": n_thr / tall,
                            "sample": "%d evaluations, OpenMP over cells"
                                      % len(tt)},
            "host_cpu": model, "nproc": ncpu}
    acc = {}
    for name, (wk, results) in checks.items():
        r = ref if wk is w else O.evaluate(O.System.from_workload(wk, flags=1))
        acc[name] = {p: rms_rel(f, r.forces) for p, f in results.items()}
    if batch:
        errs = [rms_rel(f, O.evaluate(O.System.from_workload(
            wk, flags=1)).forces) for wk, f in batch]
        acc["batch512"] = {"mixed_max": max(errs),
                           "mixed_mean": float(np.mean(errs))}
    return base, acc


def bench_batch(args, torch):
    """512 water replicas (workload 3): one graph replay per step against
    per-replica sequential launches of the single-system force."""
    from protomoladdon_amd import systems
    from protomoladdon_amd.batch import BatchedCutoffForce, OUT_SLOTS
    print("generating 512 replicas ...", file=sys.stderr, flush=True)
    sizes = systems.config3_sizes()
    ws = [systems.config3_replica(i, n) for i, n in enumerate(sizes)]
    R = len(ws)
    bf = BatchedCutoffForce.from_workloads(ws, precision="mixed")
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(np.concatenate([w.positions for w in ws])).to(dev)
    frc = torch.zeros_like(pos)
    out = torch.zeros(R * OUT_SLOTS, dtype=torch.float64, device=dev)
    bf.evaluate_device(pos, frc, out, 1)
    torch.cuda.synchronize()
    print("autotuning ...", file=sys.stderr, flush=True)
    shape, grid, tune_ms = bf.autotune(pos, frc, out, 1)
    bf.run_graph(pos, frc, out, 1, n_steps=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bf.run_graph(pos, frc, out, 1, n_steps=args.steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    P = bf.pair_count()
    fb = frc.cpu().numpy()
    # per-replica sequential launches of the single-system force
    print("sequential launches ...", file=sys.stderr, flush=True)
    singles = [w.make_force("mixed") for w in ws]
    off = bf.offsets
    frs = torch.zeros_like(pos)
    outs = torch.zeros_like(out)
    sp = torch.cuda.current_stream(dev).cuda_stream

    def seq_step():
        frs.zero_()
        outs.zero_()
        for r, (w, f) in enumerate(zip(ws, singles)):
            f.evaluate_device(w.topology, pos[off[r]:off[r + 1]],
                              frs[off[r]:off[r + 1]],
                              outs[OUT_SLOTS * r:OUT_SLOTS * (r + 1)], 1,
                              stream=sp)
    seq_step()
    torch.cuda.synchronize()
    nseq = max(3, args.steps // 50)
    t0 = time.perf_counter()
    for _ in range(nseq):
        seq_step()
    torch.cuda.synchronize()
    dseq = (time.perf_counter() - t0) / nseq
    diff = float(np.abs(frs.cpu().numpy() - fb).max())
    for f in singles:
        f.close()
    md = bench_batch_md(args, torch, ws, bf, pos)
    bf.close()
    args._batch_check = (ws, off, fb)
    return {
        "metric": "batched replicas: ns/day per replica and aggregate pair "
                  "evals/s, 512 waters boxes LJ+Ewald-real, one graph replay "
                  "per step",
        "value": ns_per_day(1.0 / dt), "unit": "ns/day per replica",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "512 periodic water replicas, 167-2667 "
                               "molecules each (density 0.0334/A^3), LJ + "
                               "Ewald real alpha 0.32, C2 7.5->9 A",
                   "n_replicas": R, "n_atoms": int(off[-1]),
                   "pairs_per_step": P},
        "aggregate_ns_per_day": R * ns_per_day(1.0 / dt),
        "pair_evals_per_s": P / dt,
        "sequential_ms_per_step": 1e3 * dseq,
        "speedup_vs_sequential": dseq / dt,
        "max_abs_force_diff_vs_sequential": diff,
        "autotune": {"shape": shape, "grid_blocks": grid, "ms": tune_ms},
        "md": md,
    }


def bench_batch_md(args, torch, ws, bf, pos0):
    """Leapfrog 1 fs + SHAKE/RATTLE of every replica (pmnb_batch_lf_*): one
    captured run(n) of all replicas, launched once."""
    from protomoladdon_amd.batch import BatchedLeapfrog, OUT_SLOTS
    from protomoladdon_amd.integrator import topology_constraints
    R = len(ws)
    masses = [w.topology.masses for w in ws]
    cons = [topology_constraints(w.topology) for w in ws]
    rng = np.random.default_rng(4321)
    v0 = np.concatenate([rng.standard_normal((w.n_atoms, 3)) * np.sqrt(
        0.001987191 * 300.0 / w.topology.masses)[:, None] for w in ws])
    lf = BatchedLeapfrog(DT_FS, bf, masses, cons, shake=(1e-5, 30),
                         rattle=(1e-5, 30))
    x = pos0.clone()
    v = torch.from_numpy(v0).to(x.device)
    f = torch.zeros_like(x)
    out = torch.zeros(R * OUT_SLOTS, dtype=torch.float64, device=x.device)
    lf.initialize(x, f, out, 1)
    lf.run(x, v, f, out, 2, 1)
    torch.cuda.synchronize()
    n = args.md_steps
    print("batched dynamics: capture of run(%d) ..." % n, file=sys.stderr,
          flush=True)
    lf.record_iterations(1 << 20)
    t0 = time.perf_counter()
    lf.run_graph(x, v, f, out, n, 1, 1)   # capture + instantiate + run
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    lf.run_graph(x, v, f, out, n, 1, 1)   # one replay: n steps
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / n
    its, n_calls, unconverged = lf.record_iterations(0)
    lf.close()
    return {"steps_per_launch": n, "ms_per_step": 1e3 * dt,
            "ns_per_day_per_replica": ns_per_day(1.0 / dt),
            "aggregate_ns_per_day": R * ns_per_day(1.0 / dt),
            "first_call_s_incl_capture": t_first,
            "unconverged_calls": int(unconverged),
            "integrator": "Leapfrog 1 fs + SHAKE/RATTLE (1e-5, 30) per "
                          "replica, one graph launch of %d steps" % n}


def bench_forcematch(args, torch):
    """Force matching on the 250k-atom chain system (workload 4): loss and
    its gradient with respect to positions, charges, sigma and epsilon."""
    from protomoladdon_amd import grad as G
    from protomoladdon_amd import systems
    print("generating chains ...", file=sys.stderr, flush=True)
    w = systems.config2()
    t = w.topology
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1234)
    q0 = t.scaled_charges / 18.2226123264
    s0 = np.array(systems.CHAIN_SIGMA)
    e0 = np.array(systems.CHAIN_EPS)
    qt = q0 * (1 + 0.05 * rng.normal(size=q0.shape))
    st = s0 * (1 + 0.05 * rng.normal(size=4))
    et = e0 * (1 + 0.05 * rng.normal(size=4))
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)
    pos = T(w.positions)
    res = {}
    grads = {}
    for prec in ("mixed", "fp64"):
        f = w.make_force(prec)
        with torch.no_grad():
            target = G.cutoff_forces(f, t, pos, T(qt), T(st), T(et), 0.5)

        def step():
            x = pos.clone().requires_grad_(True)
            q = T(q0).requires_grad_(True)
            s = T(s0).requires_grad_(True)
            e = T(e0).requires_grad_(True)
            F = G.cutoff_forces(f, t, x, q, s, e, 0.5)
            L = G.force_matching_loss(F, target)
            L.backward()
            return L, x.grad, q.grad, s.grad, e.grad
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        n = max(3, args.steps // 20)
        t0 = time.perf_counter()
        for _ in range(n):
            out = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / n
        res[prec] = {"ms_per_gradient": 1e3 * dt, "loss": float(out[0])}
        grads[prec] = [g.cpu().numpy() for g in out[1:]]
        f.close()

    def rel(a, b):
        return float(np.sqrt(np.mean((a - b) ** 2) / np.mean(b ** 2)))
    errs = {k: rel(a, b) for k, a, b in zip(("x", "q", "sigma", "epsilon"),
                                            grads["mixed"], grads["fp64"])}
    # FP64 central differences (h = 1e-6) on 64 random parameters
    f = w.make_force("fp64")
    with torch.no_grad():
        target = G.cutoff_forces(f, t, pos, T(qt), T(st), T(et), 0.5)

    def loss(th):
        with torch.no_grad():
            F = G.cutoff_forces(f, t, pos, T(th[8:]), T(th[:4]), T(th[4:8]),
                                0.5)
            return float(G.force_matching_loss(F, target))
    theta = np.concatenate([s0, e0, q0])
    an = np.concatenate([grads["fp64"][2], grads["fp64"][3], grads["fp64"][1]])
    pick = np.concatenate([np.arange(8), 8 + rng.choice(len(q0), 56,
                                                        replace=False)])
    h = 1e-6
    worst = 0.0
    gmax = np.abs(an[pick]).max()
    for k in pick:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        worst = max(worst, abs(fd - an[k]) / max(abs(fd), 1e-3 * gmax))
    f.close()
    ms = res["mixed"]["ms_per_gradient"]
    return {
        "metric": "force-matching gradients/s (loss + adjoint), 250k-atom "
                  "chains LJ+Ewald-real",
        "value": 1e3 / ms, "unit": "gradients/s", "n_gpus": 1,
        "steps": max(3, args.steps // 20), "warmup": 2, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "12,500 chains x 20 atoms (250,000), 138 A "
                               "box, LJ (4 types, LB, 1-4 x 0.5) + Ewald "
                               "real alpha 0.34, C2 8.5->10 A; target from "
                               "N(0, 0.05^2) relative parameter noise",
                   "n_atoms": w.n_atoms},
        "fp64": res["fp64"], "mixed": res["mixed"],
        "speedup_vs_fp64": res["fp64"]["ms_per_gradient"] / ms,
        "rms_rel_grad_err_mixed_vs_fp64": errs,
        "fd_check_64_params_worst_rel_err": worst,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=0,
                    help="steps of the FP64 path (default: steps/4)")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU baseline sampling (0 = skip the "
                         "CPU leg and the oracle accuracy checks)")
    ap.add_argument("--workload", default="water96k",
                    choices=["water96k", "batch512", "forcematch250k"],
                    help="water96k: the headline line (with compact batch512 "
                         "and forcematch250k summaries); batch512 and "
                         "forcematch250k alone: their own full lines")
    ap.add_argument("--batch-md-steps", type=int, default=1000,
                    help="leapfrog steps of the 512 replicas in one graph "
                         "launch")
    ap.add_argument("--no-extras", action="store_true",
                    help="leave out the batch512 / forcematch250k summaries")
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

    if args.workload == "batch512":
        args.md_steps = args.batch_md_steps
        print(json.dumps(bench_batch(args, torch)), flush=True)
        return
    if args.workload == "forcematch250k":
        print(json.dumps(bench_forcematch(args, torch)), flush=True)
        return

    print("generating workload ...", file=sys.stderr, flush=True)
    w = systems.config1()
    flags = 1  # atomic virial on, as in the per-pair FLOP convention

    print("force, mixed ...", file=sys.stderr, flush=True)
    mx = time_gpu(w, "mixed", args.steps, args.warmup, flags, torch)
    fp64_steps = args.fp64_steps or max(10, args.steps // 4)
    print("force, fp64 ...", file=sys.stderr, flush=True)
    f64 = time_gpu(w, "fp64", fp64_steps, max(2, args.warmup // 4), flags,
                   torch)

    print("dynamics ...", file=sys.stderr, flush=True)
    md_mx = time_md(w, "mixed", args.steps, max(2, args.warmup // 4), flags,
                    torch)
    md_64 = time_md(w, "fp64", fp64_steps, 2, flags, torch)

    print("ewald k-space ...", file=sys.stderr, flush=True)
    ewald = {"water3k": time_ewald_kspace(systems.config0(), -1.0, 1e-5, 20,
                                          torch),
             "water96k": time_ewald_kspace(w, 0.32, 1e-5, 3, torch)}
    print("pme k-space ...", file=sys.stderr, flush=True)
    pme = {"water96k": time_pme_kspace(w, (96, 96, 96), 4, 0.32, 20, torch)}

    # the host-array entry point (pmnb_evaluate: positions in, forces out
    # through PCIe, synchronous), for the record; never the headline value.
    # Timed as ProtoMol calls it (the C ABI, caller-owned arrays) and through
    # the Python force object.
    from protomoladdon_amd.topology import ScalarStructure
    fh = w.make_force("mixed")
    fr = np.zeros((w.n_atoms, 3))
    es = ScalarStructure(virial=True)
    fh.evaluate(w.topology, w.positions, fr, es)
    nh = max(20, args.steps // 5)
    pos = np.ascontiguousarray(w.positions)
    e2, vir, mvir = np.zeros(2), np.zeros(9), np.zeros(9)
    lib, ctx = fh._lib, fh._ctx
    args_c = (ctx, pos.ctypes.data, fr.ctypes.data, e2.ctypes.data,
              vir.ctypes.data, mvir.ctypes.data, None, 1)
    lib.pmnb_evaluate(*args_c)
    t0 = time.perf_counter()
    for _ in range(nh):
        lib.pmnb_evaluate(*args_c)
    host_ms = 1e3 * (time.perf_counter() - t0) / nh
    t0 = time.perf_counter()
    for _ in range(nh):
        fh.evaluate(w.topology, w.positions, fr, es)
    host_py_ms = 1e3 * (time.perf_counter() - t0) / nh
    fh.close()

    # forces of the other single-system workloads for the accuracy check
    checks = {"water96k": (w, {"mixed": mx["forces"], "fp64": f64["forces"]})}
    if args.cpu_budget > 0:
        for name, wk in (("water3k", systems.config0()),
                         ("chains250k", systems.config2())):
            print("forces of %s ..." % name, file=sys.stderr, flush=True)
            checks[name] = (wk, {p: time_gpu(wk, p, 1, 1, flags, torch)[
                "forces"] for p in ("mixed", "fp64")})

    extras = {}
    if not args.no_extras:
        ea = argparse.Namespace(steps=max(20, args.steps // 4),
                                warmup=max(3, args.warmup // 4),
                                md_steps=args.batch_md_steps)
        print("batch512 ...", file=sys.stderr, flush=True)
        b = bench_batch(ea, torch)
        extras["batch512"] = {k: b[k] for k in (
            "value", "unit", "ms_per_step", "aggregate_ns_per_day",
            "pair_evals_per_s", "sequential_ms_per_step",
            "speedup_vs_sequential", "max_abs_force_diff_vs_sequential",
            "autotune", "config", "md")}
        print("forcematch250k ...", file=sys.stderr, flush=True)
        fm = bench_forcematch(ea, torch)
        extras["forcematch250k"] = {k: fm[k] for k in (
            "value", "unit", "ms_per_step", "fp64", "mixed",
            "speedup_vs_fp64", "rms_rel_grad_err_mixed_vs_fp64",
            "fd_check_64_params_worst_rel_err", "config")}
        if args.cpu_budget > 0:
            ws, off, fb = ea._batch_check
            checks["batch512"] = [(wk, fb[off[r]:off[r + 1]])
                                  for r, wk in enumerate(ws)]

    sps = 1e3 / mx["ms_per_step"]
    sps64 = 1e3 / f64["ms_per_step"]
    P = mx["pairs"]
    achieved = P * FLOP_PER_PAIR / (mx["kernel_ms"] * 1e-3) / 1e12
    achieved64 = P * FLOP_PER_PAIR / (f64["kernel_ms"] * 1e-3) / 1e12
    md_achieved = P * FLOP_PER_PAIR_FORCES / (
        md_mx["pair_kernel_ms_forces_only"] * 1e-3) / 1e12

    # HBM bytes per launch of the pair kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (scaled by a copy kernel of known size)
    traffic = traffic_fo = None
    tpath = os.path.join(ROOT, "profiles", "k_tile_water96k_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        traffic = tj.get("hbm_bytes_per_launch")
        traffic_fo = tj.get("forces_only", {}).get("hbm_bytes_per_launch")

    cpu, accuracy = None, None
    if args.cpu_budget > 0:
        print("cpu baseline and oracle checks ...", file=sys.stderr,
              flush=True)
        batch = checks.pop("batch512", None)
        cpu, accuracy = cpu_baseline(w, args.cpu_budget, checks, batch)

    line = {
        "metric": METRIC,
        "value": md_mx["value"],
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": md_mx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "water96k_ewald MD: 32,000 waters (96,000 atoms), "
                        "98.6 A periodic box, LJ O-O + Ewald real space "
                        "alpha 0.32, C2 7.5->9 A; one step = leapfrog 1 fs "
                        "+ SHAKE/RATTLE + cell-list rebuild + force; energies "
                        "and virial (flag on) summed for the last step of "
                        "each run() call, the values the caller reads",
            "n_atoms": w.n_atoms,
            "pairs_per_step": P,
        },
        "pair_evals_per_s": P * sps,
        "force_only": {"value": ns_per_day(sps), "unit": "ns/day",
                       "ms_per_step": mx["ms_per_step"],
                       "pair_kernel_ms": mx["kernel_ms"],
                       "cell_build_ms": mx["build_ms"]},
        "fp64_gpu": {"value": md_64["value"], "unit": "ns/day",
                     "ms_per_step": md_64["ms_per_step"],
                     "force_only_ns_per_day": ns_per_day(sps64),
                     "force_only_ms_per_step": f64["ms_per_step"],
                     "pair_evals_per_s": f64["pairs"] * sps64,
                     "pair_kernel_ms": f64["kernel_ms"],
                     "cell_build_ms": f64["build_ms"],
                     "pair_kernel_tflops": achieved64,
                     "pairs_per_step": f64["pairs"]},
        "speedup_vs_fp64": {"md": md_mx["value"] / md_64["value"],
                            "force_only": sps / sps64,
                            "pair_kernel": f64["kernel_ms"] / mx["kernel_ms"]},
        "md": {"integrator": "Leapfrog 1 fs + SHAKE/RATTLE (1e-5, 30), "
                             "whole step on the device; energies and virial "
                             "of the last step of each run() call",
               "mixed": md_mx, "fp64": md_64},
        "ewald_kspace": ewald,
        "pme_kspace": pme,
        "host_api_ms_per_step": host_ms,
        "host_api_python_ms_per_step": host_py_ms,
        "accuracy_rms_rel_force_err_vs_fp64_oracle": accuracy,
        "roofline": {"bound": "valu", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_unit": "bytes per launch (HBM, rocprofv3)",
                     "peak_without_packed_fp32": PEAK_FP64_TFLOPS,
                     "frac_without_packed_fp32": achieved / PEAK_FP64_TFLOPS,
                     "kernel": "k_tile<mixed, Ewald-real, C2>",
                     "kernel_ms": mx["kernel_ms"],
                     "cell_build_ms": mx["build_ms"],
                     "achieved_incl_cell_build": P * FLOP_PER_PAIR / (
                         (mx["kernel_ms"] + mx["build_ms"]) * 1e-3) / 1e12,
                     "flop_per_pair": FLOP_PER_PAIR,
                     # the dynamics' pair kernel (every evaluation of a
                     # run but its last): forces only, 73 of the 90 FLOP
                     "md_kernel": {
                         "kernel": "k_tile<mixed, Ewald-real, C2, forces only>",
                         "kernel_ms": md_mx["pair_kernel_ms_forces_only"],
                         "flop_per_pair": FLOP_PER_PAIR_FORCES,
                         "achieved": md_achieved,
                         "frac": md_achieved / PEAK_FP32_TFLOPS,
                         "traffic": traffic_fo,
                         "frac_without_packed_fp32":
                             md_achieved / PEAK_FP64_TFLOPS}},
        "cpu_baseline": cpu,
    }
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
