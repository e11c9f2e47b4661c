"""Times the reference's own NonbondedCutoffSystemForce (oracle/_ref, built
from the reference sources by `make -C oracle ref`) against the FP64 C
restatement (oracle/pmnb_oracle.c) on the same host cores, one thread each,
and prints one JSON line with the median seconds per evaluation of both and
their ratio.  Needs the reference build, so it runs where the reference
sources are (never on a machine without them).

    python oracle/calibrate.py [--workload water96k|water3k|chains250k]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, sys.path[0] is this folder, whose oracle.py would shadow
# the package of the same name
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from protomoladdon_amd import systems  # noqa: E402


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="water96k",
                    choices=["water96k", "water3k", "chains250k"])
    ap.add_argument("--repeats", type=int, default=7)
    args = ap.parse_args()
    make = {"water96k": systems.config1, "water3k": systems.config0,
            "chains250k": systems.config2}[args.workload]
    w = make()
    s = O.System.from_workload(w, flags=1)
    O.build()
    if not O.ref_available():
        O.build_ref()
    # reference: R timed evaluations in one process (taskset to one core)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        O.write_ref_input(s, fin)
        res = subprocess.run(["taskset", "-c", "0", O.REF_DRIVER, fin, fout,
                              str(args.repeats)], capture_output=True,
                             text=True, check=True)
    ref_ms = [float(l.split()[1]) for l in res.stdout.splitlines()
              if l.startswith("ms ")]
    # restatement on the same core
    os.sched_setaffinity(0, {0})
    t = []
    for _ in range(args.repeats):
        t0 = time.perf_counter()
        r = O.evaluate(s)
        t.append(time.perf_counter() - t0)
    ref_s = float(np.median(ref_ms)) * 1e-3
    port_s = float(np.median(t))
    print(json.dumps({
        "workload": w.name, "n_atoms": w.n_atoms, "pairs": r.n_pairs,
        "reference_s_per_eval": ref_s, "restatement_s_per_eval": port_s,
        "reference_over_restatement": ref_s / port_s,
        "threads": 1, "cpu": cpu_model(), "nproc": os.cpu_count(),
        "repeats": args.repeats}))


if __name__ == "__main__":
    main()
