"""Batched form study: 64 TFIM systems, (D, chi) uniform over {(3,24),(4,32),(5,48),(6,64)}, h ~ U[2,4].

    python -m carcassonne_amd.batched_study --systems 64 --dtype bf16 --out result.json

Prints one JSON line: aggregate eigensolves/s of one HIP-graph replay (all systems), the eager setup time
per system, the sequential adaptive mixed and plain FP64 rates on the same systems, the autotuned
workgroup counts per shape bucket, and every system's energy error against its plain FP64 solve.
"""
import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--dtype", default="f16")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--replays", type=int, default=5)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    from .batched import BatchedSolve, draw_specs
    from .device import synchronize
    specs = draw_specs(args.systems, seed=args.seed)
    b = BatchedSolve(specs, dtype=args.dtype)
    b.load()
    b.autotune()
    t0 = time.perf_counter()
    b.build()
    t_prep = time.perf_counter() - t0
    print("built", len(specs), "systems in %.2f s" % t_prep, file=sys.stderr, flush=True)
    b.tune_work()
    print("tuning", b.tuning, file=sys.stderr, flush=True)
    ref = b.reference_energies()
    synchronize()
    # sequential baselines on the same prepared operators
    t0 = time.perf_counter()
    ref2 = b.reference_energies()
    synchronize()
    t_fp64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    for s in b.systems:
        s["solver"].solve(s["x0"])
    synchronize()
    t_mixed = time.perf_counter() - t0
    b.capture()
    b.replay()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.replays):
        b.replay()
    synchronize()
    t_graph = (time.perf_counter() - t0) / args.replays
    e = b.energies()
    errs = [abs(x - r) / abs(r) for x, r in zip(e, ref)]
    n = len(specs)
    res = {
        "workload": "batched form: {} TFIM systems, (D,chi) uniform over (3,24),(4,32),(5,48),(6,64), h~U[2,4], seed {}".format(n, args.seed),
        "dtype": args.dtype, "graph_eigensolves_per_s": n / t_graph, "graph_ms": t_graph * 1e3,
        "setup_ms_per_system": t_prep * 1e3 / n,
        "setup": "stage 1/2 operators + normalization factor + low-precision plan per system, eager (not in the graph)",
        "end_to_end_solves_per_s": n / (t_graph + t_prep),
        "sequential_mixed_eigensolves_per_s": n / t_mixed, "sequential_fp64_eigensolves_per_s": n / t_fp64,
        "max_rel_energy_err": max(errs), "rel_energy_errs": errs,
        "reference_repeatability": max(abs(a - c) / abs(a) for a, c in zip(ref, ref2)),
        "shapes": [s[:2] for s in specs],
        "tuning": {"P{}_alpha{}".format(*k): v for k, v in b.tuning.items()},
    }
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
