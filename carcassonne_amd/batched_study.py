"""Batched form study: 64 TFIM systems, (D, chi) uniform over {(3,24),(4,32),(5,48),(6,64)}, h ~ U[2,4].

    python -m carcassonne_amd.batched_study --systems 64 --dtype bf16 --out result.json

Prints one JSON line with end-to-end rates (setup + eigensolve, inputs resident) of
* one HIP-graph replay with every system's setup inside the graph (the batched form),
* eager setup per system followed by one replay of a graph holding only the eigensolves,
* a sequential loop of adaptive mixed solves (pipeline.solve_mixed) and of plain FP64 solves
  (pipeline.solve) over the same systems,
plus the autotuned work per shape bucket, the convergence record's verdicts and every system's energy
error against its plain FP64 solve.
"""
import argparse
import json
import sys
import time


def _timed(fn, reps=1):
    from .device import synchronize
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    synchronize()
    return (time.perf_counter() - t0) / reps


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--dtype", default="bf16")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--replays", type=int, default=5)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    from . import pipeline
    from .batched import BatchedSolve, draw_specs
    specs = draw_specs(args.systems, seed=args.seed)
    n = len(specs)
    b = BatchedSolve(specs, dtype=args.dtype)
    b.load()
    b.autotune()
    t_build = _timed(b.build)
    print("built", n, "systems in %.2f s" % t_build, file=sys.stderr, flush=True)
    b.tune_work()
    print("tuning", b.tuning, file=sys.stderr, flush=True)
    ref = b.reference_energies()
    max_outer = 24 if args.dtype == "bf16" else 8
    seq_mixed = lambda: [pipeline.solve_mixed(c, s, o, x, dtype=args.dtype, max_outer=max_outer) for c, s, o, x in b.inputs]
    seq_fp64 = lambda: [pipeline.solve(c, s, o, x) for c, s, o, x in b.inputs]
    seq_mixed()                                   # first-use allocations and stage-1/2 graphs
    t_seq_mixed = _timed(seq_mixed)
    t_seq_fp64 = _timed(seq_fp64)
    print("sequential done", file=sys.stderr, flush=True)
    b.capture(setup=False)
    b.replay()
    t_graph_solves = _timed(b.replay, args.replays)
    t_setup = _timed(lambda: [s["solver"].refresh(*b._fresh_operators(i)) for i, s in enumerate(b.systems)])
    b.capture(setup=True)
    b.replay()
    t_graph_full = _timed(b.replay, args.replays)
    e = b.energies()
    conv = b.convergence()
    errs = [abs(x - r) / abs(r) for x, r in zip(e, ref)]
    res = {
        "workload": "batched form: {} TFIM systems, (D,chi) uniform over (3,24),(4,32),(5,48),(6,64), h~U[2,4], "
                    "seed {}".format(n, args.seed),
        "dtype": args.dtype,
        "graph_with_setup_solves_per_s": n / t_graph_full, "graph_with_setup_ms": t_graph_full * 1e3,
        "eager_setup_plus_graph_solves_per_s": n / (t_setup + t_graph_solves),
        "eager_setup_ms": t_setup * 1e3, "graph_eigensolves_ms": t_graph_solves * 1e3,
        "sequential_mixed_solves_per_s": n / t_seq_mixed, "sequential_fp64_solves_per_s": n / t_seq_fp64,
        "max_rel_energy_err": max(errs), "rel_energy_errs": errs,
        "flagged_by_record": [i for i, c in enumerate(conv) if not c[0]],
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
