"""Train and evaluate the learned warm start on one environment family (one GPU).

    python -m carcassonne_amd.warm_start_study --pairs 20000 --D 4 --chi 32 --out result.json

Generates (features, converged state) pairs with the plain FP64 solve over h ~ U[2, 4] and seeded
perturbations of a base environment, trains the 3-layer FP16 MLP with the HIP kernels, and on held-out
perturbation seeds (h walking 2.0 -> 4.0 in order, as consecutive sweep steps) reports the Lanczos
iterations to a 1e-10 residual and the end-to-end solve times from random, previous-step and learned
starts.  Prints one JSON line; optionally saves the trained weights (.npz).
"""
import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--D", type=int, default=4)
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--model", default="ising")
    ap.add_argument("--sigma", type=float, default=0.2)
    ap.add_argument("--base-seed", type=int, default=11)
    ap.add_argument("--hidden", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=40)
    ap.add_argument("--out", default=None)
    ap.add_argument("--save", default=None)
    ap.add_argument("--chains", type=int, default=0,
                    help="also train the correction network on this many walks of consecutive environments")
    args = ap.parse_args(argv)

    from .device import synchronize
    from .learned import Family, WarmStartMLP, evaluate_starts, generate_pairs, _pad64
    fam = Family(args.D, args.chi, model=args.model, base_seed=args.base_seed, sigma=args.sigma)
    net = WarmStartMLP(8, args.chi, _pad64(fam.n), hidden=args.hidden, seed=0)
    rng = np.random.default_rng(args.base_seed)
    hs = rng.uniform(2.0, 4.0, args.pairs)
    t0 = time.perf_counter()
    X, T, iters = [], [], []
    chunk = 1000
    for c0 in range(0, args.pairs, chunk):     # progress lines while the pairs are produced
        seeds = list(range(1000 + c0, 1000 + min(args.pairs, c0 + chunk)))
        x, t, it = generate_pairs(fam, net, seeds, hs[c0:c0 + len(seeds)])
        X.append(x)
        T.append(t)
        iters += it
        print("pairs", c0 + len(seeds), "of", args.pairs, "%.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    import torch
    X, T = torch.cat(X), torch.cat(T)
    synchronize()
    t_gen = time.perf_counter() - t0
    t1 = time.perf_counter()
    hist = net.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    synchronize()
    t_train = time.perf_counter() - t1
    print("trained: loss", ["%.3g" % h for h in hist], "%.1f s" % t_train, file=sys.stderr, flush=True)
    cnet = None
    if args.chains:
        from .learned import generate_correction_pairs
        t2 = time.perf_counter()
        cnet = WarmStartMLP.for_residual(fam.n, _pad64(fam.n), hidden=args.hidden, seed=1)
        Xc, Tc, npairs = generate_correction_pairs(fam, cnet, args.chains, np.random.default_rng(7), seed0=500000)
        synchronize()
        print("correction pairs", npairs, "%.1f s" % (time.perf_counter() - t2), file=sys.stderr, flush=True)
        chist = cnet.train(Xc, Tc, epochs=args.epochs, batch=min(args.batch, npairs), lr=args.lr)
        print("correction net: loss", ["%.3g" % h for h in chist], file=sys.stderr, flush=True)
    eval_seeds = list(range(900000, 900000 + args.eval))
    ev = evaluate_starts(fam, net, eval_seeds, np.linspace(2.0, 4.0, args.eval), correction_net=cnet)
    kinds = ("random", "previous", "learned", "combined") + (("corrected",) if cnet is not None else ())
    res = {
        "workload": "learned warm start, 2D TFIM family D={} chi={} sigma={} h~U[2,4]".format(args.D, args.chi, args.sigma),
        "pairs": args.pairs, "generation_s": t_gen, "train_s": t_train, "train_loss": hist,
        # means over the environments that have a previous one (all but the first), for every start
        "mean_iterations": {k: float(np.mean(ev[k][1:] if k in ("random", "learned") else ev[k])) for k in kinds},
        "solves_per_s": {k: 1e3 / float(np.mean(ev["ms_" + k][1:] if k in ("random", "learned") else ev["ms_" + k]))
                         for k in kinds},
        "mean_overlap_with_converged": {k: float(np.mean(ev["overlap_" + k])) for k in kinds if k != "random"},
        "max_energy_spread": float(max(ev["energy_spread"])),
        "iterations": {k: ev[k] for k in kinds},
    }
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    if args.save:
        np.savez(args.save, **net.state())


if __name__ == "__main__":
    main()
