"""The batched form: many independent systems of uneven (D, chi) and fields solved by one HIP-graph launch.

No reference counterpart (the reference solves one system at a time).  `prepare` allocates every
system's buffers and plans (stage-1/2 operators, normalization factor, FP64 packed slabs, low-precision
slabs and plan) and runs the setup once eagerly.  `capture` records every system's solve into one HIP
graph, each system on its own stream branch forked from the capture stream, so one `replay` launches every
solve and the GPU overlaps the small systems' kernels.  With setup=True (the default) each branch first
rebuilds its system's setup from the environment tensors into those buffers -- stage 1/2, the
normalization factor, the FP64 and low-precision slabs -- so a replay is a whole solve of every system,
and `load_environments` + `replay` solves new environments of the same shapes; with setup=False only the
eigensolves are recorded.  The eigensolve is the fixed-work mixed solve (low-precision Lanczos plus FP64
Rayleigh-Ritz refinement, no host reads: MixedSolver.solve_fixed), which also writes a convergence record
per system; `convergence` reads the records after a replay and `resolve_unconverged` solves the flagged
systems again with the adaptive solve.  An autotuner picks, per shape bucket, the low-precision kernel's
workgroup count by timing candidates on the bucket's first system, and the fixed work (Lanczos steps,
refinement passes) from adaptive solves of every system of the bucket with a safety margin.  The plain
FP64 solve of every system is the reference for the batched energies.
"""
import time

import numpy as np
import torch

from . import pipeline, synthetic as syn
from .device import empty, synchronize
from .precision import LowPrecisionOperator, MixedSolver
from .tensors._2d import sparse as S2

SHAPES = ((3, 24), (4, 32), (5, 48), (6, 64))


def draw_specs(n, seed=0, shapes=SHAPES, h_range=(2.0, 4.0)):
    """(D, chi, h, env_seed) of n systems: shapes uniform over `shapes`, h uniform, seeded."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = rng.integers(0, len(shapes), n)
    hs = rng.uniform(h_range[0], h_range[1], n)
    return [(shapes[i][0], shapes[i][1], float(h), 1000 + k) for k, (i, h) in enumerate(zip(idx, hs))]


def _time_lp(op, x, y, reps=20):
    for _ in range(3):
        op.apply_flat(x, y)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        op.apply_flat(x, y)
    synchronize()
    return (time.perf_counter() - t0) / reps


def autotune_groups(H_operator, dtype, candidates=None):
    """Workgroup count of the low-precision stage-3 kernel for this operator's shape: the fastest of the
    candidates (alpha per workgroup 2..32) by measured matvec time."""
    na = H_operator.n_alpha
    if candidates is None:
        candidates = sorted({max(1, na // k) for k in (2, 4, 8, 16, 32)})
    x = empty((H_operator.n,)).fill_(1.0)
    y = empty((H_operator.n,))
    best, best_t = None, None
    timings = {}
    for g in candidates:
        op = LowPrecisionOperator(H_operator, dtype, groups=g)
        t = _time_lp(op, x, y)
        timings[g] = t
        if best_t is None or t < best_t:
            best, best_t = g, t
    return best, timings


class BatchedSolve:
    def __init__(self, specs, dtype="f16", model="ising"):
        self.specs, self.dtype, self.model = list(specs), dtype, model
        self.systems = []
        self.tuning = {}
        self.graph = None

    def load(self):
        """Generate every environment on the host and upload it (the input data; not part of a solve)."""
        self.inputs = []
        for (D, chi, h, seed) in self.specs:
            c, s, st = syn.make_environment(D, chi, model=self.model, seed=seed)
            corners, sides, opc = pipeline.device_environment(c, s, syn.operator_center(self.model, h=h))
            x0 = torch.from_numpy(st.ravel()).to(empty((1,)).device)
            self.inputs.append((corners, sides, opc, x0))
        synchronize()

    def autotune(self):
        """Workgroup count of the low-precision kernel per shape bucket (on the bucket's first system)."""
        for (corners, sides, opc, _) in self.inputs:
            H, _ = S2.formExpectationAndNormalizationMultipliers(corners, sides, opc)
            key = (H.operator.P, H.operator.n_alpha)
            if key not in self.tuning:
                g, timings = autotune_groups(H.operator, self.dtype)
                self.tuning[key] = dict(groups=g, timings_us={k: v * 1e6 for k, v in timings.items()})
        synchronize()

    def build(self):
        """Per system: stage 1/2 operators, normalization factor, low-precision plan (the solve's setup)."""
        self.systems = []
        for spec, (corners, sides, opc, x0) in zip(self.specs, self.inputs):
            H, N = S2.formExpectationAndNormalizationMultipliers(corners, sides, opc)
            key = (H.operator.P, H.operator.n_alpha)
            solver = MixedSolver(H.operator, N.operator, dtype=self.dtype, groups=self.tuning[key]["groups"])
            self.systems.append(dict(spec=spec, H=H, N=N, solver=solver, x0=x0, key=key))
        synchronize()
        for sysm in self.systems:
            sysm["solver"].factor.check()

    def prepare(self):
        self.load()
        self.autotune()
        self.build()

    def tune_work(self, margin_iters=8, margin_outer=2, max_outer=None):
        """Fixed work per shape bucket: the most low-precision Lanczos steps and FP64 passes that an
        adaptive solve of any system of the bucket needed, plus margins.  BF16's correction equation
        gains less per FP64 pass than FP16's, so its adaptive solves may take up to 24 passes."""
        if max_outer is None:
            max_outer = 24 if self.dtype == "bf16" else 8
        for key in self.tuning:
            lp, passes = 0, 0
            for sysm in (s for s in self.systems if s["key"] == key):
                _, _, info = sysm["solver"].solve(sysm["x0"], max_outer=max_outer)
                lp, passes = max(lp, int(info["lowprec_iterations"])), max(passes, int(info["fp64_passes"]))
            self.tuning[key]["lp_iters"] = int(lp * 1.25) + margin_iters
            self.tuning[key]["outer"] = passes + margin_outer
        synchronize()

    def _solve_all(self, streams, setup):
        cur = torch.cuda.current_stream()
        outs, keep = [], []
        for sysm, st, (corners, sides, opc, _) in zip(self.systems, streams, self.inputs):
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                if setup:
                    H, N = S2.formExpectationAndNormalizationMultipliers(corners, sides, opc)
                    sysm["solver"].refresh(H.operator, N.operator)
                    keep.append((H, N))
                t = self.tuning[sysm["key"]]
                outs.append(sysm["solver"].solve_fixed(sysm["x0"], lp_iters=t["lp_iters"], outer=t["outer"]))
        for st in streams:
            cur.wait_stream(st)
        self._setup_outputs = keep
        return outs

    def capture(self, setup=True):
        """Capture every system's solve into one HIP graph (each system on its own stream branch); with
        setup, each branch rebuilds its system's setup from the environment tensors first."""
        self.setup_in_graph = bool(setup)
        self._streams = [torch.cuda.Stream() for _ in self.systems]
        self._solve_all(self._streams, setup)      # warm-up outside the graph: first-use allocations
        synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = self._solve_all(self._streams, setup)
        synchronize()

    def load_environments(self, envs):
        """Copies new environments -- (corners, sides) lists of tag -> host array, the shapes and tags of
        the loaded ones, as synthetic.make_environment returns -- into the tensors the graph reads; the
        next replay solves them (needs a graph captured with setup)."""
        if not getattr(self, "setup_in_graph", False):
            raise RuntimeError("load_environments needs a graph captured with setup=True")
        for (corners, sides, _, _), (c_new, s_new) in zip(self.inputs, envs):
            for dev, host in zip(list(corners) + list(sides), list(c_new) + list(s_new)):
                for key, arr in host.items():
                    t = dev[pipeline.synthetic_tag(key)]
                    if tuple(t.shape) != tuple(arr.shape):
                        raise ValueError("load_environments needs the loaded shapes")
                    t.__dict__.pop("_host", None)
                    t._t.copy_(torch.from_numpy(np.ascontiguousarray(arr)))
        synchronize()

    def replay(self):
        self.graph.replay()

    def energies(self):
        synchronize()
        return [float(rho.item()) for _, rho, _ in self.outputs]

    def convergence(self, tol_change=1e-11, tol_residual=1e-7):
        """Per system, from the graph's in-device record (read back after a replay): whether the last FP64
        pass moved the energy by at most tol_change relative and the final relative residual
        |H x - rho N x| / (|rho| |N x|) is at most tol_residual; returns (ok, change, residual) tuples."""
        synchronize()
        out = []
        for _, _, rec in self.outputs:
            h = rec.tolist()
            rho, rr, nn = h[-3], h[-2], h[-1]
            delta = abs(h[-3] - h[-4]) / abs(rho) if len(h) >= 4 else float("inf")
            resid = float(np.sqrt(max(rr, 0.0) / max(nn, 1e-300))) / max(abs(rho), 1e-300)
            out.append((bool(delta <= tol_change and resid <= tol_residual), delta, resid))
        return out

    def _fresh_operators(self, i):
        """System i's (H, N) built eagerly from its current environment tensors (new stage-2 tensors, so no
        cached packing or conversion of an earlier environment is reused)."""
        corners, sides, opc, _ = self.inputs[i]
        H, N = S2.formExpectationAndNormalizationMultipliers(corners, sides, opc)
        return H.operator, N.operator

    def resolve_unconverged(self, conv=None):
        """The guard's fallback: every system the record flags is solved again eagerly by the adaptive
        mixed solve with the FP16 operator (its energy replaces the graph's).  Returns {index: energy}."""
        conv = self.convergence() if conv is None else conv
        out = {}
        for i, (ok, _, _) in enumerate(conv):
            if ok:
                continue
            H, N = self._fresh_operators(i)
            _, rho, _ = MixedSolver(H, N, dtype="f16").solve(self.systems[i]["x0"])
            out[i] = rho
        return out

    def reference_energies(self):
        """Plain FP64 solve of every system's current environment (the reference energies)."""
        from .solver import GroundStateSolver
        out = []
        for i, sysm in enumerate(self.systems):
            H, N = self._fresh_operators(i)
            gs = GroundStateSolver(H, N)
            x, _, _, _ = gs.solve(sysm["x0"])
            out.append(gs.rayleigh_quotient(x))
        return out
