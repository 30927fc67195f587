#!/usr/bin/env python
"""Benchmark: ground-state solves/s of the center-site effective Hamiltonian on one MI355X.

Workload: 2D transverse-field Ising model H = -sum ZZ - 3.0 sum X, center bond D = 6, fused
environment bond chi = 64, d = 2, w = 3 operator channels on the sides; seeded synthetic environment
(carcassonne_amd.synthetic, seed 0) resident in HBM before timing.  One step = one solve:
stage 1/2 on the GPU, the stage-3 operators, the normalization factor (N_env = L L^T, W = L^-1), the
eigensolve to a 1e-10 relative energy, and the FP64 Rayleigh quotient of the result; the energy is
checked against the oracle's eigh energy (tests/golden).

Two forms are timed on the same inputs:
  * mixed-f16 (the headline `value`): FP16 stage-3 matvecs with FP32 accumulation inside the Lanczos,
    FP64 Jacobi-Davidson refinement, FP64 Rayleigh-Ritz energy;
  * fp64: the plain FP64 Lanczos form (`plain_fp64_*` fields).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = dict(D=6, chi=64, d=2, model="ising", h=3.0, seed=0)
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
COPY_CALIBRATION_DOUBLES = 64 << 20   # 512 MiB read + 512 MiB written per launch
TIMING_EVERY = 8   # stage-3 launches per HIP-event-timed launch inside the timed region


def measured_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc summaries (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[kernel_key]
    except (OSError, KeyError, ValueError):
        return None
    return 1024.0 * (t["fetch_kib_per_launch"] * t["fetch_factor"] + t["write_kib_per_launch"] * t["write_factor"])


def golden_energy():
    path = os.path.join(ROOT, "tests", "golden", "energies.json")
    key = "D{D}_chi{chi}_{model}_seed{seed}".format(**WORKLOAD)
    with open(path) as f:
        return json.load(f)[key]["ground_energy"]


def _oracle_operator(D, chi, seed, h=3.0):
    from carcassonne_amd import synthetic as syn
    from oracle import sparse2d
    from tests.oracle_bridge import to_oracle
    c, s, st = syn.make_environment(D, chi, model=WORKLOAD["model"], seed=seed)
    return sparse2d.form_multipliers(*to_oracle(c, s, syn.operator_center(WORKLOAD["model"], h=h))), st


def _median_time(fn, reps):
    import numpy as np
    fn()                                   # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(gpu_iterations, sample_iterations=6, reps=5):
    """The oracle (NumPy/SciPy FP64) timed on this host, as a bounded sample:

    * headline size (D=6, chi=64): stage 1/2, the Cholesky factor of N_env and `sample_iterations` Lanczos
      steps, the solve extrapolated to the plain FP64 GPU form's iteration count (all BLAS threads; a full
      CPU solve there takes about half a minute);
    * D=4, chi=32, h=3: complete solves -- the reference's relaxOver (restarted Arnoldi, k=3, 100
      multiplications, its cost rule choosing dense H and LU of N) and the residual-controlled Lanczos
      port to 1e-10 -- one warm-up, then the median of `reps` repetitions, with one BLAS thread and with
      all of them."""
    import numpy as np
    from scipy.linalg import cholesky, solve_triangular
    from threadpoolctl import threadpool_info, threadpool_limits
    from oracle import dense, solver as OS
    threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])

    def headline_sample():
        t0 = time.perf_counter()
        op, st = _oracle_operator(WORKLOAD["D"], WORKLOAD["chi"], WORKLOAD["seed"], WORKLOAD["h"])
        Lc = cholesky(dense.normalization_env_matrix(op.AN, op.BN), lower=True)
        t1 = time.perf_counter()
        d, ne = op.d, Lc.shape[0]
        q = (Lc.T @ st.reshape(ne, d)).ravel()
        Q = [q / np.linalg.norm(q)]
        t2 = time.perf_counter()
        for _ in range(sample_iterations):
            z = solve_triangular(Lc, Q[-1].reshape(ne, d), lower=True, trans='T')
            w = solve_triangular(Lc, op.apply_H(z).reshape(ne, d), lower=True).ravel()
            Qm = np.array(Q)
            w -= Qm.T @ (Qm @ w)
            w -= Qm.T @ (Qm @ w)
            Q.append(w / np.linalg.norm(w))
        t3 = time.perf_counter()
        return (t1 - t0) + gpu_iterations * (t3 - t2) / sample_iterations, t3 - t0, t3 - t2

    t_solve, t_sample, t_steps = headline_sample()
    op4, st4 = _oracle_operator(4, 32, 0)
    H4, N4 = op4.matrix_H(), op4.matrix_N()
    use_dense = OS.is_cheaper_to_form_matrix(6, op4.cost_H, op4.cost_H_form, op4.n)

    def relax():
        OS.relax_over(st4.ravel(), lambda x: op4.apply_H(x).ravel(), lambda x: op4.apply_N(x).ravel(), max_mults=100,
                      tol=1e-7, H_matrix=op4.matrix_H() if use_dense else None, N_matrix=op4.matrix_N(),
                      use_dense_H=use_dense, use_lu_N=True)

    def lanczos():
        OS.lanczos_generalized(lambda x: op4.apply_H(x).ravel(), op4.matrix_N(), st4.ravel(), tol=1e-10)

    d4 = {}
    for name, fn in (("relaxover", relax), ("lanczos", lanczos)):
        with threadpool_limits(1):
            d4[name + "_1_thread_s"] = _median_time(fn, reps)
        d4[name + "_all_threads_s"] = _median_time(fn, reps)
    d4_mv = {}
    x4 = st4.ravel()
    for label, limit in (("1_thread", 1), ("all_threads", None)):
        if limit is None:
            d4_mv[label] = 1.0 / _median_time(lambda: op4.apply_H(x4), reps)
        else:
            with threadpool_limits(limit):
                d4_mv[label] = 1.0 / _median_time(lambda: op4.apply_H(x4), reps)
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": int(threads), "kind": "port",
            "cpu_model": cpu_model, "env_threads": {k: os.environ.get(k) for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS")},
            "d4_chi32_matvecs_per_s": dict(d4_mv, statistic="median of {} after one warm-up".format(reps)),
            "d6_chi64_lanczos_steps_per_s_all_threads": sample_iterations / t_steps,
            "sample": "oracle stage 1/2 + Cholesky of N_env + {} Lanczos steps at D=6, chi=64 ({:.1f} s); solve time "
                      "extrapolated to the plain FP64 form's {} iterations".format(sample_iterations, t_sample,
                                                                                 gpu_iterations),
            "host_cpus": len(os.sched_getaffinity(0)),
            "d4_chi32_full_solves": dict(d4, reps=reps, statistic="median after one warm-up",
                                         solves_per_s={k[:-2]: 1.0 / v for k, v in d4.items()})}


def measured_peaks():
    """The machine's own ceilings, measured here beside the published figures the roofline uses: a 512 MiB
    copy (unary_kernel, read + write) and a 512 MiB read through global->LDS loads (HBM), the same read of an
    85 MB buffer repeated (resident in the 256 MB Infinity Cache, the size of the FP16 kernel's slabs), and
    the dense MFMA rates (f16 16x16x32 and f64 16x16x4, eight independent accumulators per wave)."""
    import torch
    from carcassonne_amd import _lib
    st = torch.cuda.current_stream()

    def seconds(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps

    n = COPY_CALIBRATION_DOUBLES
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    sink = torch.empty(1024 * 256, dtype=torch.float32, device="cuda")
    out = {}
    t = seconds(lambda: _lib.call("cc_unary_f64", st.cuda_stream, 3, n, src.data_ptr(), 1.0, dst.data_ptr()), 5)
    out["hbm_copy_gbs"] = 2 * n * 8 / t / 1e9
    t = seconds(lambda: _lib.call("cc_calibration_lds_read", st.cuda_stream, n * 8, src.data_ptr(), sink.data_ptr(), 1024), 5)
    out["hbm_read_global_to_lds_gbs"] = n * 8 / t / 1e9
    small = (85 << 20) // 16384 * 16384
    t = seconds(lambda: _lib.call("cc_calibration_lds_read", st.cuda_stream, small, src.data_ptr(), sink.data_ptr(), 1024), 20)
    out["cache_resident_read_85MB_gbs"] = small / t / 1e9
    for kind, name, iters, flop in ((0, "mfma_f16_tflops", 20000, 16384), (1, "mfma_f64_tflops", 5000, 2048),
                                    (2, "mfma_f16_k16_tflops", 20000, 8192)):
        t = seconds(lambda: _lib.call("cc_calibration_mfma", st.cuda_stream, kind, iters, 1024, sink.data_ptr()), 3)
        out[name] = 1024 * 4 * iters * 8 * flop / t / 1e12
    out["published"] = {"hbm_gbs": PEAK_HBM_GBS, "fp16_dense_tflops": 2500.0, "fp64_dense_tflops": PEAK_FP64_TFLOPS}
    del src, dst, sink
    return out


def _apply_rate(apply, x, reps):
    import torch
    from carcassonne_amd.device import synchronize
    y = torch.empty_like(x)
    for _ in range(3):
        apply(x, y)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        apply(x, y)
    synchronize()
    return (time.perf_counter() - t0) / reps


def isolated_matvecs(op, x0, reps=50):
    """Back-to-back applications of the stage-3 H operator alone (kernel + partial-sum reduction, inputs
    resident): the FP64 plan, the FP32 plan and the FP16 plan, in matvecs/s."""
    from carcassonne_amd import _lib, pipeline
    from carcassonne_amd.device import ptr, stream
    from carcassonne_amd.precision import LowPrecisionOperator
    x = (x0._t if hasattr(x0, "_t") else x0).reshape(-1).contiguous()
    out = {}
    for name in ("f64", "f32"):
        h = op.plan_handle(name)
        out[name] = 1.0 / _apply_rate(lambda a, b: _lib.call("cc_stage3_plan_apply", stream(), h, ptr(a), ptr(b)), x, reps)
    lp = LowPrecisionOperator(op, "f16")
    out["f16"] = 1.0 / _apply_rate(lp.apply_flat, x, reps)
    return {"matvecs_per_s": out, "flops_per_matvec": pipeline.flops_per_matvec(op),
            "note": "one application = the fused kernel + the fixed-order FP64 sum of its workgroup partials"}


def plain_gpu_d4_chi32(seed=0):
    """The plain GPU workload: TFIM with h drawn from U[2.5, 3.5] (seeded), D=4, chi=32: isolated H matvecs/s
    (FP64 and FP32 operands) and the whitened Lanczos to a 1e-10 Ritz residual in both, each from the same
    seeded start (operators and normalization factor built once, outside the timing)."""
    import numpy as np
    import torch
    from carcassonne_amd import pipeline, synthetic as syn
    from carcassonne_amd.device import synchronize
    from carcassonne_amd.solver import GroundStateSolver
    h = float(np.random.default_rng(seed).uniform(2.5, 3.5))
    c, s_, st = syn.make_environment(4, 32, model="ising", seed=seed)
    corners, sides, opc = pipeline.device_environment(c, s_, syn.operator_center("ising", h=h))
    H, N = pipeline.multipliers(corners, sides, opc, graph=False)
    x0 = torch.from_numpy(st.ravel()).cuda()
    res = {"h": h, "isolated": isolated_matvecs(H.operator, x0, reps=100)}
    factor = None
    for prec in ("f64", "f32"):
        gs = GroundStateSolver(H.operator, N.operator, factor=factor, precision=prec)
        factor = gs.factor
        gs.solve(x0, tol_residual=1e-10)
        synchronize()
        t0 = time.perf_counter()
        x, theta, resid, iters = gs.solve(x0, tol_residual=1e-10)
        synchronize()
        res["lanczos_" + prec] = {"ms": (time.perf_counter() - t0) * 1e3, "iterations": iters, "ritz_value": theta,
                                  "fp64_rayleigh_quotient": gs.rayleigh_quotient(x)}
    e64 = res["lanczos_f64"]["fp64_rayleigh_quotient"]
    res["lanczos_f32"]["rel_diff_vs_f64"] = {k: abs(res["lanczos_f32"][k] - e64) / abs(e64)
                                             for k in ("ritz_value", "fp64_rayleigh_quotient")}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", choices=["both", "mixed", "fp64", "absorb"], default="both",
                    help="absorb: time the fused-bond sweep step (System.contractAndCompressTowards) instead")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first launch a plain copy kernel (unary_kernel, scale by 1) over 512 MiB three times, a read "
                         "of the same 512 MiB through global->LDS loads (cc_calibration_lds_read) three times and a "
                         "matrix-core-bound loop (cc_calibration_mfma) three times: their counters calibrate the HBM "
                         "traffic and the MFMA busy fraction of the solve's kernels")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve is a single-GPU kernel path "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import torch
    from carcassonne_amd import _lib, pipeline, synthetic as syn
    from carcassonne_amd.data import HipData
    from carcassonne_amd.device import synchronize
    from carcassonne_amd.sparse import Identity, Complete, OneSiteOperator, TwoSiteOperator, CENTER

    def tag(k):
        return {"I": lambda: Identity(), "C": lambda: Complete(), "O": lambda: OneSiteOperator(None),
                "T": lambda: TwoSiteOperator(k[1], CENTER),
                "TT": lambda: TwoSiteOperator(k[1], k[2], 0)}[k[0]]()

    c, s, st = syn.make_environment(WORKLOAD["D"], WORKLOAD["chi"], model=WORKLOAD["model"], seed=WORKLOAD["seed"])
    opc = syn.operator_center(WORKLOAD["model"], h=WORKLOAD["h"])
    corners = [{tag(k): HipData.fromArray(v) for k, v in x.items()} for x in c]
    sides = [{tag(k): HipData.fromArray(v) for k, v in x.items()} for x in s]
    op_center = {tag(k): HipData.fromArray(v) for k, v in opc.items()}
    x0 = HipData.fromArray(st)
    synchronize()
    e_ref = golden_energy()
    if args.copy_calibration:
        src = torch.ones(COPY_CALIBRATION_DOUBLES, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            _lib.call("cc_unary_f64", torch.cuda.current_stream().cuda_stream, 3, src.numel(), src.data_ptr(), 1.0,
                      dst.data_ptr())
        sink = torch.empty(1024, dtype=torch.float32, device="cuda")
        for _ in range(3):   # the same 512 MiB read through direct global->LDS loads
            _lib.call("cc_calibration_lds_read", torch.cuda.current_stream().cuda_stream, src.numel() * 8,
                      src.data_ptr(), sink.data_ptr(), 1024)
        sink_m = torch.empty(1024 * 256, dtype=torch.float32, device="cuda")
        for _ in range(3):   # a matrix-core-bound loop: what the MFMA busy counters read at the full rate
            _lib.call("cc_calibration_mfma", torch.cuda.current_stream().cuda_stream, 0, 2000, 1024, sink_m.data_ptr())
        synchronize()
        del src, dst, sink, sink_m

    if args.form == "absorb":
        from carcassonne_amd.system import System
        system = System(corners, sides, x0, op_center)
        directions = iter(range(10 ** 9))
        step = lambda: system.contractAndCompressTowards(next(directions) % 4, WORKLOAD["chi"])
        for _ in range(args.warmup):
            step()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        e, n = system.computeExpectationAndNormalization()
        print(json.dumps({"metric": "fused-bond sweep steps/s (absorb + compress) at D=6,chi=64", "value": 1e3 / ms,
                          "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": ms, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic (seeded environment, carcassonne_amd.synthetic seed 0)",
                          "config": {"workload": "contractAndCompressTowards on the 2D TFIM environment, D=6, "
                                                 "chi=64, bonds cut back to 64", "D": 6, "chi": 64},
                          "energy_after": e, "normalization_after": n}), flush=True)
        return

    def run(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        synchronize()
        # HIP events on every TIMING_EVERY-th stage-3 launch of each kind (launches of one kind do the same work)
        _lib.call("cc_profile_begin_sampled", TIMING_EVERY)
        t0 = time.perf_counter()
        results = [fn() for _ in range(steps)]
        synchronize()
        t1 = time.perf_counter()
        kms, launches = _lib.f64_array([0, 0]), _lib.i64_array([0, 0])
        _lib.call("cc_profile_end", 2, kms, launches)
        return (t1 - t0) * 1e3 / steps, results[-1], list(kms), list(launches)

    out = {}
    if args.form in ("both", "fp64"):
        ms, r, kms, nl = run(lambda: pipeline.solve(corners, sides, op_center, x0), args.steps, args.warmup)
        r_fp64 = r
        op = r["operator"]
        flops = pipeline.flops_per_matvec(op)
        kernel_ms = kms[0] / max(1, nl[0])
        out["fp64"] = dict(ms=ms, energy=r["energy"], err=abs(r["energy"] - e_ref) / abs(e_ref),
                           iterations=r["iterations"], kernel_ms=kernel_ms, launches=nl[0], flops=flops,
                           tflops=flops / (kernel_ms * 1e-3) / 1e12, terms=len(op._A))
    if args.form in ("both", "mixed"):
        ms, r, kms, nl = run(lambda: pipeline.solve_mixed(corners, sides, op_center, x0, dtype="f16"),
                             args.steps, args.warmup)
        lp = r["lp_operator"]
        kernel_ms = kms[1] / max(1, nl[1])
        lp_bytes = lp.bytes_per_matvec
        out["mixed"] = dict(ms=ms, energy=r["energy"], err=abs(r["energy"] - e_ref) / abs(e_ref),
                            iterations=r["iterations"], info=r["info"], kernel_ms=kernel_ms, launches=nl[1],
                            fp64_kernel_ms=kms[0] / max(1, nl[0]), fp64_launches=nl[0], bytes=lp_bytes,
                            gbs=lp_bytes / (kernel_ms * 1e-3) / 1e9, flops=pipeline.flops_per_matvec(r["operator"]),
                            kernel="lp_kernel<36,2,{},{}-split>".format(lp.dtype, lp.split))

    if "fp64" in out:
        out["isolated"] = isolated_matvecs(r_fp64["operator"], x0)
        out["plain_gpu_d4"] = plain_gpu_d4_chi32()

    head = out.get("mixed") or out["fp64"]
    form = "mixed-f16" if "mixed" in out else "fp64-lanczos"
    line = {
        "metric": "ground-state solves/sec at D=6,chi=64 to 1e-10 rel. energy vs plain FP64 GPU",
        "value": 1e3 / head["ms"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16+f64" if form == "mixed-f16" else "f64",
        "data": "synthetic (seeded environment, carcassonne_amd.synthetic seed 0)",
        "config": {"workload": "2D TFIM h=3.0, D=6, chi=64, d=2, w=3; one solve = stage 1/2 + normalization "
                               "factor + eigensolve to 1e-10 rel. energy + FP64 Rayleigh quotient",
                   "form": form, "D": 6, "chi": 64},
        "energy": head["energy"], "energy_rel_err_vs_oracle": head["err"],
        "learned_warm_start": "not used by the headline form: at the learned-start workload (D=4, chi=32) the trained "
                              "network's start takes 50 Lanczos iterations against 46 from the previous-step start "
                              "and 67 from a random one (profiles/warm_start_D4_chi32_study.md); both forms here "
                              "start from the same seeded state",
    }
    if "fp64" in out:
        f = out["fp64"]
        line.update(plain_fp64_solves_per_s=1e3 / f["ms"], plain_fp64_ms_per_solve=f["ms"],
                    plain_fp64_energy_rel_err=f["err"], plain_fp64_lanczos_iterations=f["iterations"],
                    plain_fp64_matvec_kernel_ms=f["kernel_ms"], plain_fp64_matvec_tflops=f["tflops"],
                    plain_fp64_matvec_frac_of_fp64_peak=f["tflops"] / PEAK_FP64_TFLOPS)
    if "mixed" in out:
        m = out["mixed"]
        line.update(mixed_lowprec_iterations=m["iterations"], mixed_info=m["info"],
                    lp_matvec_kernel_ms=m["kernel_ms"], lp_matvec_launches_timed=m["launches"],
                    lp_matvec_timed_every=TIMING_EVERY,
                    lp_matvec_effective_tflops=m["flops"] / (m["kernel_ms"] * 1e-3) / 1e12)
        if "fp64" in out:
            line["speedup_vs_plain_fp64"] = out["fp64"]["ms"] / m["ms"]
        line["roofline"] = {"bound": "hbm", "achieved": m["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": m["gbs"] / PEAK_HBM_GBS,
                            "traffic": measured_traffic(m["kernel"] + " D=6 chi=64"),
                            "kernel": m["kernel"] + " (stage-3 matvec, FP16 slabs)",
                            "algorithmic_bytes_per_launch": m["bytes"]}
    else:
        f = out["fp64"]
        line["roofline"] = {"bound": "mfma", "achieved": f["tflops"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                            "frac": f["tflops"] / PEAK_FP64_TFLOPS, "traffic": None,
                            "kernel": "fused_f64_kernel<36,2> (stage-3 matvec)",
                            "algorithmic_flops_per_launch": f["flops"]}
    if "isolated" in out:
        line["isolated_matvecs_d6_chi64"] = out["isolated"]
        line["plain_gpu_d4_chi32"] = out["plain_gpu_d4"]
    line["measured_peaks"] = measured_peaks()
    if not args.no_cpu_baseline:
        its = out["fp64"]["iterations"] if "fp64" in out else 80
        line["cpu_baseline"] = cpu_baseline(its)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
