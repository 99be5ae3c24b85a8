#!/usr/bin/env python3
"""Time-to-T of TimeEvolver's Krylov evolution on one MI355X, with the SpMV roofline.

Workload (default): ring Bose-Hubbard, L = N = 14 modes/particles, D = 20,058,300
occupation states, nnz = 311,275,100 complex128 CSR entries, seeded couplings
(eps_k ~ U[0,1], J_k ~ U[-1,1], U = 0.5), initial Fock state |14,0,...,0>, evolved to
T = 10 with Krylov dimension m = 40, tol = 1e-8, sampling step 0.5, observables n_k for all
14 modes. Basis, Hamiltonian and observables are built on the device (timed separately);
a "step" is one complete timeEvolve() from psi0 to T with all inputs resident in HBM.

Prints one JSON line. `value` is the time-to-T in seconds (lower is better). `roofline` is
the fused Lanczos SpMV kernel: algorithmic bytes per launch = the matrix bytes its storage
format reads once (sliced ELL: 4 per stored slot, padding included, + 4 per slot of a slice
for its base column + slice offsets + 16 D diagonal; value-dictionary rows: 6 per off-diagonal
entry + row pointers + 16 D; plain CSR: 20 nnz + 4 (D+1)) + 16 D (x = v_j, read once, also for
<v_j, w>) + 16 D (w written) + 16 D (v_{j-1}), divided by its mean duration from HIP events
recorded around every apply launch on the library's stream during the timed runs (the other
kernels are not timed there: `kernel_ms_per_run` comes from one more run with every kind timed,
outside the timed region).
`csr_equivalent_GBps` prices the same launches at the reference's CSR bytes
(20 nnz + 4 (D+1) + 48 D); `spmv_unit` compares the launch time with the time the 60 %-of-peak
bar allows for one CSR SpMV (20 nnz + 4 (D+1) + 32 D bytes at 4.8 TB/s). `cpu_baseline` times
the oracle's CPU restatement of the reference evolver on this host, its SpMV and BLAS on MKL as
the reference's build runs them (OpenMP loops when libmkl_rt is absent), on a bounded sample (the
first two Krylov steps of the same workload) and extrapolates to the full time-to-T.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "time-to-T + SpMV HBM GB/s at dim 2e7, tol 1e-8; batched/adjoint vs plain"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed time-to-T runs")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=14, help="L = N of the ring (14 = the 2e7 workload)")
    ap.add_argument("--T", type=float, default=10.0)
    ap.add_argument("--sampling-step", type=float, default=0.5)
    ap.add_argument("--m", type=int, default=40)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2, help="Krylov steps the CPU baseline runs")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batched-only", action="store_true", help="only the batched measurements")
    ap.add_argument("--shared-count", type=int, default=64, help="initial states of the shared-matrix batch")
    ap.add_argument("--shared-size", type=int, default=10, help="L = N of the shared-matrix batch")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient measurement")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint-gradient measurement")
    ap.add_argument("--adjoint-size", type=int, default=12)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from timeevolver_amd import capi
    if args.batched_only:
        ctx = capi.Context(0)
        print(json.dumps({"batched": bench_batched(np, capi, ctx, args),
                          "batched_shared": bench_batched_shared(np, capi, ctx, args)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": bench_adjoint(np, capi, capi.Context(0), args)}), flush=True)
        return
    from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
    from timeevolver_amd.workloads import (fock_state, ring_bose_hubbard_terms, ring_couplings, ring_dimension,
                                           ring_nnz)

    L = N = args.size
    D = ring_dimension(L, N)
    eps, J = ring_couplings(L, args.seed)
    terms = as_tuples(ring_bose_hubbard_terms(L, eps, J, 0.5))
    obs_terms = [as_tuples([createNumberOperator(k, 1)]) for k in range(L)]

    ctx = capi.Context(0)
    log("building basis/H/observables on the device: L=N=%d D=%d" % (L, D))
    t0 = time.perf_counter()
    basis = capi.Basis.basis(ctx, N, L)
    H = capi.operator_build(ctx, basis, terms)
    obs = [capi.operator_build(ctx, basis, t) for t in obs_terms]
    ctx.synchronize()
    build_s = time.perf_counter() - t0
    assert basis.D == D and H.nnz == ring_nnz(L, N), (basis.D, H.nnz)
    log("built in %.2f s: nnz=%d" % (build_s, H.nnz))

    psi0 = fock_state(D)
    ev = capi.Evolver(ctx, H, psi0, args.T, args.sampling_step, obs, tol=args.tol, m=args.m)
    del psi0
    ev.set_profiling(["apply"])                     # the roofline's kernel, inside the timed runs
    for w in range(args.warmup):
        t = time.perf_counter()
        r = ev.run(want_state=False)
        log("warmup %d: %.3f s, n_steps=%d err=%.3e status=%d" % (w, time.perf_counter() - t, r["n_steps"], r["err"],
                                                                  r["statusCode"]))
    times, kts = [], []
    for k in range(args.steps):
        ctx.synchronize()
        t = time.perf_counter()
        r = ev.run(want_state=False)
        ctx.synchronize()
        times.append(time.perf_counter() - t)
        kts.append(ev.kernel_times())
        log("step %d: %.3f s, n_steps=%d err=%.3e status=%d" % (k, times[-1], r["n_steps"], r["err"], r["statusCode"]))
    time_to_T = float(np.median(times))              # the median of the timed runs
    # every kernel kind timed in one more run, outside the timed region (its events cost time)
    ev.set_profiling(True)
    ev.run(want_state=False)
    kt_all = ev.kernel_times()
    ev.set_profiling(False)

    # roofline of the fused Lanczos SpMV
    apply_ms = sum(kt["apply"][0] for kt in kts)
    apply_n = sum(kt["apply"][1] for kt in kts)
    t_apply = apply_ms / 1e3 / max(apply_n, 1)
    fmt, n_values, b_matrix = H.storage()
    b_spmv = b_matrix + 32 * D                      # y = A x: matrix, x once, y written
    b_apply = b_matrix + 48 * D                     # + v_{j-1}
    b_csr = 20 * H.nnz + 4 * (D + 1) + 48 * D
    b_unit = 20 * H.nnz + 4 * (D + 1) + 32 * D      # one CSR SpMV, y = A x
    achieved = b_apply / t_apply / 1e9
    kname = {0: "te::lanczos_apply_kernel", 1: "te::lanczos_apply_stream_kernel",
             2: "te::lanczos_apply_vd_kernel", 3: "te::lanczos_apply_vd_kernel",
             4: "te::lanczos_apply_sell_kernel"}[fmt]
    kernel_ms = {k: round(kt_all[k][0], 3) for k in kt_all}
    kernel_launches = {k: kt_all[k][1] for k in kt_all}
    # whole-run figure: algorithmic bytes of every apply, projection (48 D; 32 D for the last
    # of a step, which stores nothing) and combine pass (the basis columns read, currentVec and
    # the sampled state written, the observables' diagonals) over the wall time
    n_steps_run = r["n_steps"]
    diag_row = ev.diag_bytes_per_row()              # 14 one-byte observable indices (16 B a row) or 8 L
    b_run = (kernel_launches.get("apply", 0) * b_apply
             + kernel_launches.get("project", 0) * 48 * D - n_steps_run * 16 * D
             + kernel_launches.get("combine", 0) * (16 * args.m * D + 32 * D + diag_row * D))

    line = {
        "metric": METRIC,
        "value": round(time_to_T, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(time_to_T * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded ring Bose-Hubbard couplings, Fock initial state)",
        "config": {"workload": "ring Bose-Hubbard L=N=%d, D=%d, nnz=%d, T=%g, m=%d, tol=%g, sampling step %g, "
                               "observables n_k x %d" % (L, D, H.nnz, args.T, args.m, args.tol, args.sampling_step, L),
                   "dim": D, "nnz": H.nnz, "T": args.T, "m": args.m, "tol": args.tol,
                   "sampling_step": args.sampling_step},
        "run_times_s": [round(t, 4) for t in times],
        "n_krylov_steps": r["n_steps"],
        "err": r["err"],
        "status_code": r["statusCode"],
        "build_s": round(build_s, 3),
        "spmv_GBps": round(b_spmv / t_apply / 1e9, 1),
        "csr_equivalent_GBps": round(b_csr / t_apply / 1e9, 1),
        "spmv_unit": {"bytes": b_unit, "ms_allowed_at_60pct_peak": round(b_unit / (0.6 * PEAK_HBM_GBS * 1e9) * 1e3, 4),
                      "apply_ms_per_launch": round(t_apply * 1e3, 4),
                      "within": bool(t_apply <= b_unit / (0.6 * PEAK_HBM_GBS * 1e9))},
        "storage": {"format": ["csr-lanes", "csr-rowblocks", "value-dictionary",
                               "value-dictionary, locality order",
                               "value-dictionary sliced ELL, locality order"][fmt], "n_values": n_values,
                    "matrix_bytes": b_matrix},
        "kernel_ms_per_run": kernel_ms,
        "kernel_launches_per_run": kernel_launches,
        "run_roofline": {"bytes": int(b_run), "achieved_GBps": round(b_run / time_to_T / 1e9, 1),
                         "frac": round(b_run / time_to_T / 1e9 / PEAK_HBM_GBS, 4),
                         "basis": "algorithmic bytes of a run's applies, projections and combine passes "
                                  "over time-to-T",
                         "observable_bytes_per_row": diag_row},
        "roofline": {"kernel": kname + " (fused SpMV + axpy + dot)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(kname, D, b_apply),
                     "bytes_per_launch": b_apply, "mean_launch_us": round(t_apply * 1e6, 2),
                     "launches_timed": apply_n},
    }

    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(np, capi, H, obs, D, args, r["n_steps"])
    del ev, H, obs, basis
    if not args.no_batched:
        line["batched"] = bench_batched(np, capi, ctx, args)
        line["batched_shared"] = bench_batched_shared(np, capi, ctx, args)
    if not args.no_adjoint:
        line["adjoint"] = bench_adjoint(np, capi, ctx, args)
    print(json.dumps(line), flush=True)


def bench_batched(np, capi, ctx, args):
    """Batched workload: `batch_count` ring Bose-Hubbard instances, instance i from a SplitMix64
    stream seeded with seed + i: L uniform in {6..10} (D 462 to 92,378), N = L, its own eps_k,
    J_k and U in [0.2, 1.0], a random complex initial state; T = 5, m = 30, tol = 1e-8, sampling
    step 0.25, observable n_0. One batched run (te_batch_run) against the same instances run one
    after the other through the plain GPU evolver (te_evolver_run, evolvers created beforehand).
    Metric: aggregate Krylov steps per second."""
    from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
    from timeevolver_amd.workloads import batch_instance, random_state, ring_bose_hubbard_terms

    T, m, tol, step = 5.0, 30, 1e-8, 0.25
    t0 = time.perf_counter()
    Hs, psis, obs = [], [], []
    for i in range(args.batch_count):
        L, N, eps, J, U, sseed = batch_instance(args.seed, i)
        b = capi.Basis.basis(ctx, N, L)
        Hs.append(capi.operator_build(ctx, b, as_tuples(ring_bose_hubbard_terms(L, eps, J, U))))
        obs.append([capi.operator_build(ctx, b, as_tuples([createNumberOperator(0, 1)]))])
        psis.append(random_state(b.D, sseed))
    ctx.synchronize()
    build_s = time.perf_counter() - t0
    dims = [H.n_rows for H in Hs]
    mat_bytes = [H.storage()[2] for H in Hs]          # matrix stream of one product, per instance
    log("batched: %d instances built in %.2f s, sum D = %d, sum nnz = %d" %
        (len(Hs), build_s, sum(dims), sum(H.nnz for H in Hs)))
    be = capi.BatchEvolver(ctx, Hs, psis, T, step, obs, tol=tol, m=m)
    rb = be.run(want_states=False)                       # warmup (tunes the shape, captures the round graph)
    shape = be.shape()
    log("batched: launch shape %s" % shape)
    times = []
    for _ in range(max(args.steps, 1)):
        ctx.synchronize()
        t = time.perf_counter()
        rb = be.run(want_states=False)
        ctx.synchronize()
        times.append(time.perf_counter() - t)
    t_batch = float(np.median(times))
    steps_batch = sum(r["n_steps"] for r in rb)
    rounds = be.rounds()
    log("batched: %.3f s, %d Krylov steps in %d rounds" % (t_batch, steps_batch, rounds))
    # every launch kind timed with HIP events in one more run, outside the timed region (the
    # rounds are then launched kernel by kernel, not replayed from the graph)
    be.set_profiling(True)
    ctx.synchronize()
    t = time.perf_counter()
    be.run(want_states=False)
    ctx.synchronize()
    t_prof = time.perf_counter() - t
    kt = be.kernel_times()
    per_launch = be.kernel_launch_times()
    rounds_prof = be.rounds()
    be.set_profiling(False)
    be.close()
    # per round: the m Lanczos iterations' applies and projections, the controller, the combine
    # that follows it (the first combine takes the t = 0 sample); live instances and rows from
    # the instances' Krylov step counts (instance i steps in rounds 0 .. n_steps_i - 1)
    per_it = len(per_launch["apply"]) // max(rounds_prof, 1)
    per_round = {
        "live": [sum(1 for r in rb if r["n_steps"] > q) for q in range(rounds_prof)],
        "live_rows": [sum(d for r, d in zip(rb, dims) if r["n_steps"] > q) for q in range(rounds_prof)],
        "apply_ms": [round(float(per_launch["apply"][q * per_it:(q + 1) * per_it].sum()), 3)
                     for q in range(rounds_prof)],
        "project_ms": [round(float(per_launch["project"][q * per_it:(q + 1) * per_it].sum()), 3)
                       for q in range(rounds_prof)],
        "control_ms": [round(float(x), 3) for x in per_launch["control"]],
        "combine_ms": [round(float(x), 3) for x in per_launch["combine"][1:]],
        "t0_combine_ms": round(float(per_launch["combine"][0]), 3) if len(per_launch["combine"]) else None,
    }

    evs = [capi.Evolver(ctx, H, p, T, step, o, tol=tol, m=m) for H, p, o in zip(Hs, psis, obs)]
    evs[0].run(want_state=False)                         # warmup
    ctx.synchronize()
    t = time.perf_counter()
    rp = [ev.run(want_state=False) for ev in evs]
    ctx.synchronize()
    t_plain = time.perf_counter() - t
    steps_plain = sum(r["n_steps"] for r in rp)
    log("plain, one instance after another: %.3f s, %d Krylov steps" % (t_plain, steps_plain))
    same = all(a["n_steps"] == b["n_steps"] and a["statusCode"] == b["statusCode"] for a, b in zip(rb, rp))
    # algorithmic bytes of a run's Lanczos passes: per instance and Krylov step, krylovDim applies
    # (matrix stream + 48 D) and projections (48 D), as in the single-instance kernels
    b_apply = sum(r["n_steps"] * r["krylovDim"] * (mb + 48 * d) for r, mb, d in zip(rb, mat_bytes, dims))
    b_proj = sum(r["n_steps"] * r["krylovDim"] * 48 * d for r, d in zip(rb, dims))
    dobs = max(float(np.abs(a["observables"] - b["observables"]).max()) for a, b in zip(rb, rp))
    for ev in evs:
        ev.close()
    return {"metric": "aggregate Krylov steps per second", "unit": "steps/s",
            "workload": "%d ring Bose-Hubbard instances, L=N uniform in {6..10}, sum D=%d, T=%g, m=%d, tol=%g, "
                        "sampling step %g, observable n_0" % (len(dims), sum(dims), T, m, tol, step),
            "value": round(steps_batch / t_batch, 1), "plain_value": round(steps_plain / t_plain, 1),
            "speedup_vs_plain": round((steps_batch / t_batch) / (steps_plain / t_plain), 2),
            "spmv_per_s": round(sum(r["n_steps"] * r["krylovDim"] for r in rb) / t_batch, 1),
            "spmv_per_s_note": "Krylov steps x Krylov dimension per instance (at most: a lucky breakdown "
                               "ends a step's Lanczos early)",
            "batched_s": round(t_batch, 4), "plain_s": round(t_plain, 4), "krylov_steps": steps_batch,
            "rounds": rounds, "same_steps_and_status_as_plain": same, "max_obs_diff_vs_plain": dobs,
            "build_s": round(build_s, 2),
            "launch_shape": shape,
            "apply_GBps": round(b_apply / (kt["apply"][0] / 1e3) / 1e9, 1),
            "project_GBps": round(b_proj / (kt["project"][0] / 1e3) / 1e9, 1),
            "pass_bytes_note": "algorithmic bytes of the run's applies (matrix stream + 48 D per instance and "
                               "iteration) and projections (48 D) over their event-timed totals (finish launches "
                               "included); an instance that broke down moved fewer",
            "kernel_ms_per_run": {k: round(v[0], 3) for k, v in kt.items()},
            "per_round": per_round,
            "kernel_launches_per_run": {k: v[1] for k, v in kt.items()},
            "profiled_run_s": round(t_prof, 4),
            "kernel_note": "one more run with HIP events around every launch kind (apply and projection with "
                           "their finish launches, controller, combine with observable sums and live count), "
                           "launched kernel by kernel; outside the timed runs"}


def bench_batched_shared(np, capi, ctx, args):
    """Many initial states of one Hamiltonian: `shared_count` seeded random complex states of the
    ring Bose-Hubbard L = N = `shared_size` (couplings from the bench seed, U = 0.5), T = 5,
    m = 30, tol = 1e-8, sampling step 0.25, observable n_0, as one batch whose instances share
    the matrix (served four at a time by one pass over its words), as the same batch with every
    instance on its own (TE_BATCH_SHARED=0 at creation) and one after the other through the plain
    evolver. Metric: aggregate Krylov steps per second."""
    from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
    from timeevolver_amd.workloads import random_state, ring_bose_hubbard_terms, ring_couplings

    T, m, tol, step = 5.0, 30, 1e-8, 0.25
    L = N = args.shared_size
    eps, J = ring_couplings(L, args.seed)
    b = capi.Basis.basis(ctx, N, L)
    H = capi.operator_build(ctx, b, as_tuples(ring_bose_hubbard_terms(L, eps, J, 0.5)))
    obs = [capi.operator_build(ctx, b, as_tuples([createNumberOperator(0, 1)]))]
    psis = [random_state(b.D, 7000 + q) for q in range(args.shared_count)]
    ctx.synchronize()
    k = len(psis)

    def timed_batch(shared):
        old = os.environ.get("TE_BATCH_SHARED")
        os.environ["TE_BATCH_SHARED"] = "1" if shared else "0"
        try:
            be = capi.BatchEvolver(ctx, [H] * k, psis, T, step, [obs] * k, tol=tol, m=m)
        finally:
            if old is None:
                del os.environ["TE_BATCH_SHARED"]
            else:
                os.environ["TE_BATCH_SHARED"] = old
        rb = be.run(want_states=False)                   # warmup (tuner, graph capture)
        times = []
        for _ in range(max(args.steps, 1)):
            ctx.synchronize()
            t = time.perf_counter()
            rb = be.run(want_states=False)
            ctx.synchronize()
            times.append(time.perf_counter() - t)
        be.set_profiling(True)
        be.run(want_states=False)
        kt = be.kernel_times()
        be.close()
        return float(np.median(times)), rb, kt

    t_sh, r_sh, kt_sh = timed_batch(True)
    t_own, r_own, kt_own = timed_batch(False)
    evs = [capi.Evolver(ctx, H, p, T, step, obs, tol=tol, m=m) for p in psis]
    evs[0].run(want_state=False)
    ctx.synchronize()
    t = time.perf_counter()
    rp = [ev.run(want_state=False) for ev in evs]
    ctx.synchronize()
    t_plain = time.perf_counter() - t
    for ev in evs:
        ev.close()
    steps = sum(r["n_steps"] for r in r_sh)
    same = all(a["n_steps"] == c["n_steps"] and np.array_equal(a["observables"], c["observables"])
               for a, c in zip(r_sh, r_own))
    log("batched, shared matrix: %.4f s (every instance on its own: %.4f s, plain: %.3f s), apply %.1f ms "
        "against %.1f ms" % (t_sh, t_own, t_plain, kt_sh["apply"][0], kt_own["apply"][0]))
    return {"workload": "%d random initial states of one ring Bose-Hubbard L=N=%d (D=%d), T=%g, m=%d, tol=%g, "
                        "sampling step %g, observable n_0" % (k, L, b.D, T, m, tol, step),
            "metric": "aggregate Krylov steps per second", "unit": "steps/s",
            "value": round(steps / t_sh, 1), "krylov_steps": steps,
            "shared_s": round(t_sh, 4), "own_s": round(t_own, 4), "plain_s": round(t_plain, 4),
            "speedup_vs_own": round(t_own / t_sh, 3), "speedup_vs_plain": round(t_plain / t_sh, 2),
            "apply_ms_per_run": {"shared": round(kt_sh["apply"][0], 3), "own": round(kt_own["apply"][0], 3)},
            "bitwise_same_as_own": same}


def bench_adjoint(np, capi, ctx, args):
    """Adjoint workload: ring Bose-Hubbard L = N = adjoint_size (12: D = 1,352,078), seeded eps_k and
    J_k, U = 0.5, Fock |N,0,...,0>, T = 5, m = 40, tol = 1e-8, loss <psi(T)|n_0|psi(T)>, gradient with
    respect to the L eps_k and L J_k. Forward cost = the plain evolver's time-to-T for the same
    run (samples at 0 and T); the adjoint is timed (forward + backward) for checkpoint intervals
    K in {0, 1, 4, 16}; the K = 0 gradient is checked against central differences (h = 1e-6) of
    the plain evolver's loss at tol 1e-12."""
    from timeevolver_amd.adjoint import ring_problem
    from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
    from timeevolver_amd.workloads import fock_state, ring_bose_hubbard_terms, ring_couplings

    L = N = args.adjoint_size
    T, m, tol, U = 5.0, 40, 1e-8, 0.5
    eps, J = ring_couplings(L, args.seed)
    theta = np.array(eps + J)
    t0 = time.perf_counter()
    prob = ring_problem(ctx, L, N, U, T, tol=tol, m=m, checkpoint_interval=0, theta0=theta)
    D = prob.basis.D
    psi0 = fock_state(D)
    n0 = capi.operator_build(ctx, prob.basis, as_tuples([createNumberOperator(0, 1.0)]))
    log("adjoint: L=N=%d D=%d set up in %.2f s" % (L, D, time.perf_counter() - t0))

    def plain_loss(th, tl):
        H = capi.operator_build(ctx, prob.basis, as_tuples(ring_bose_hubbard_terms(L, list(th[:L]), list(th[L:]), U)))
        ev = capi.Evolver(ctx, H, psi0, T, T, [n0], tol=tl, m=m)
        r = ev.run(want_state=False)
        ev.close()
        return r["observables"][0, 1], r

    ev = capi.Evolver(ctx, prob.H, psi0, T, T, [n0], tol=tol, m=m)
    ev.run(want_state=False)
    fwd = []
    for _ in range(3):
        ctx.synchronize()
        t = time.perf_counter()
        r = ev.run(want_state=False)
        ctx.synchronize()
        fwd.append(time.perf_counter() - t)
    t_fwd = float(np.median(fwd))
    loss_plain = r["observables"][0, 1]
    ev.close()
    out = {"workload": "ring Bose-Hubbard L=N=%d, D=%d, T=%g, m=%d, tol=%g, loss <n_0>(T), gradient wrt %d eps_k + %d "
                       "J_k" % (L, D, T, m, tol, L, L),
           "forward_s": round(t_fwd, 4), "forward_steps": r["n_steps"], "K": {}}
    grad0 = None
    for K in (0, 1, 4, 16):
        pk = prob if K == 0 else ring_problem(ctx, L, N, U, T, tol=tol, m=m, checkpoint_interval=K, theta0=theta)
        pk.adj.run(psi0)
        ts = []
        for _ in range(3):
            ctx.synchronize()
            t = time.perf_counter()
            loss, grad, res, info = pk.adj.run(psi0)
            ctx.synchronize()
            ts.append(time.perf_counter() - t)
        tk = float(np.median(ts))
        if K == 0:
            grad0 = grad
            out["loss"] = loss
            out["loss_minus_plain"] = loss - loss_plain
        out["K"][str(K)] = {"total_s": round(tk, 4), "ratio_vs_forward": round(tk / t_fwd, 3),
                            "device_forward_ms": round(info["forward_ms"], 2),
                            "device_backward_ms": round(info["backward_ms"], 2), "nodes": info["nodes"],
                            "node_states": info["node_states"],
                            "max_phase": round(info["max_phase"], 1), "lambda_err": info["backward_err"],
                            "psi_err": info["psi_err"],
                            "grad_rel_diff_vs_K0": float(np.linalg.norm(grad - grad0) / np.linalg.norm(grad0))}
        log("adjoint K=%d: %.3f s (forward %.3f s) ratio %.2f nodes %d, node states %d"
            % (K, tk, t_fwd, tk / t_fwd, info["nodes"], info["node_states"]))
        if K != 0:
            pk.adj.close()
    # central differences of the plain evolver's loss, tight tolerance
    t = time.perf_counter()
    h = 1e-6
    fd = np.zeros_like(theta)
    for p in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (plain_loss(tp, 1e-12)[0] - plain_loss(tm, 1e-12)[0]) / (2 * h)
    out["fd_rel_err"] = float(np.linalg.norm(grad0 - fd) / np.linalg.norm(fd))
    out["fd_max_abs_err"] = float(np.abs(grad0 - fd).max())
    out["grad_norm"] = float(np.linalg.norm(grad0))
    log("adjoint: FD check %.2e relative (%.1f s)" % (out["fd_rel_err"], time.perf_counter() - t))
    out["ratio_vs_forward"] = out["K"]["0"]["ratio_vs_forward"]
    return out


def measured_traffic(kernel, D, algorithmic_bytes):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE scaled by a plain-copy calibration), when
    it was measured on this workload with this storage (same algorithmic bytes per launch);
    None otherwise."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not t or ("D=%d" % D) not in t.get("workload", ""):
        return None
    if t.get("algorithmic_bytes") != algorithmic_bytes:
        return None
    return t["fetch_bytes"] + t["write_bytes"]


def copy_calibration():
    """A plain copy of 1 GiB on the device, repeated 3 times: rocprofv3's FETCH_SIZE/WRITE_SIZE for
    this kernel give the counter-to-bytes factor used for the SpMV traffic figure."""
    import torch
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "repeats": 3}), flush=True)


def host_description():
    """CPU model, logical CPUs (all, and usable by this process) and RAM of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ram_gib = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    ram_gib = round(int(line.split()[1]) / 1048576.0, 1)
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "ram_GiB": ram_gib}


def cpu_baseline(np, capi, H, obs, D, args, gpu_steps):
    """The reference's CPU path on this host's cores: the oracle's restatement of
    krylovTimeEvolver::timeEvolve (core/krylovTimeEvolver.cpp:261-512) with its products and
    level-1/2 kernels on MKL when libmkl_rt is present, as the reference's MKL build runs them
    (mkl_sparse_z_mv on an ILP64 handle from create_coo -> convert_csr -> order -> hints ->
    optimize, core/matrixDataTypes.cpp:169-174, :196-224; cblas axpy/dotc/nrm2/scal/copy/gemv),
    else on the oracle's OpenMP loops. Threads: OMP_NUM_THREADS when set (the host's share of
    this machine), else every CPU this process may use; MKL gets the same count. Timed: the
    first `cpu_steps` Krylov steps of the same workload with their samples (the t = 0 sample
    included), timeEvolve's loop only (the handles' set-up is the constructor's, timed apart);
    extrapolated to the GPU run's Krylov steps by seconds per step."""
    import oracle
    from timeevolver_amd.workloads import fock_state

    host = host_description()
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or host["usable_cpus"]
    oracle.set_threads(threads)
    rc = oracle.mkl_enable(threads=threads)
    if rc == 0:
        spmv = "mkl"
        backend = "MKL mkl_sparse_z_mv + cblas (%s)" % oracle.mkl_version()
    else:
        spmv = "openmp-loop"
        backend = "oracle OpenMP loops (libmkl_rt %s)" % {1: "not found", 2: "lacks an entry point",
                                                          3: "refused the ILP64 interface"}.get(rc, "error %d" % rc)
    log("cpu baseline: %s, %d threads; copying CSR to host" % (backend, threads))
    ptr, col, val = H.to_csr()
    diags = []
    for o in obs:
        p, c, v = o.to_csr()
        d = np.zeros(D)
        d[c] = v.real        # diagonal observables: one entry per row
        diags.append(d)
    try:
        r = oracle.evolve_csr(ptr, col, val, fock_state(D), args.T, args.sampling_step, diags, tol=args.tol,
                              m=args.m, max_steps=args.cpu_steps)
    finally:
        oracle.mkl_disable()
    dt = r["run_s"]
    per_step = dt / max(r["n_steps"], 1)
    log("cpu baseline: %.2f s for %d Krylov step(s), %.2f s set-up" % (dt, r["n_steps"], r["setup_s"]))
    out = {"value": round(per_step * gpu_steps, 2), "unit": "s", "cores": threads, "kind": "port",
           "spmv": spmv, "backend": backend,
           "sample": "first %d Krylov step(s) of the same workload (m=%d Lanczos iterations each, the samples "
                     "they cover of %d observables, t=0 sample included): %.2f s of timeEvolve on %d threads; "
                     "value = seconds per Krylov step x %d Krylov steps of the GPU run (extrapolated)"
                     % (r["n_steps"], args.m, len(diags), dt, threads, gpu_steps),
           "extrapolated": True, "seconds_per_krylov_step": round(per_step, 3), "krylov_steps_timed": r["n_steps"],
           "samples_timed": r["numSamples"], "setup_s": round(r["setup_s"], 2)}
    out.update(host)
    return out


if __name__ == "__main__":
    main()
