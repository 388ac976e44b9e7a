#!/usr/bin/env python
"""Throughput of the MI355X Heston ADI engine (one GPU).

Default workload (the line's ``value``): a single FP64 Heston European call
on a 2048 x 1024 (spot x variance) grid, rho = -0.7, Hundsdorfer-Verwer
theta = 0.5, nt = 2000 steps (S0 = K = 100, r = 0.06, kappa = 1,
theta = v0 = 0.04, sigma = 0.2, spot mesh on [0, 5K], variance mesh on
[0, 5 theta]).  One bench "step" is one full option solve (2000 ADI steps) on
device-resident inputs; ``value`` is options/s over the K timed solves.

The line also carries
* ``roofline``: the dominant kernel's algorithmic bytes per launch over its
  HIP-event-measured average duration (events on the engine's stream, inside
  the timed region), against the 8 TB/s HBM peak; ``traffic`` is read from
  the rocprofv3 counter summary committed under profiles/ when present;
* ``cpu_baseline``: the NumPy/SciPy restatement of the reference CPU engine
  (oracle/), one thread, on a bounded sample of the same workload;
* ``extra``: the 150 x 150 single option (plain loop) and the 4096-option
  batched Craig-Sneyd workload in options/s.
"""

import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("options/sec at ≤1e-4 abs err vs FP64 ADI: surrogate+fallback vs batched "
          "vs plain")
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per grid point for each HV stage (reads + writes of state
# vectors; coefficients and factors are regenerated and count 0).
HV_STAGE_BYTES = [16, 16, 24, 16]
HV_STAGE_NAMES = ["spot_sweep_kernel<SM_FIRST>", "var_sweep_kernel<VM_PLAIN>",
                  "spot_sweep_kernel<SM_HV>", "var_sweep_kernel<VM_PLAIN>"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def config1():
    from fd_adi_amd.workloads import single
    return single(2048, 1024, 2000, -0.7)


def run_single(opt, scheme, theta, K, W, smoothing=0, graph=False):
    """Time K solves of one option (state reset from a device-resident copy).

    Every solve rebuilds the factorisations (configure) and runs all nt
    steps.  With ``graph`` the steps of a solve are one HIP graph replay
    (captured once, untimed); the last timed solve launches its stages one by
    one with a HIP event pair around each (the per-kernel durations)."""
    import numpy as np
    import fd_adi_amd
    hb = fd_adi_amd.HestonBatch([opt.params], [opt.spots], [opt.vars], [opt.dt])
    U0 = opt.payoff()
    hb.set_state([U0])
    init = hb.state_view(0).copy(True)          # payoff kept in HBM
    view = hb.state_view(0)
    n = opt.nt - smoothing
    if graph:
        assert not smoothing
        hb.configure(scheme, theta, 1.0)
        hb.capture(n)

    def solve(timed=False):
        view.copy_from(init)
        stages = None
        if smoothing:
            hb.configure("implicit", 1.0, 0.5)
            hb.step(2 * smoothing)
        hb.configure(scheme, theta, 1.0)
        if timed:
            stages = hb.step_timed(n)
        elif graph:
            hb.replay()
        else:
            hb.step(n)
        return stages

    for w in range(W):
        solve(timed=(w == 0))  # also creates the event pool outside the timed region
    hb.sync()
    t0 = time.perf_counter()
    stages = None
    for k in range(K):
        st = solve(timed=(k == K - 1))
        if st is not None:
            stages = st
    hb.sync()
    wall = time.perf_counter() - t0
    price = hb.get_state()[0][opt.idx]
    return wall, stages, price, hb


def batched(n_opt, seed, K, W):
    import numpy as np
    import fd_adi_amd
    from fd_adi_amd.workloads import batch_draws
    t = time.perf_counter()
    opts = batch_draws(n_opt, seed)
    log("batch: %d options drawn in %.1fs" % (n_opt, time.perf_counter() - t))
    hb = fd_adi_amd.HestonBatch([o.params for o in opts], [o.spots for o in opts],
                                [o.vars for o in opts], [o.dt for o in opts])
    hb.set_state([o.payoff() for o in opts])
    init = hb.state_view().copy(True)
    view = hb.state_view()
    hb.configure("cs", 0.5, 1.0)
    shapes = hb.autotune()
    log("batch: autotuned launch shapes %s" % shapes)
    hb.step(2)
    hb.capture(200)
    for _ in range(W):
        view.copy_from(init)
        hb.replay()
    hb.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        view.copy_from(init)
        hb.replay()
    hb.sync()
    wall = time.perf_counter() - t0
    npts = sum(o.spots.size * o.vars.size for o in opts)
    return n_opt * K / wall, wall / K, npts, shapes


def surrogate_run(n_labels, n_queries, members, steps, seed=17, subsample=1.0):
    """Configs 3-4 at reduced scale: FP64 ADI labels from the batched engine
    (256 x 128, nt = 200, Craig-Sneyd), an ensemble of MLPs trained on the
    GPU, the spread gate calibrated on held-out labels for 1e-4 absolute
    error, then serving of fresh queries (5% outside the training box) with
    the ADI fallback; errors against FP64 ADI on a sample of the queries."""
    import numpy as np
    import fd_adi_amd
    from fd_adi_amd import surrogate as S
    out = {"labels": n_labels, "queries": n_queries, "members": members,
           "train_steps": steps, "batch": 8192, "grid": "256x128 nt=200 cs, spot mesh K x fixed sinh mesh, fixed "
                                     "variance power mesh, bicubic read-out at (S0, v0)"}
    t = time.perf_counter()
    P = S.draw_params(n_labels, seed)
    y = S.adi_prices(P)
    fd_adi_amd._lib.sync()
    t_lab = time.perf_counter() - t
    ok = S.within_call_bounds(P, y)
    out["label_options_per_s"] = round(n_labels / t_lab, 1)
    out["labels_dropped_unstable"] = int((~ok).sum())
    P, y = P[ok], y[ok]
    nval = max(len(y) // 10, 1)
    Pv, yv, Pt, yt = P[:nval], y[:nval], P[nval:], y[nval:]
    t = time.perf_counter()
    feller_min, max_resid = 0.0, 0.05
    out["train_subsample"] = subsample
    out["gate_feller_min"], out["train_max_resid"] = feller_min, max_resid
    ens, hists = S.train_ensemble(Pt, yt, members=members, steps=steps, log_every=steps // 4,
                                  max_resid=max_resid, schedule="cosine", subsample=subsample)
    fd_adi_amd._lib.sync()
    t_train = time.perf_counter() - t
    flop = 4.75e6 * 8192 * steps * members
    out["train_sec"] = round(t_train, 2)
    out["train_TFLOPs"] = round(flop / t_train / 1e12, 1)
    out["train_final_loss"] = [round(h[-1][1], 6) for h in hists]
    out["baseline"] = "semi-analytic Heston price (device), network learns ADI - it"
    base_v = S.cf_price(Pv)
    out["baseline_only_val_rmse"] = float(np.sqrt(np.mean((base_v - yv) ** 2)))
    out["baseline_only_val_max_abs_err"] = float(np.max(np.abs(base_v - yv)))
    mean_v, spread_v = ens.predict(Pv)
    out["val_rmse"] = float(np.sqrt(np.mean((mean_v - yv) ** 2)))
    out["val_max_abs_err"] = float(np.max(np.abs(mean_v - yv)))
    # ungated: how many held-out labels the ensemble mean (and the analytic
    # price alone) meets at each tolerance
    out["val_frac_within"] = {"%g" % a: round(float(np.mean(np.abs(mean_v - yv) <= a)), 4)
                              for a in (1e-4, 1e-3, 1e-2)}
    out["baseline_only_val_frac_within"] = {
        "%g" % a: round(float(np.mean(np.abs(base_v - yv) <= a)), 4) for a in (1e-4, 1e-3, 1e-2)}
    tau, acc = S.calibrate_gate(ens, Pv, yv, 1e-4, feller_min)
    out["gate_tau"] = tau
    out["gate_val_accept"] = round(acc, 4)
    # the same gate calibrated for looser tolerances (not used for serving)
    out["val_accept_at_atol"] = {"%g" % a: round(S.calibrate_gate(ens, Pv, yv, a,
                                                                   feller_min)[1], 4)
                                 for a in (1e-3, 1e-2, 1e-1)}
    out["val_accept_at_atol_feller_ge_1.5"] = {
        "%g" % a: round(S.calibrate_gate(ens, Pv, yv, a, 1.5)[1], 4) for a in (1e-4, 1e-3, 1e-2)}
    # a looser gate (not used for serving): the largest spread threshold at
    # which at most 0.1% of the accepted held-out queries exceed 1e-4
    order = np.argsort(spread_v)
    fails = np.cumsum(np.abs(mean_v - yv)[order] > 1e-4)
    okq = np.where(fails <= 1e-3 * np.arange(1, len(order) + 1))[0]
    out["val_accept_1e-4_at_0.1pct_failures"] = round((okq[-1] + 1) / len(order), 4) \
        if okq.size else 0.0
    reg = S.feller_ratio(Pv) >= feller_min
    out["val_rmse_in_gate_region"] = float(np.sqrt(np.mean((mean_v[reg] - yv[reg]) ** 2)))
    out["baseline_only_val_rmse_in_gate_region"] = float(
        np.sqrt(np.mean((base_v[reg] - yv[reg]) ** 2)))
    pricer = S.SurrogatePricer(ens, tau, feller_min=feller_min)
    Q = S.draw_params(n_queries, seed + 1, outside_frac=0.05)
    t = time.perf_counter()
    price, okq = pricer.price(Q)
    fd_adi_amd._lib.sync()
    t_serve = time.perf_counter() - t
    out["options_per_s"] = round(n_queries / t_serve, 1)
    out["fallback_rate"] = round(1.0 - okq.mean(), 4)
    t = time.perf_counter()
    _ = ens.predict(Q)
    fd_adi_amd._lib.sync()
    out["surrogate_only_options_per_s"] = round(n_queries / (time.perf_counter() - t), 1)
    t = time.perf_counter()
    _ = S.cf_price(Q)
    fd_adi_amd._lib.sync()
    out["semi_analytic_options_per_s"] = round(n_queries / (time.perf_counter() - t), 1)
    rng = np.random.default_rng(seed + 2)
    idx = rng.choice(n_queries, min(1024, n_queries), replace=False)
    truth = S.adi_prices(Q[idx])
    err = np.abs(price[idx] - truth)
    out["err_sample"] = len(idx)
    out["max_abs_err"] = float(np.max(err))
    out["mean_abs_err"] = float(np.mean(err))
    return out


def single_loops(n_op, n_fused, seed=2024):
    """Config-2 options solved one at a time (host payoff in, host grid out):
    the operator-algebra path (per-step BandedOperator calls, the reference's
    GPU loop) and the fused one-option engine; options/s for each."""
    from fd_adi_amd.workloads import batch_draws
    from fd_adi_amd.FiniteDifferenceEngineGPU import HestonFiniteDifferenceEngine
    from fd_adi_amd.Option import HestonOption
    opts = batch_draws(max(n_op, n_fused) + 1, seed)
    out = {}
    for fused, n in ((False, n_op), (True, n_fused)):
        runs = []
        for o in opts[:n + 1]:
            opt = HestonOption(spot=o.S0, strike=o.K, interest_rate=o.r, variance=o.v0,
                               tenor=o.T, mean_reversion=o.kappa, mean_variance=o.theta,
                               vol_of_variance=o.sigma, correlation=o.rho)
            E = HestonFiniteDifferenceEngine(opt, spots=o.spots, vars=o.vars)
            E.fused = fused
            runs.append((E, o))
        E, o = runs[0]
        E.solve_craigsneyd(o.nt, o.dt, o.payoff(), 0.5)  # warm-up (kernel attributes)
        t = time.perf_counter()
        for E, o in runs[1:]:
            E.solve_craigsneyd(o.nt, o.dt, o.payoff(), 0.5)
        out["op_path" if not fused else "fused"] = round(n / (time.perf_counter() - t), 2)
    return out


def cpu_baseline(steps):
    """Oracle (NumPy/SciPy restatement of the reference CPU engine), one
    thread, ``steps`` HV steps of the 2048x1024 option; extrapolated to the
    full 2000-step solve."""
    code = r"""
import os, sys, time, json
sys.path.insert(0, %r)
from fd_adi_amd.workloads import single
from oracle.engine import HestonOracle, HestonParams
o = single(2048, 1024, 2000, -0.7)
E = HestonOracle(HestonParams(spot=o.S0, strike=o.K, r=o.r, v0=o.v0, T=o.T, kappa=o.kappa,
                 theta=o.theta, sigma=o.sigma, rho=o.rho), o.spots, o.vars)
U0 = o.payoff()
E.solve_hundsdorferverwer(1, o.dt, U0)
t = time.perf_counter()
E.solve_hundsdorferverwer(%d, o.dt, U0)
print(json.dumps({"sec_per_step": (time.perf_counter() - t) / %d}))
""" % (ROOT, steps, steps)
    env = dict(os.environ)
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS",
              "NUMEXPR_NUM_THREADS", "BLIS_NUM_THREADS"):
        env[v] = "1"
    env["HIP_VISIBLE_DEVICES"] = ""
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=600)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-2000:])
    sec = json.loads(out.stdout.strip().splitlines()[-1])["sec_per_step"]
    return sec


def probe():
    """Fixed dispatch sequence for counter runs (see --probe)."""
    import fd_adi_amd
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    opt = config1()
    hb = fd_adi_amd.HestonBatch([opt.params], [opt.spots], [opt.vars], [opt.dt])
    hb.set_state([opt.payoff()])
    hb.configure("hv", 0.5, 1.0)
    hb.step(20)
    hb.sync()
    x = SizedArrayPtr().alloc(1 << 25)
    for _ in range(5):
        x.timeseq_scalar_from_host(1.0)
    fd_adi_amd._lib.sync()
    log("probe: 20 HV steps (80 sweep launches) + 5 scale launches on 2^25 doubles done")


def probe_mlp():
    """Fixed MLP dispatch sequence for kernel traces (see --probe-mlp)."""
    import numpy as np
    import fd_adi_amd
    from fd_adi_amd import surrogate as S
    rng = np.random.default_rng(0)
    P = S.draw_params(8192, 1)
    y = rng.normal(10.0, 5.0, 8192)
    m = S.MLP(S.BOX_LO, S.BOX_HI, 10.0, 5.0, seed=0, max_rows=8192)
    m.fit(P, y, 50, batch=8192, lr=1e-3)
    fd_adi_amd._lib.sync()
    log("probe-mlp: 50 training steps at batch 8192 done")


def probe_cf():
    """Fixed semi-analytic pricer dispatches for kernel traces (see --probe-cf)."""
    import fd_adi_amd
    from fd_adi_amd import surrogate as S
    from fd_adi_amd.pricing import heston_call_device
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    import numpy as np
    P = S.draw_params(1 << 20, 3)
    rows = np.ascontiguousarray(np.stack([np.full(len(P), 100.0), 100.0 * P[:, 7], P[:, 5],
                                          P[:, 4], P[:, 6], P[:, 0], P[:, 1], P[:, 2],
                                          P[:, 3]], axis=1))
    d_rows, d_out = SizedArrayPtr(rows), SizedArrayPtr().alloc(len(P))
    for _ in range(5):
        heston_call_device(d_rows, len(P), d_out)
    fd_adi_amd._lib.sync()
    log("probe-cf: 5 launches pricing 2^20 options done")


def read_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_config1.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*%s\s+bytes_per_launch\s*=\s*([0-9.eE+]+)" % re.escape(kernel), line)
        if m:
            return float(m.group(1))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed option solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--surrogate-labels", type=int, default=1 << 18,
                    help="ADI-labelled training draws for the surrogate (0 = skip)")
    ap.add_argument("--surrogate-queries", type=int, default=1 << 15)
    ap.add_argument("--surrogate-steps", type=int, default=8000)
    ap.add_argument("--surrogate-subsample", type=float, default=1.0,
                    help="fraction of the labels each ensemble member trains on (bagging)")
    ap.add_argument("--probe", action="store_true",
                    help="profiling probe for rocprofv3 counter runs: 20 HV steps of the "
                         "2048x1024 option, then 5 launches of the scale kernel on 2^25 "
                         "doubles (reads and writes 256 MiB each: the copy reference); "
                         "prints no JSON line")
    ap.add_argument("--probe-mlp", action="store_true",
                    help="profiling probe: 50 Adam steps of one surrogate network at batch "
                         "8192 on random features and targets; prints no JSON line")
    ap.add_argument("--probe-cf", action="store_true",
                    help="profiling probe: 5 launches of the semi-analytic Heston pricer on "
                         "2^20 parameter rows; prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(options shard across GPUs with no exchange; see the design notes)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import fd_adi_amd  # noqa: F401

    if args.probe:
        probe()
        return
    if args.probe_mlp:
        probe_mlp()
        return
    if args.probe_cf:
        probe_cf()
        return
    opt = config1()
    N = opt.spots.size * opt.vars.size
    log("config1: %dx%d, nt=%d, HV; warmup %d, timed %d" % (opt.spots.size, opt.vars.size,
                                                           opt.nt, args.warmup, args.steps))
    wall, stages, price, _ = run_single(opt, "hv", 0.5, args.steps, args.warmup, graph=True)
    ms_per = 1e3 * wall / args.steps
    value = args.steps / wall
    per_launch_ms = stages / opt.nt
    stage_gbs = [HV_STAGE_BYTES[k] * N / (per_launch_ms[k] * 1e-3) / 1e9 for k in range(4)]
    dom = int(np.argmax(stages))
    achieved = stage_gbs[dom]
    log("config1: %.2f ms/option, price %.12f; stage ms/launch %s; GB/s %s" % (
        ms_per, price, np.round(per_launch_ms, 4).tolist(), np.round(stage_gbs, 1).tolist()))
    algo_step = sum(HV_STAGE_BYTES) * N
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(HV_STAGE_NAMES[dom]),
                "kernel": HV_STAGE_NAMES[dom], "bytes_per_launch": HV_STAGE_BYTES[dom] * N,
                "avg_launch_us": round(per_launch_ms[dom] * 1e3, 3),
                "stage_GBs": [round(x, 1) for x in stage_gbs],
                "step_GBs": round(algo_step / (wall / args.steps / opt.nt) / 1e9, 1)}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: %d HV steps of the same option on one core" % args.cpu_sample_steps)
        sec = cpu_baseline(args.cpu_sample_steps)
        cpu = {"value": 1.0 / (sec * opt.nt), "unit": "options/s", "cores": 1, "kind": "port",
               "sample": "%d of 2000 HV steps of the 2048x1024 option (oracle, NumPy/SciPy "
                         "dia.dot + LAPACK solve_banded), extrapolated" % args.cpu_sample_steps,
               "sec_per_step": sec}

    extra = {}
    if not args.no_extra:
        from fd_adi_amd.workloads import single
        o0 = single(150, 150, 150, 0.0, spot_max=500.0, var_max=0.2)
        w0, _, p0, _ = run_single(o0, "douglas", 0.5, 20, 2, smoothing=2)
        extra["plain_150x150_douglas_smoothed_options_per_s"] = round(20 / w0, 2)
        extra["plain_150x150_price"] = p0
        loops = single_loops(2, 16)
        extra["single_option_loop_cs_options_per_s"] = loops
        rate, sec, npts, shapes = batched(4096, 2024, 2, 1)
        extra["batched_4096_shapes"] = {"spot_threads_cols": list(shapes[0]),
                                        "var_lanes": shapes[1][0]}
        extra["batched_4096_cs_options_per_s"] = round(rate, 1)
        extra["batched_4096_cs_sec"] = round(sec, 4)
        extra["batched_4096_cs_GBs_algorithmic"] = round(72 * npts * 200 / sec / 1e9, 1)
        extra["batched_over_op_path_loop"] = round(rate / loops["op_path"], 1)
        extra["batched_over_fused_loop"] = round(rate / loops["fused"], 1)
        log("extra: %s" % extra)
    if args.surrogate_labels:
        sur = surrogate_run(args.surrogate_labels, args.surrogate_queries, 4,
                            args.surrogate_steps, subsample=args.surrogate_subsample)
        extra["surrogate"] = sur
        log("surrogate: %s" % sur)

    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "options/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "heston_call_2048x1024_hv_nt2000_rho-0.7", "nS": 2048,
                   "nv": 1024, "nt": 2000, "scheme": "hundsdorferverwer", "theta": 0.5,
                   "price": price},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
