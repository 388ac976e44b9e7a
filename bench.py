#!/usr/bin/env python
"""Throughput of the MI355X Heston ADI engine (one GPU).

The metric is options/s at <= 1e-4 absolute error against FP64 ADI, for the
surrogate with its ADI fallback, the batched ADI engine and the plain
single-option engine.  ``value`` is the first of these at the surrogate
serving configuration's full size: 10^6 fresh seed-drawn queries (5% of them
outside the training box) answered by the surrogate trained on 2^20 FP64 ADI
labels (256 x 128 grids, nt = 200, Craig-Sneyd), every query the gate
rejects priced by the batched FP64 ADI engine on the same meshes.  One bench
"step" serves one batch of 50,000 fresh queries, so the default 20 timed
steps are those 10^6 queries (after 5 untimed warm-up steps); ``value`` =
timed queries / wall time (queries on the host, prices returned to it).

The line also carries
* ``roofline``: the dominant kernel of the served workload (a spot sweep of
  the batched fallback engine), its algorithmic bytes per launch over its
  mean duration from HIP events recorded on the engine's stream inside the
  timed region, against the 8 TB/s HBM peak; ``traffic`` from the committed
  rocprofv3 counter summary;
* ``cpu_baseline``: the oracle (the NumPy/SciPy restatement of the reference
  CPU engine) pricing options of the same workload, one option per process
  over the host cores, on a bounded sample;
* ``extra``: the 2048 x 1024 single option (roofline per ADI stage), the
  150 x 150 plain option, the 4096-option batched workload and the
  single-option loops, label throughput, training, the gate and the errors of
  the admitted answers against FP64 ADI, the other CPU baselines.
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
# Algorithmic bytes per grid point of each stage (reads + writes of state
# vectors; coefficients and factors are regenerated and count 0).
HV_STAGE_BYTES = [16, 16, 24, 16]
HV_STAGE_NAMES = ["spot_sweep_kernel<SM_FIRST>", "var_sweep_kernel<VM_PLAIN>",
                  "spot_sweep_kernel<SM_HV>", "var_sweep_kernel<VM_PLAIN>"]
CS_STAGE_BYTES = [16, 16, 24, 16]
CS_STAGE_NAMES = ["spot_sweep_kernel<SM_FIRST>", "var_sweep_kernel<VM_PLAIN>",
                  "spot_sweep_kernel<SM_CS>", "var_sweep_kernel<VM_PLAIN>"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def config1():
    from fd_adi_amd.workloads import single
    return single(2048, 1024, 2000, -0.7)


def run_single(opt, scheme, theta, K, W, smoothing=0, graph=False, tune=False):
    """Time K solves of one option (state reset from a device-resident copy).

    Every solve rebuilds the factorisations (configure) and runs all nt
    steps.  With ``graph`` a solve is one HIP graph replay (captured once,
    untimed: smoothing steps, both refactorisations and the scheme's steps),
    and the per-stage durations come from one eager solve before the capture;
    without, the last timed solve launches its stages one by one with a HIP
    event pair around each.  ``tune``: the launch shapes are autotuned for
    the scheme first (untimed)."""
    import fd_adi_amd
    hb = fd_adi_amd.HestonBatch([opt.params], [opt.spots], [opt.vars], [opt.dt])
    hb.set_state([opt.payoff()])
    if tune:
        hb.configure(scheme, theta, 1.0)
        log("single %dx%d: autotuned launch shapes %s" % (opt.spots.size, opt.vars.size,
                                                         hb.autotune()))
    init = hb.state_view(0).copy(True)          # payoff kept in HBM
    view = hb.state_view(0)
    n = opt.nt - smoothing

    def eager(timed=False):
        view.copy_from(init)
        if smoothing:
            hb.configure("implicit", 1.0, 0.5)
            hb.step(2 * smoothing)
        hb.configure(scheme, theta, 1.0)
        if timed:
            return hb.step_timed(n)
        hb.step(n)
        return None

    graph_stages = None
    if graph:
        graph_stages = eager(timed=True)
        hb.capture_solve(scheme, opt.nt, theta=theta, smoothing_steps=smoothing)

    def solve(timed=False):
        if graph:
            view.copy_from(init)
            hb.replay()
            return None
        return eager(timed)

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
    return wall, (graph_stages if graph else stages), price


def batched(n_opt, seed, K, W):
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
    t = time.perf_counter()
    hb.configure("cs", 0.5, 1.0)
    configure_s = time.perf_counter() - t
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
    return n_opt * K / wall, wall / K, npts, shapes, configure_s


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


def surrogate_run(n_labels, q_per_step, K, W, steps, members, max_fail, check, seed=17,
                  after_labels=None):
    """Configurations 3 and 4 at their sizes: FP64 ADI labels from the
    device labeller, the gated surrogate (ensemble, error model, calibrated
    threshold), then serving: W untimed and K timed steps, each a batch of
    ``q_per_step`` fresh queries (5% outside the training box) answered by
    the surrogate or, where the gate rejects them, by the batched FP64 ADI
    engine.  The admitted answers of the timed steps are checked against
    FP64 ADI on a sample of ``check`` of them (all with check < 0)."""
    import numpy as np
    import fd_adi_amd
    from fd_adi_amd import surrogate as S
    out = {"labels": n_labels, "queries": K * q_per_step, "queries_per_step": q_per_step,
           "members": members, "train_steps": steps, "batch": 8192, "atol": 1e-4,
           "gate_max_fail": max_fail,
           "grid": "256x128, nt=200, Craig-Sneyd; spot mesh K x one unit sinh mesh, one "
                   "power-law variance mesh; bicubic read-out at (S0, v0)"}
    lab = S.labeler().warm()                      # engines, launch shapes, graphs: untimed
    P = S.draw_params(n_labels, seed)
    lab.price(P[:4096])
    fd_adi_amd._lib.sync()
    t = time.perf_counter()
    y = lab.price(P)
    fd_adi_amd._lib.sync()
    t_lab = time.perf_counter() - t
    out["label_sec"] = round(t_lab, 2)
    out["label_options_per_s"] = round(n_labels / t_lab, 1)
    out["label_GBs_algorithmic"] = round(72 * 256 * 128 * 200 * n_labels / t_lab / 1e9, 1)
    log("surrogate: %d labels in %.1fs" % (n_labels, t_lab))
    if after_labels is not None:
        out["cpu_baselines"] = after_labels()
    t = time.perf_counter()
    pricer, info = S.fit_surrogate(P, y, members=members, steps=steps, max_fail=max_fail,
                                   seed=seed)
    fd_adi_amd._lib.sync()
    t_fit = time.perf_counter() - t
    out["fit_sec"] = round(t_fit, 2)
    out.update(info)
    log("surrogate: fit %.1fs %s" % (t_fit, info))
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    Q = S.draw_params((W + K) * q_per_step, seed + 1, outside_frac=0.05)
    # every query in HBM before the timed region; answers and flags stay there
    Qd = SizedArrayPtr(np.ascontiguousarray(Q))
    out_d = SizedArrayPtr().alloc(Q.shape[0])
    ok_d = S.DeviceInts(np.zeros(Q.shape[0], np.int32))

    def serve(k, timed=False):
        r0 = k * q_per_step
        pricer.price_device(S._offset(Qd, 8 * r0), q_per_step, S._offset(out_d, r0),
                            S._offset(ok_d, r0, 4), timed=timed)

    for w in range(W):
        serve(w)
    fd_adi_amd._lib.sync()
    fallback = 0
    t = time.perf_counter()
    for k in range(W, W + K):
        serve(k, timed=(k == W))
        fallback += pricer.last_fallback
    fd_adi_amd._lib.sync()
    t_serve = time.perf_counter() - t
    price = out_d.to_numpy()[W * q_per_step:]
    ok = ok_d.to_numpy()[W * q_per_step:].astype(bool)
    Qt = Q[W * q_per_step:]
    out["inputs"] = ("queries in HBM before the timed region; the gate's trees, the ensemble, "
                     "the semi-analytic baseline and the FP64 ADI fallback on the device; "
                     "answers left in HBM")
    out["serve_sec"] = round(t_serve, 2)
    out["options_per_s"] = round(len(Qt) / t_serve, 1)
    out["fallback_rate"] = round(fallback / float(len(Qt)), 4)
    inside = np.all((Qt >= S.BOX_LO) & (Qt <= S.BOX_HI), axis=1)
    out["admitted_inside_box"] = round(float(ok[inside].mean()), 4)
    out["fallback_stage_ms"] = [round(float(x), 4) for x in lab.last_stage_ms]
    out["fallback_stage_batch"] = lab.last_batch
    t = time.perf_counter()
    pricer.ensemble.predict(Qt)
    fd_adi_amd._lib.sync()
    out["surrogate_only_options_per_s"] = round(len(Qt) / (time.perf_counter() - t), 1)
    t = time.perf_counter()
    S.cf_price(Qt)
    fd_adi_amd._lib.sync()
    out["semi_analytic_options_per_s"] = round(len(Qt) / (time.perf_counter() - t), 1)
    # the admitted answers against FP64 ADI (the fallback answers are FP64 ADI)
    acc = np.nonzero(ok)[0]
    if check >= 0 and acc.size > check:
        acc = np.random.default_rng(seed + 2).choice(acc, check, replace=False)
    out["checked_admitted"] = int(acc.size)
    if acc.size:
        err = np.abs(price[acc] - lab.price(Qt[acc]))
        out["admitted_max_abs_err"] = float(err.max())
        out["admitted_mean_abs_err"] = float(err.mean())
        out["admitted_beyond_atol"] = int((err > 1e-4).sum())
        out["admitted_frac_beyond_atol"] = float(np.mean(err > 1e-4))
        out["admitted_err_quantiles"] = {"%g" % q: float(np.quantile(err, q))
                                         for q in (0.5, 0.99, 0.9999)}
    return out


def cpu_baselines(procs, n_label, n_batch, c1_steps):
    """The oracle timed on the host cores, in child processes (started early,
    run while the GPU works): config 0 and config 1 on one core; the
    surrogate's fallback options (label meshes) and the batched workload's
    options one per process over ``procs`` processes."""
    env = dict(os.environ)
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS",
              "NUMEXPR_NUM_THREADS", "BLIS_NUM_THREADS"):
        env[v] = "1"
    env["HIP_VISIBLE_DEVICES"] = ""
    script = os.path.join(ROOT, "oracle", "cpu_baseline.py")
    cmds = {"config0": ["single", "config0"], "config1": ["single", "config1", str(c1_steps)],
            "labels": ["labels", str(n_label), str(procs)],
            "batch": ["batch", str(n_batch), str(procs)]}
    return {k: subprocess.Popen([sys.executable, script] + a, env=env, stdout=subprocess.PIPE,
                                stderr=subprocess.PIPE, text=True)
            for k, a in cmds.items()}


def collect(procs):
    out = {}
    for k, p in procs.items():
        so, se = p.communicate(timeout=900)
        if p.returncode != 0:
            raise RuntimeError("cpu baseline %s failed: %s" % (k, se[-2000:]))
        out[k] = json.loads(so.strip().splitlines()[-1])
    return out


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


def probe_labels():
    """Fixed fallback-engine dispatches for counter runs (see --probe-labels)."""
    import fd_adi_amd
    from fd_adi_amd import surrogate as S
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    lab = S.AdiLabeler(tune=True)
    P = S.draw_params(4096, 5)
    lab.price(P)
    hb, _ = lab._engine(4096)
    hb.step(20)
    hb.sync()
    x = SizedArrayPtr().alloc(1 << 25)
    for _ in range(5):
        x.timeseq_scalar_from_host(1.0)
    fd_adi_amd._lib.sync()
    log("probe-labels: one 4096-option batch (graph), 20 CS steps (80 sweep launches), "
        "5 scale launches on 2^25 doubles done; shapes %s %s" % (hb.get_shape(0), hb.get_shape(1)))


def probe_batch():
    """Fixed config-2 dispatches for kernel traces (see --probe-batch)."""
    import fd_adi_amd
    from fd_adi_amd.workloads import batch_draws
    opts = batch_draws(4096, 2024)
    hb = fd_adi_amd.HestonBatch([o.params for o in opts], [o.spots for o in opts],
                                [o.vars for o in opts], [o.dt for o in opts])
    hb.set_state([o.payoff() for o in opts])
    hb.configure("cs", 0.5, 1.0)
    hb.set_shape(0, 512, 16)
    hb.set_shape(1, 16)
    hb.step(20)
    hb.set_shape(0, -1, 16)
    hb.set_shape(1, -1)
    hb.step(20)
    hb.sync()
    log("probe-batch: 4096 uneven options, 20 CS steps with one launch shape (512, 16) / 16 "
        "lanes, then 20 with per-option shapes (%d launches per step)" % hb.launches_per_step())


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
    import numpy as np
    import fd_adi_amd
    from fd_adi_amd import surrogate as S
    from fd_adi_amd.pricing import heston_call_device
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    P = S.draw_params(1 << 20, 3)
    rows = np.ascontiguousarray(np.stack([np.full(len(P), 100.0), 100.0 * P[:, 7], P[:, 5],
                                          P[:, 4], P[:, 6], P[:, 0], P[:, 1], P[:, 2],
                                          P[:, 3]], axis=1))
    d_rows, d_out = SizedArrayPtr(rows), SizedArrayPtr().alloc(len(P))
    for _ in range(5):
        heston_call_device(d_rows, len(P), d_out)
    fd_adi_amd._lib.sync()
    log("probe-cf: 5 launches pricing 2^20 options done")


def read_traffic(profile, kernel):
    """HBM bytes per launch of ``kernel`` from a committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", profile)
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
    ap.add_argument("--steps", type=int, default=20,
                    help="timed serving steps (each a batch of --queries-per-step fresh queries)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed serving steps before them")
    ap.add_argument("--queries-per-step", type=int, default=50000,
                    help="queries per serving step (20 steps: the 10^6 queries of the "
                         "serving configuration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="processes of the per-option CPU baselines (0: host cores - 2, <= 14)")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--config1-steps", type=int, default=5, help="timed 2048x1024 solves")
    ap.add_argument("--surrogate-labels", type=int, default=1 << 20)
    ap.add_argument("--surrogate-steps", type=int, default=48000)
    ap.add_argument("--surrogate-members", type=int, default=8)
    ap.add_argument("--gate-max-fail", type=float, default=0.0,
                    help="failure budget of the gate's calibration (0: no calibration query "
                         "beyond 1e-4 is admitted)")
    ap.add_argument("--check-admitted", type=int, default=131072,
                    help="admitted answers checked against FP64 ADI (-1: all of them)")
    ap.add_argument("--probe", action="store_true",
                    help="profiling probe: 20 HV steps of the 2048x1024 option, then 5 "
                         "launches of the scale kernel on 2^25 doubles (the copy reference); "
                         "prints no JSON line")
    ap.add_argument("--probe-labels", action="store_true",
                    help="profiling probe: one 4096-option batch of the fallback engine and "
                         "20 more CS steps of it, then the scale kernel; no JSON line")
    ap.add_argument("--probe-mlp", action="store_true",
                    help="profiling probe: 50 Adam steps of one surrogate network at batch "
                         "8192 on random features and targets; prints no JSON line")
    ap.add_argument("--probe-batch", action="store_true",
                    help="profiling probe: the 4096-option uneven batch, 20 CS steps with one "
                         "launch shape, then 20 with per-option shapes; prints no JSON line")
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

    for flag, fn in ((args.probe, probe), (args.probe_labels, probe_labels),
                     (args.probe_mlp, probe_mlp), (args.probe_cf, probe_cf),
                     (args.probe_batch, probe_batch)):
        if flag:
            fn()
            return

    extra = {}
    # the 2048 x 1024 single option: per-stage roofline of the fused sweeps
    opt = config1()
    N = opt.spots.size * opt.vars.size
    # with the surrogate off the 2048 x 1024 solves are the line's steps
    K1, W1 = ((args.config1_steps, 1) if args.surrogate_labels else (args.steps, args.warmup))
    wall, stages, price = run_single(opt, "hv", 0.5, K1, W1, graph=True)
    per_launch_ms = stages / opt.nt
    stage_gbs = [HV_STAGE_BYTES[k] * N / (per_launch_ms[k] * 1e-3) / 1e9 for k in range(4)]
    payoff_max = float(opt.spots[-1] - opt.K)
    extra["config1_2048x1024_hv"] = {
        "options_per_s": round(K1 / wall, 4),
        "ms_per_option": round(1e3 * wall / K1, 3),
        "stage_us": [round(1e3 * x, 2) for x in per_launch_ms],
        "stage_GBs_algorithmic": [round(x, 1) for x in stage_gbs],
        "step_GBs_algorithmic": round(sum(HV_STAGE_BYTES) * N /
                                      (wall / K1 / opt.nt) / 1e9, 1),
        "price": price, "price_valid": bool(np.isfinite(price) and 0 <= price <= payoff_max),
        "note": "the reference algorithm diverges on this grid (tests/test_oracle_stability.py);"
                " a throughput measurement, not a price"}
    log("config1: %s" % extra["config1_2048x1024_hv"])

    if not args.no_extra:
        from fd_adi_amd.workloads import single
        o0 = single(150, 150, 150, 0.0, spot_max=500.0, var_max=0.2)
        w0, _, p0 = run_single(o0, "douglas", 0.5, 20, 2, smoothing=2, graph=True)
        extra["plain_150x150_douglas_smoothed"] = {
            "options_per_s": round(20 / w0, 2), "price": p0,
            "launches_per_option": 2 * 4 + 2 * 148 + 6, "graph": "one replay per solve"}
        loops = single_loops(2, 16)
        extra["single_option_loop_cs_options_per_s"] = loops
        rate, sec, npts, shapes, cfg_s = batched(4096, 2024, 3, 1)
        extra["batched_4096_cs"] = {
            "options_per_s": round(rate, 1), "sec_per_batch": round(sec, 4),
            "configure_sec": round(cfg_s, 4),
            "options_per_s_with_configure": round(4096 / (sec + cfg_s), 1),
            "GBs_algorithmic": round(72 * npts * 200 / sec / 1e9, 1),
            "shapes": {"spot_threads_cols": list(shapes[0]), "var_lanes": shapes[1][0]},
            "over_op_path_loop": round(rate / loops["op_path"], 1),
            "over_fused_loop": round(rate / loops["fused"], 1)}
        log("extra: %s" % {k: extra[k] for k in extra if k != "config1_2048x1024_hv"})

    # The CPU baselines run in child processes while the labels are generated
    # (GPU-bound; the host waits): they would slow the host-bound
    # measurements above, and are collected before anything timed on the
    # host runs again.
    cpu_procs = None
    if not args.no_cpu:
        nproc = args.cpu_procs or max(1, min(12, (os.cpu_count() or 4) - 4))
        cpu_procs = cpu_baselines(nproc, 2 * nproc, 2 * nproc, 30)
        log("cpu baselines started (%d processes for the per-option ones)" % nproc)

    sur = None
    if args.surrogate_labels:
        sur = surrogate_run(args.surrogate_labels, args.queries_per_step, args.steps,
                            args.warmup, args.surrogate_steps, args.surrogate_members,
                            args.gate_max_fail, args.check_admitted,
                            after_labels=lambda: collect(cpu_procs) if cpu_procs else None)
        extra["surrogate"] = sur
        log("surrogate: %s" % sur)
        if "batched_4096_cs" in extra:
            sur["over_batched_4096"] = round(sur["options_per_s"] /
                                             extra["batched_4096_cs"]["options_per_s"], 2)
        sur["over_fallback_engine"] = round(sur["options_per_s"] / sur["label_options_per_s"],
                                            2)

    cpu = None
    if cpu_procs is not None:
        base = sur.pop("cpu_baselines") if sur is not None else collect(cpu_procs)
        extra["cpu_baselines"] = base
        lab = base["labels"]
        cpu = {"value": lab["options_per_s"], "unit": "options/s", "cores": lab["procs"],
               "kind": "port",
               "sample": "%d options on the serving workload's 256x128 meshes (nt=200, CS), one "
                         "per process, oracle (NumPy/SciPy dia.dot + LAPACK solve_banded), "
                         "host CPU %s" % (lab["options"], lab["cpu_model"])}

    if sur is not None:
        value = sur["options_per_s"]
        ms_per = 1e3 * sur["serve_sec"] / args.steps
        nq = sur["queries"]
        st = np.array(sur["fallback_stage_ms"]) / 200.0      # ms per launch
        pts = sur["fallback_stage_batch"] * 256 * 128
        gbs = [CS_STAGE_BYTES[k] * pts / (st[k] * 1e-3) / 1e9 for k in range(4)]
        dom = int(np.argmax(st))
        roofline = {"bound": "hbm", "achieved": round(gbs[dom], 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs[dom] / HBM_PEAK_GBS, 4),
                    "traffic": read_traffic("hbm_traffic_labels_4096x256x128.txt",
                                            CS_STAGE_NAMES[dom]),
                    "kernel": CS_STAGE_NAMES[dom],
                    "bytes_per_launch": CS_STAGE_BYTES[dom] * pts,
                    "avg_launch_us": round(st[dom] * 1e3, 2),
                    "stage_us": [round(x * 1e3, 2) for x in st],
                    "stage_GBs": [round(x, 1) for x in gbs]}
        workload = "heston_surrogate_serving_%dq_%dlabels_256x128_cs_nt200" % (
            nq, sur["labels"])
        cfg = {"workload": workload, "queries": nq, "queries_per_step": args.queries_per_step,
               "outside_box_frac": 0.05,
               "labels": sur["labels"], "members": sur["members"],
               "gate": "error model (largest of 4 boosted-tree estimates, on the device), calibrated at %g failures "
                       "above 0.9e-4" % sur["gate_max_fail"],
               "fallback": "batched FP64 ADI, 256x128, nt=200, Craig-Sneyd"}
        steps, warmup = args.steps, args.warmup
    else:
        value = extra["config1_2048x1024_hv"]["options_per_s"]
        ms_per = extra["config1_2048x1024_hv"]["ms_per_option"]
        dom = int(np.argmax(stages))
        roofline = {"bound": "hbm", "achieved": round(stage_gbs[dom], 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(stage_gbs[dom] / HBM_PEAK_GBS, 4),
                    "traffic": read_traffic("hbm_traffic_config1.txt", HV_STAGE_NAMES[dom]),
                    "kernel": HV_STAGE_NAMES[dom],
                    "bytes_per_launch": HV_STAGE_BYTES[dom] * N,
                    "avg_launch_us": round(per_launch_ms[dom] * 1e3, 3)}
        cfg = {"workload": "heston_call_2048x1024_hv_nt2000_rho-0.7", "nS": 2048, "nv": 1024,
               "nt": 2000, "scheme": "hundsdorferverwer", "theta": 0.5}
        steps, warmup = K1, W1

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "options/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms_per, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": cfg, "roofline": roofline, "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
