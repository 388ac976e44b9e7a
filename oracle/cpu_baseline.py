"""CPU baselines for bench.py -- the oracle timed on the host cores.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py's
``cpu_baseline`` leg runs these functions in child processes, as the
reference CPU engine it stands in for (the NumPy/SciPy restatement of
FiniteDifference/FiniteDifferenceEngine.pyx; the reference itself is Python
2 and cannot run here).  Every solve uses one BLAS/OpenMP thread.

* ``single``: one option of a configuration on one core, a bounded number of
  its steps, extrapolated to the whole solve;
* ``per_process``: one option per process over ``procs`` processes, the way
  the reference runs many options (run_parameters.py:16-36, one
  price_one_option.py process per option), on the first ``n`` options of the
  seeded draw.
"""

import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREAD_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS",
               "NUMEXPR_NUM_THREADS", "BLIS_NUM_THREADS")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_of(o):
    from oracle.engine import HestonOracle, HestonParams
    return HestonOracle(HestonParams(spot=o.S0, strike=o.K, r=o.r, v0=o.v0, T=o.T,
                                     kappa=o.kappa, theta=o.theta, sigma=o.sigma, rho=o.rho),
                        o.spots, o.vars)


def _solve_option(args):
    """One whole option solve (Craig-Sneyd, the batched configuration's
    scheme, or config 0's composition); returns (seconds, price)."""
    kind, o = args
    E = _oracle_of(o)
    t = time.perf_counter()
    if kind == "config0":
        V = E.solve_implicit(4, o.dt * 0.5, o.payoff())
        V = E.solve_douglas(o.nt - 2, o.dt, V, theta=0.5)
    else:
        V = E.solve_craigsneyd(o.nt, o.dt, o.payoff(), 0.5)
    return time.perf_counter() - t, float(V[o.idx])


def single(kind, steps=None):
    """Seconds per option of one option on one core: "config0" (150 x 150,
    smoothing + Douglas, whole solve) or "config1" (2048 x 1024 HV, ``steps``
    of its 2000 steps, extrapolated)."""
    from fd_adi_amd.workloads import single as one
    if kind == "config0":
        o = one(150, 150, 150, 0.0, spot_max=500.0, var_max=0.2)
        _solve_option(("config0", o))              # warm-up
        sec, price = _solve_option(("config0", o))
        return {"sec_per_option": sec, "price": price, "steps": o.nt}
    o = one(2048, 1024, 2000, -0.7)
    E = _oracle_of(o)
    U0 = o.payoff()
    E.solve_hundsdorferverwer(1, o.dt, U0)
    t = time.perf_counter()
    E.solve_hundsdorferverwer(steps, o.dt, U0)
    sec = (time.perf_counter() - t) / steps
    return {"sec_per_option": sec * o.nt, "sec_per_step": sec, "steps": steps}


def per_process(n, procs, seed=2024, label_mesh=False):
    """Options/s of the first ``n`` options of the batched draw (or, with
    ``label_mesh``, of the surrogate's label meshes), one option per process
    over ``procs`` processes."""
    from fd_adi_amd.workloads import batch_draws, label_draws
    opts = label_draws(n, seed) if label_mesh else batch_draws(n, seed)[:n]
    ctx = mp.get_context("spawn")
    t = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_solve_option, [("cs", o) for o in opts], chunksize=1)
    wall = time.perf_counter() - t
    return {"options": len(opts), "procs": procs, "wall_s": wall,
            "options_per_s": len(opts) / wall,
            "mean_sec_per_option": sum(r[0] for r in res) / len(res)}


def main():
    for v in THREAD_VARS:
        os.environ[v] = "1"
    sys.path.insert(0, ROOT)
    kind = sys.argv[1]
    if kind == "single":
        out = single(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else None)
    else:
        out = per_process(int(sys.argv[2]), int(sys.argv[3]), label_mesh=(kind == "labels"))
    out["cpu_model"] = cpu_model()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
