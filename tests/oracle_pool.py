"""Oracle solves in a pool of CPU processes (test infrastructure).

``solve_many`` runs independent oracle solves in spawned interpreters with
one BLAS thread each (nothing of the parent's GPU state is shared), so a GPU
test can hold a large grid or many label rows against the oracle within its
time limit.

A label is the oracle's Craig-Sneyd solve on the surrogate's label meshes
(spot mesh K x one unit sinh mesh, fixed power-law variance mesh) read at
(S0 = 100, v0) by bicubic Lagrange interpolation (oracle/label.py).  Each
label solve is repeated under the rounding-order probes of tests/parity.py;
the largest move of the grid is the floor that sets the tolerance.
"""

import os
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np


def _solve(job):
    """job = (HestonParams fields, spots, vars, oracle switches, method, args)."""
    pf, spots, vars_, kw, method, args = job
    from oracle.engine import HestonOracle, HestonParams
    O = HestonOracle(HestonParams(**pf), spots, vars_, **kw)
    args = tuple(O.payoff() if a is None else a for a in args)
    return getattr(O, method)(*args)


def solve_many(jobs, workers=None):
    """Results of the oracle solves ``jobs`` (see ``_solve``), in order."""
    workers = workers or min(16, max(1, (os.cpu_count() or 2) - 1), len(jobs))
    keep = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                           "MKL_NUM_THREADS")}
    for k in keep:
        os.environ[k] = "1"
    try:
        with ProcessPoolExecutor(max_workers=workers,
                                 mp_context=mp.get_context("spawn")) as ex:
            return list(ex.map(_solve, jobs))
    finally:
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def label_params(r):
    """HestonParams fields of feature row r = (kappa, theta, sigma, rho, v0,
    r, T, K/S0) with S0 = 100."""
    return dict(spot=100.0, strike=100.0 * r[7], r=r[5], v0=r[4], T=r[6], kappa=r[0],
                theta=r[1], sigma=r[2], rho=r[3])


def oracle_labels(P, spots, vars_, nt, probes=None):
    """Oracle grids and labels of feature rows P on the given meshes.

    Returns (grids, labels, floors, label_floors): the Craig-Sneyd grid of
    each row, its read-out at (100, v0), and the largest move of the grid and
    of the read-out under ``probes`` (the switches of tests/parity.PROBES by
    default)."""
    from oracle.label import bicubic_at
    from tests.parity import PROBES
    probes = PROBES if probes is None else probes
    jobs = []
    for r, s, v in zip(P, spots, vars_):
        pf = label_params(np.asarray(r, float))
        for kw in ({},) + tuple(probes):
            jobs.append((pf, np.asarray(s, float), np.asarray(v, float), dict(kw),
                         "solve_craigsneyd", (int(nt), r[6] / nt, None, 0.5)))
    out = solve_many(jobs)
    m = 1 + len(probes)
    grids, labels, floors, lfloors = [], [], [], []
    for k, (r, s, v) in enumerate(zip(P, spots, vars_)):
        W = out[k * m]
        alts = out[k * m + 1:(k + 1) * m]
        y = bicubic_at(W, s, v, 100.0, r[4])
        grids.append(W)
        labels.append(y)
        floors.append(max([float(np.max(np.abs(W - a))) for a in alts] or [0.0]))
        lfloors.append(max([abs(y - bicubic_at(a, s, v, 100.0, r[4])) for a in alts] or [0.0]))
    return grids, np.array(labels), np.array(floors), np.array(lfloors)
