"""Parity criteria shared by the GPU tests.

Price: relative error at the (S0, v0) node <= 1e-12 (strict).
Grid: max |V_gpu - V_oracle| <= max(1e-10, 10 * floor), where ``floor`` is
the oracle's own rounding noise: the difference between the oracle and the
same oracle summing its explicit operators in the opposite order (the
reference's order is Python 2 dict order, i.e. arbitrary).  Far-field rows
(S near Smax, values ~1e3, coefficients ~1e4) make that floor exceed 1e-10
on the reference's own grids, so a fixed 1e-10 would reject the oracle
against itself.
"""

import numpy as np

from oracle.engine import HestonOracle

PRICE_RTOL = 1e-12
GRID_ATOL = 1e-10
FLOOR_FACTOR = 10.0


def reordered(O):
    return HestonOracle(O.p, O.spots, O.vars, firsts_reversed=True)


def grid_tolerance(W, W_alt):
    floor = float(np.max(np.abs(W - W_alt)))
    return max(GRID_ATOL, FLOOR_FACTOR * floor), floor


def check(V, W, W_alt, idx):
    tol, floor = grid_tolerance(W, W_alt)
    err = float(np.max(np.abs(V - W)))
    rel = abs(V[idx] - W[idx]) / abs(W[idx])
    msg = "grid err %.3e (tol %.3e, oracle floor %.3e), price rel %.3e" % (err, tol, floor, rel)
    print(msg)
    assert err <= tol, msg
    assert rel <= PRICE_RTOL, msg
    return err, rel, floor
