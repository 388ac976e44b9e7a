"""Loader of the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): inputs of one Heston solve and the oracle's
value grid for it."""

import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Case(object):
    def __init__(self, path):
        z = np.load(path)  # plain arrays only; no pickles
        self.meta = json.loads(str(z["meta"]))
        self.name = self.meta["name"]
        self.scheme = self.meta["scheme"]
        self.n, self.dt, self.theta = self.meta["n"], self.meta["dt"], self.meta["theta"]
        self.p = self.meta["params"]
        self.barrier = self.meta["barrier"]
        self.idx = tuple(self.meta["idx"])
        for k in ("spots", "vars", "U0", "V", "floors"):
            setattr(self, k, z[k])
        self.floor = float(np.max(self.floors))
        self.price_cf = float(z["price_cf"])


def names():
    return sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    return Case(os.path.join(GOLDEN, name + ".npz"))


def solve(E, c, U0):
    """Run fixture ``c``'s scheme on engine ``E`` (the oracle or the GPU engine,
    which share the solve_* signatures)."""
    if c.scheme == "implicit":
        return E.solve_implicit(c.n, c.dt, U0)
    if c.scheme == "smooth":
        return E.solve_smooth(c.n, c.dt, U0)
    if c.scheme == "douglas_smoothed":
        V = E.solve_implicit(4, c.dt * 0.5, U0)
        return E.solve_douglas(c.n - 2, c.dt, V, theta=c.theta)
    return getattr(E, "solve_" + c.scheme)(c.n, c.dt, U0, theta=c.theta)
