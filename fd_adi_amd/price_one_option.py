"""Price one Heston option on the GPU engine from the command line.

Mirrors the reference driver price_one_option.py:48-208 (same options and
defaults, same engine construction, the same "FD: <price>" line and the same
result file naming), for the GPU engine only:

    python -m fd_adi_amd.price_one_option -nx 150 150 -nt 150 --gpu --scheme hv

The result file is the reference's: ``[grid.mesh, grid.domain[-1]]`` pickled
with the highest protocol into ``<name>.txt`` (price_one_option.py:193,
199-208), written to ``--out-dir`` (default: the current directory).
Differences: ``--cpu`` and ``--mc`` (the reference's CPU engine and Monte
Carlo) are not part of this package and are rejected; ``--npz`` writes a
NumPy ``.npz`` (``spots``, ``vars``, ``V``) instead of the pickle;
``--scheme`` also accepts ``cs`` and ``mcs``.
"""

import argparse
import os
import pickle
import sys

import numpy as np

SCHEMES = {"i": "solve_implicit", "d": "solve_douglas", "hv": "solve_hundsdorferverwer",
           "s": "solve_smooth", "cs": "solve_craigsneyd", "mcs": "solve_craigsneyd2"}


def read_args(argv=None):
    """Argument parser of price_one_option.py:48-77."""
    p = argparse.ArgumentParser(description="Price one Heston option with the ADI engine")
    p.add_argument("--scheme", default="hv", choices=sorted(SCHEMES))
    p.add_argument("-s", "--spot", metavar="FLOAT", type=float, default=100.0)
    p.add_argument("-k", "--strike", metavar="FLOAT", type=float, default=99.0)
    p.add_argument("-t", "--tenor", metavar="FLOAT", type=float, default=1.0)
    p.add_argument("-r", "--interest-rate", metavar="FLOAT", type=float, default=0.06)
    p.add_argument("--mean-reversion", metavar="FLOAT", type=float, default=1.0)
    p.add_argument("--variance", metavar="FLOAT", type=float, default=0.04)
    p.add_argument("-o", "--vol-of-var", metavar="FLOAT", type=float, default=0.001)
    p.add_argument("--mean-variance", metavar="FLOAT", type=float, default=-1.0)
    p.add_argument("-p", "--correlation", metavar="FLOAT", type=float, default=0.0)
    p.add_argument("--top", type=float, default=None)
    p.add_argument("--barrier", action="store_true")
    p.add_argument("-nx", required=True, metavar="int", nargs=2, type=int)
    p.add_argument("-nt", required=True, metavar="int", type=int)
    p.add_argument("-v", action="count", dest="verbose", default=0)
    p.add_argument("--gpu", action="store_true")
    p.add_argument("--cpu", action="store_true")
    p.add_argument("--mc", metavar="int", type=int, default=0)
    p.add_argument("--out-dir", default=".")
    p.add_argument("--npz", action="store_true",
                   help="write spots, vars and V as .npz instead of the pickled list")
    opt = p.parse_args(argv)
    if opt.top is not None and not opt.barrier:
        p.error("--barrier required with --top")
    if opt.cpu or opt.mc:
        p.error("--cpu and --mc (the reference's CPU engine and Monte Carlo) are not part of "
                "this package; use --gpu")
    if opt.mean_variance == -1:
        opt.mean_variance = opt.variance
    return opt


def new_option(opt):
    from .Option import HestonOption, HestonBarrierOption
    kw = dict(spot=opt.spot, strike=opt.strike, interest_rate=opt.interest_rate,
              variance=opt.variance, tenor=opt.tenor, mean_reversion=opt.mean_reversion,
              mean_variance=opt.mean_variance, vol_of_variance=opt.vol_of_var,
              correlation=opt.correlation)
    if opt.barrier:
        return HestonBarrierOption(top=(False, opt.top), **kw)
    return HestonOption(**kw)


def new_engine(opt):
    """Engine construction of price_one_option.py:98-121."""
    from .FiniteDifferenceEngineGPU import HestonFiniteDifferenceEngine
    return HestonFiniteDifferenceEngine(
        new_option(opt), grid=None, spot_max=1500.0, spot_min=0.0, spots=None, vars=None,
        var_max=10.0, nspots=opt.nx[0], nvols=opt.nx[1], spotdensity=7.0, varexp=4.0,
        force_exact=True, verbose=opt.verbose)


def filestring(opt, engine):
    """Result file name of price_one_option.py:166-193 (``.txt``, or ``.npz``
    with --npz)."""
    idxs, idxv = engine.idx
    parts = ["gpu", "moneyness", opt.spot / opt.strike, "tenor", opt.tenor,
             "ir", opt.interest_rate, "var", opt.variance, "volofvar", opt.vol_of_var,
             "corr", opt.correlation, "rev", opt.mean_reversion, "nspot", opt.nx[0],
             "nvols", opt.nx[1], "idxs", idxs, "idxv", idxv, "ntime", opt.nt]
    ext = ".npz" if getattr(opt, "npz", False) else ".txt"
    return os.path.join(opt.out_dir, "_".join(map(str, parts)) + ext)


def run(opt):
    """Solve with the chosen scheme and return (engine, price)
    (price_one_option.py:124-145)."""
    e = new_engine(opt)
    s, v = e.idx
    found = (e.spots[s], e.vars[v])
    if not np.allclose((opt.spot, opt.variance), found, rtol=0, atol=1e-10):
        raise AssertionError("We have the wrong indices! %s %s"
                             % ((opt.spot, opt.variance), found))
    getattr(e, SCHEMES[opt.scheme])(opt.nt, opt.tenor / opt.nt)
    return e, float(e.grid.domain[-1][s, v])


def main(argv=None):
    opt = read_args(argv)
    e, price = run(opt)
    print("FD:", price)
    fn = filestring(opt, e)
    if opt.npz:
        np.savez(fn, spots=e.spots, vars=e.vars, V=e.grid.domain[-1])
    else:
        with open(fn, "wb") as f:
            pickle.dump([e.grid.mesh, e.grid.domain[-1]], f, pickle.HIGHEST_PROTOCOL)
    if opt.verbose:
        print("wrote", fn)
    return 0


if __name__ == "__main__":
    sys.exit(main())
