"""Host-side mesh generation for the Heston grids.

The reference builds its meshes on the host before uploading them
(FiniteDifference/FiniteDifferenceEngineGPU.pyx:658-686); this module restates
those generators so the engine produces identical grids.
"""

from bisect import bisect_left

import numpy as np
import scipy.optimize


def sinh_space(exact, high, density, size, force_exact=True):
    """Sinh-clustered mesh on [0, ~high] with nodes dense around ``exact``.

    FiniteDifference/utils.py:102-147.  The mesh is
    ``K + c*sinh(eps_k)`` shifted so its minimum is 0, with the eps_k uniformly
    spaced between asinh(-K/c) and asinh((Smax-K)/c).  ``high`` is retuned
    inside [0.95, 1.05]*high (bounded scalar minimisation) so that the node
    nearest above ``exact`` is as close to it as possible; with
    ``force_exact`` that node is then set to ``exact``.
    """
    density = float(density)
    if size == 1:
        return np.asarray(exact)

    def newspace(smax):
        deps = 1.0 / size * (np.arcsinh((smax - exact) * (1 / density))
                             - np.arcsinh(-exact / density))
        eps = np.arcsinh(-exact / density) + np.arange(size) * deps
        space = exact + density * np.sinh(eps)
        space -= min(space)
        return space

    def energy(smax):
        space = newspace(smax)
        try:
            best = space[bisect_left(space, exact)]
        except IndexError:
            best = space[-2]
        return abs(exact - best)

    newmax, _, failed, _ = scipy.optimize.fminbound(
        energy, high * 0.95, high * 1.05, full_output=True)
    if failed:
        raise ValueError("Unable to create a suitable space. %s" %
                         ((exact, high, density, size),))
    space = newspace(newmax)
    if force_exact:
        idx = bisect_left(space, exact)
        if idx < len(space):
            space[idx] = exact
    return space


def exponential_space(low, exact, high, ex, size, force_exact=True):
    """Power-law mesh v_i = low**(1/ex) + (i*dv)**ex.

    FiniteDifference/utils.py:161-190.  With ``force_exact`` the top end is
    raised so that node j (the last node below ``exact``) lands on ``exact``.
    """
    if size == 1:
        return np.array([exact])
    v = np.zeros(size)
    lo = pow(low, 1. / ex)
    hi = pow(high, 1. / ex)
    x = pow(exact, 1. / ex)
    dv = (hi - lo) / (size - 1)
    if force_exact:
        j = 0
        for i in range(size):
            if lo + i * dv > x:
                j = i - 1
                break
        if j == 0:
            raise AssertionError("Did not find suitable value.")
        dx = x - (lo + j * dv)
        hi += (size - 1) * dx / j
    dv = (hi - lo) / (size - 1)
    for i in range(size):
        v[i] = lo + pow(i * dv, ex)
    return v


def nearest_index(mesh, value, decimals=4):
    """Index of ``value`` on ``mesh`` the way the engines' ``idx`` finds it
    (bisect on the mesh rounded to ``decimals``,
    FiniteDifferenceEngineGPU.pyx:727-734)."""
    return bisect_left(list(np.round(mesh, decimals)), round(float(value), decimals))


def sinh_space_batch(exact, high, density, size):
    """Sinh meshes for many (exact, high) rows at once: the mesh family of
    ``sinh_space`` on [0, high] with the node nearest ``exact`` forced onto it,
    without the retuning of the top end (which needs a scalar bounded
    minimisation per mesh).  Used for the surrogate's label grids, where the
    same function builds the grids of the training labels and of the ADI
    fallback."""
    exact = np.asarray(exact, dtype=float)
    high = np.asarray(high, dtype=float) * np.ones_like(exact)
    density = float(density)
    k = np.arange(size)
    deps = 1.0 / size * (np.arcsinh((high - exact) / density) - np.arcsinh(-exact / density))
    eps = np.arcsinh(-exact / density)[:, None] + k[None, :] * deps[:, None]
    sp = exact[:, None] + density * np.sinh(eps)
    sp -= sp.min(axis=1, keepdims=True)
    rows = np.arange(sp.shape[0])
    sp[rows, np.argmin(np.abs(sp - exact[:, None]), axis=1)] = exact
    return sp


def exponential_space_batch(low, exact, high, ex, size):
    """Vectorised exponential_space (force_exact) for many (exact, high) rows."""
    exact = np.asarray(exact, dtype=float)
    high = np.asarray(high, dtype=float) * np.ones_like(exact)
    lo = pow(low, 1.0 / ex)
    hi = high ** (1.0 / ex)
    x = exact ** (1.0 / ex)
    dv = (hi - lo) / (size - 1)
    i = np.arange(size)
    above = (lo + i[None, :] * dv[:, None]) > x[:, None]
    j = np.argmax(above, axis=1) - 1
    if np.any(j <= 0) or not np.all(above.any(axis=1)):
        raise AssertionError("Did not find suitable value.")
    dx = x - (lo + j * dv)
    hi = hi + (size - 1) * dx / j
    dv = (hi - lo) / (size - 1)
    return lo + (i[None, :] * dv[:, None]) ** ex
