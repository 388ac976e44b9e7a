"""Pins of the oracle's operator restatement (CPU only).

Each check is chosen so that a plausible slip (wrong sign, swapped h-/h+,
a transposed band, an off-by-one row) fails it:
* stencils are exact on quadratics on a nonuniform mesh;
* the fold/unfold of the reference's 5x5 known-answer matrix
  (FiniteDifference/tests/test_BandedOperatorCPU.py:341-353) equals the
  elimination matrix of utils.foldMatFor (FiniteDifference/utils.py:505-519);
* apply/solve equal dense numpy algebra;
* the cross operator equals the reference test's closed-form COO
  construction (FiniteDifference/tests/test_BandedOperatorGPU.py:349-420) and
  the np.gradient identity (tests/test_FiniteDifferenceEngine.py:508-566).
"""

import numpy as np
import numpy.testing as npt
import pytest
import scipy.sparse

from oracle import banded as BO
from oracle.engine import HestonOracle, HestonParams
from fd_adi_amd.utils import sinh_space, exponential_space


def nonuniform(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(np.hstack(([0.0], rng.uniform(0.5, 1.5, n - 1))))


# -- stencils ---------------------------------------------------------------------------
@pytest.mark.parametrize("derivative", [1, 2])
def test_center_stencil_exact_on_quadratics(derivative):
    x = nonuniform(12, 1)
    B = BO.for_vector(x, derivative=derivative)
    for f, d1, d2 in [(x ** 2, 2 * x, 2 + 0 * x), (3 * x - 1, 3 + 0 * x, 0 * x),
                      (np.ones_like(x), 0 * x, 0 * x)]:
        got = B.D.dot(f)[1:-1]
        want = (d1 if derivative == 1 else d2)[1:-1]
        npt.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_center_first_matches_numpy_gradient():
    x = nonuniform(15, 2)
    f = np.sin(x)
    B = BO.for_vector(x, derivative=1)
    npt.assert_allclose(B.D.dot(f)[1:-1], np.gradient(f, x)[1:-1], rtol=1e-13, atol=1e-14)


# -- folding (reference known answer) -------------------------------------------------
KA = np.array([[-1.5, 0.5, 2, 0, 0],
               [-1, 2, -1, 0, 0],
               [0, -1, 2, -1, 0],
               [0, 0, -1, 2, -1],
               [0, 0, -1.5, 0.5, 2]])


def ka_operator(blocks=2):
    A = scipy.sparse.block_diag([KA] * blocks).toarray()
    D = scipy.sparse.dia_matrix(A)
    order = np.argsort(D.offsets)[::-1]
    B = BO.BandedOperator(D.data[order], D.offsets[order])
    B.blocks = blocks
    return B, A


def fold_matrix(A, blocks):
    """utils.foldMatFor: the elimination matrix X with X*A tridiagonal."""
    n = A.shape[0]
    l = n // blocks
    X = np.eye(n)
    for b in range(blocks):
        s, e = b * l, (b + 1) * l - 1
        X[s, s + 1] = -A[s, s + 2] / A[s + 1, s + 2] if A[s + 1, s + 2] else 0
        X[e, e - 1] = -A[e, e - 2] / A[e - 1, e - 2] if A[e, e - 2] else 0
    return X


def test_fold_known_answer():
    B, A = ka_operator(2)
    X = fold_matrix(A, 2)
    XA = X.dot(A)
    B.diagonalize()
    assert B.is_tridiagonal() and B.is_folded()
    npt.assert_allclose(B.dense(), XA, rtol=0, atol=1e-15)
    # fold factors are the off-diagonal entries of X
    npt.assert_allclose(B.top_factors, [X[0, 1], X[5, 6]])
    npt.assert_allclose(B.bottom_factors, [X[4, 3], X[9, 8]])
    v = np.random.default_rng(3).random(10)
    npt.assert_allclose(B.fold_vector(v.copy()), X.dot(v), rtol=0, atol=1e-15)
    npt.assert_allclose(B.fold_vector(B.fold_vector(v.copy()), unfold=True), v, atol=1e-14)


def test_diagonalize_roundtrip_and_folded_algebra():
    B, A = ka_operator(2)
    v = np.random.default_rng(4).random(10)
    ref_apply = A.dot(v)
    ref_solve = np.linalg.solve(A, v)
    C = B.copy()
    C.diagonalize()
    npt.assert_allclose(C.apply(v), ref_apply, rtol=1e-13)
    npt.assert_allclose(C.solve(v), ref_solve, rtol=1e-12)
    C.undiagonalize()
    assert not C.is_tridiagonal()
    npt.assert_allclose(C.dense(), A, atol=1e-15)


# -- apply / solve with Dirichlet and residual ----------------------------------------
def test_apply_solve_dense_with_dirichlet():
    rng = np.random.default_rng(5)
    n0, n1 = 7, 5
    x = nonuniform(n0, 6)
    B = BO.for_vector(x, derivative=2, axis=0)
    B.applyboundary(((0, 0.0), (None, 1.0)))
    Bs = BO.flatten_blocks([B.copy() for _ in range(n1)])
    Bs.dirichlet[0] = tuple([0.0] * n1)
    Bs = Bs.mul(-0.1).add(1)
    V = rng.random((n0, n1))
    # dense reference: operator acts on the spot axis of every variance column
    A = B.mul(-0.1).add(1).dense()
    for j in range(n1):
        col = V[:, j].copy()
        col[0] = 0.0
        npt.assert_allclose(Bs.apply(V)[:, j], A.dot(col) + Bs.R[j * n0:(j + 1) * n0],
                            rtol=1e-13)
        sol = Bs.solve(V)[:, j]
        rhs = V[:, j].copy()
        rhs[0] = 0.0
        npt.assert_allclose(A.dot(sol), rhs - Bs.R[j * n0:(j + 1) * n0], rtol=1e-12, atol=1e-13)
        assert sol[0] == 0.0


def test_boundary_rows_explicit():
    x = nonuniform(8, 7)
    d = np.diff(x)
    # (0,0) with Dirichlet low and known first derivative 1 at the top
    B = BO.for_vector(x, derivative=2)
    B.applyboundary(((0, 0.0), (None, 1.0)))
    A = B.dense()
    assert A[0, 0] == 1 and A[0, 1] == 0
    npt.assert_allclose(A[-1, -2:], [2 / d[-1] ** 2, -2 / d[-1] ** 2])
    npt.assert_allclose(B.R[-1], 2 / d[-1])
    # (1,): free first-order one-sided rows
    B = BO.for_vector(x, derivative=1)
    B.applyboundary(((None, None), (None, None)))
    A = B.dense()
    npt.assert_allclose(A[0, :2], [-1 / d[0], 1 / d[0]])
    npt.assert_allclose(A[-1, -2:], [-1 / d[-1], 1 / d[-1]])
    # (1,1): VN 0 at the bottom, second-order backward (with -2 band) at the top
    B = BO.for_vector(x, derivative=2, force_bandwidth=(-2, 1))
    B.applyboundary(((1, 0.0), (None, None)))
    A = B.dense()
    assert np.all(A[0] == 0) and B.R[0] == 0
    f = x ** 2  # the one-sided second derivative is exact on quadratics
    npt.assert_allclose(A[-1].dot(f), 2.0, rtol=1e-12)
    assert B.bottom_fold_status == BO.CAN_FOLD


# -- cross derivative --------------------------------------------------------------------
def test_cross_matches_reference_closed_form():
    m0 = nonuniform(6, 8)
    m1 = nonuniform(5, 9) * 0.1
    n0, n1 = len(m0), len(m1)
    d0 = np.hstack((np.nan, np.diff(m0)))
    d1 = np.hstack((np.nan, np.diff(m1)))
    sup = d1[1:n1 - 1] / (d1[2:n1] * (d1[1:n1 - 1] + d1[2:n1]))
    mid = (-d1[1:n1 - 1] + d1[2:n1]) / (d1[1:n1 - 1] * d1[2:n1])
    sub = -d1[2:n1] / (d1[1:n1 - 1] * (d1[1:n1 - 1] + d1[2:n1]))
    dsup = np.repeat(d0[1:n0 - 1] / (d0[2:n0] * (d0[1:n0 - 1] + d0[2:n0])), n1 - 2)
    dmid = np.repeat((-d0[1:n0 - 1] + d0[2:n0]) / (d0[1:n0 - 1] * d0[2:n0]), n1 - 2)
    dsub = np.repeat(-d0[2:n0] / (d0[1:n0 - 1] * (d0[1:n0 - 1] + d0[2:n0])), n1 - 2)
    t = lambda a: np.tile(a, n0 - 2)
    vals = [t(sub) * dsub, t(mid) * dsub, t(sup) * dsub,
            t(sub) * dmid, t(mid) * dmid, t(sup) * dmid,
            t(sub) * dsup, t(mid) * dsup, t(sup) * dsup]
    row = np.tile(np.arange(1, n1 - 1), n0 - 2) + np.repeat(np.arange(1, n0 - 1), n1 - 2) * n1
    offs = [-n1 - 1, -n1, -n1 + 1, -1, 0, 1, n1 - 1, n1, n1 + 1]
    A = np.zeros((n0 * n1, n0 * n1))
    for v, o in zip(vals, offs):
        A[row, row + o] = v
    B = BO.mixed_operator(m0, m1)
    npt.assert_array_equal(B.dense(), A)


def test_cross_matches_numpy_gradient():
    m0 = np.linspace(0, 3, 9) ** 1.5
    m1 = np.linspace(0, 1, 7) ** 2
    X, Y = np.meshgrid(m0, m1, indexing="ij")
    V = np.sin(X) * np.cos(2 * Y) + X * Y ** 2
    B = BO.mixed_operator(m0, m1)
    got = B.apply(V)
    want = np.gradient(np.gradient(V, m0, axis=0), m1, axis=1)
    npt.assert_allclose(got[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-12, atol=1e-12)
    assert np.all(got[0] == 0) and np.all(got[-1] == 0)
    assert np.all(got[:, 0] == 0) and np.all(got[:, -1] == 0)


# -- engine construction -------------------------------------------------------------
def small_engine(rho=0.4, barrier=None):
    p = HestonParams(spot=100, strike=100, r=0.03, v0=0.04, T=1.0, kappa=1.0, theta=0.12,
                     sigma=0.3, rho=rho, barrier=barrier)
    spots = sinh_space(100, 500, 7.0, 12)
    vars_ = exponential_space(0, 0.04, 1.0, 4.0, 9)
    return HestonOracle(p, spots, vars_)


def test_heston_coefficients_closed_form():
    E = small_engine()
    p = E.p
    S, v = np.meshgrid(E.spots, E.vars, indexing="ij")
    npt.assert_array_equal(E.coefficient_vector((0,), 0), (p.r * S).T.ravel())
    npt.assert_allclose(E.coefficient_vector((0, 0), 0), (0.5 * v * S ** 2).T.ravel(), rtol=1e-15)
    want = p.kappa * (p.theta - v)
    want[S == 0] = 0
    npt.assert_allclose(E.coefficient_vector((1,), 1), want.ravel(), rtol=1e-15)
    npt.assert_allclose(E.coefficient_vector((0, 1), 1), (p.rho * p.sigma * S * v).ravel(),
                        rtol=1e-15)


def test_combined_operators_dense():
    """A0 = r S d/dS + 1/2 v S^2 d2/dS2 - r/2 with identity Dirichlet rows,
    rebuilt densely from the 1-D stencils for one variance column."""
    E = small_engine()
    nS, nv = E.shape
    j = 4
    v = E.vars[j]
    D1 = BO.for_vector(E.spots, 1).dense()
    D2 = BO.for_vector(E.spots, 2).dense()
    h = E.spots[-1] - E.spots[-2]
    D2[-1, -2:] = [2 / h ** 2, -2 / h ** 2]
    want = np.diag(E.p.r * E.spots).dot(D1) + np.diag(0.5 * v * E.spots ** 2).dot(D2)
    want -= np.diag(np.full(nS, E.p.r / 2))
    want[0] = 0
    want[0, 0] = 1
    A0 = E.operators[0].dense()[j * nS:(j + 1) * nS, j * nS:(j + 1) * nS]
    npt.assert_allclose(A0, want, rtol=1e-13, atol=1e-13)
    R = E.operators[0].R[j * nS:(j + 1) * nS]
    npt.assert_allclose(R[-1], E.p.r * E.spots[-1] + 0.5 * v * E.spots[-1] ** 2 * 2 / h)
    assert np.all(R[:-1] == 0)
