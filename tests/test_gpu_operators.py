"""GPU operator algebra vs the oracle (reference tests
FiniteDifference/tests/test_BandedOperatorGPU.py and test_SizedArray.py).

Construction is compared bit for bit (the reference asserts array_equal for
for_vector, test_BandedOperatorGPU.py:292-337); apply at decimal 12
(:140-163); solve and algebra at rtol 1e-12.
"""

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu

from oracle import banded as OB
from oracle.engine import HestonOracle, HestonParams
from fd_adi_amd.utils import sinh_space, exponential_space


@pytest.fixture(scope="module")
def fd(gpu):
    import fd_adi_amd
    return fd_adi_amd


def engine_pair(fd, nS=20, nv=12, rho=0.4, barrier=None):
    from fd_adi_amd.FiniteDifferenceEngineGPU import HestonFiniteDifferenceEngine
    from fd_adi_amd.Option import HestonOption, HestonBarrierOption
    if barrier:
        opt = HestonBarrierOption(spot=100, strike=100, interest_rate=0.03, variance=0.04,
                                  tenor=1.0, mean_reversion=1.0, mean_variance=0.12,
                                  vol_of_variance=0.3, correlation=rho, top=(False, barrier))
    else:
        opt = HestonOption(spot=100, strike=100, interest_rate=0.03, variance=0.04, tenor=1.0,
                           mean_reversion=1.0, mean_variance=0.12, vol_of_variance=0.3,
                           correlation=rho)
    G = HestonFiniteDifferenceEngine(opt, nspots=nS, nvols=nv, spot_max=500.0, var_max=1.0)
    p = HestonParams(spot=100, strike=100, r=0.03, v0=0.04, T=1.0, kappa=1.0, theta=0.12,
                     sigma=0.3, rho=rho, barrier=barrier)
    O = HestonOracle(p, G.spots, G.vars)
    return G, O


def same_layout(host, ref):
    npt.assert_array_equal(host.offsets, ref.offsets)
    npt.assert_array_equal(host.data, ref.data)
    npt.assert_array_equal(host.R, ref.R)


# -- device arrays (test_SizedArray.py) -------------------------------------------
def test_sized_array_roundtrip_and_ops(fd):
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    rng = np.random.default_rng(0)
    a, b = rng.random((7, 5)), rng.random((7, 5))
    A, B = SizedArrayPtr(a), SizedArrayPtr(b)
    npt.assert_array_equal(A.to_numpy(), a)
    npt.assert_array_equal(A.copy(True).pluseq(B).to_numpy(), a + b)
    npt.assert_array_equal(A.copy(True).minuseq(B).to_numpy(), a - b)
    npt.assert_array_equal(A.copy(True).minuseq_over2(B).to_numpy(), (a - b) * 0.5)
    npt.assert_array_equal(A.copy(True).timeseq(B).to_numpy(), a * b)
    npt.assert_array_equal(A.copy(True).pluseq_scalar_from_host(23.214).to_numpy(), a + 23.214)
    npt.assert_array_equal(A.copy(True).timeseq_scalar_from_host(-2.5).to_numpy(), a * -2.5)
    s = SizedArrayPtr(np.array([1.5, 3.0]))
    npt.assert_array_equal(A.copy(True).minuseq_scalar(s, 1).to_numpy(), a - 3.0)
    X = SizedArrayPtr().alloc(A.size)
    X.copy_from(A)
    npt.assert_array_equal(X.to_numpy(), a)
    T = A.copy(True).transpose()
    npt.assert_array_equal(T.to_numpy(), a.T)


def test_shallow_copy_dies_with_owner(fd):
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    S = SizedArrayPtr(np.ones(8))
    V = S.copy(False)
    npt.assert_array_equal(V.to_numpy(), np.ones(8))
    del S
    with pytest.raises(RuntimeError):
        V.to_numpy()


# -- construction: bit-for-bit -------------------------------------------------------
@pytest.mark.parametrize("barrier", [None, 170.0])
def test_for_vector_matches_oracle_bitwise(fd, barrier):
    import fd_adi_amd.BandedOperatorGPU as BOG
    G, O = engine_pair(fd, barrier=barrier)
    m0, m1 = G.spots, G.vars
    cases = [((0,), m0, m1.size, 1, 0), ((0, 0), m0, m1.size, 2, 0),
             ((1,), m1, m0.size, 1, 1), ((1, 1), m1, m0.size, 2, 1)]
    for d, mesh, blocks, der, axis in cases:
        host = BOG.for_vector(mesh, blocks, der, axis, bool(barrier)).immigrate()
        ref = O.simple_operators[d]
        same_layout(host, ref)
        assert host.bottom_fold_status == ref.bottom_fold_status
        lo = None if ref.dirichlet[0] is None else np.atleast_1d(ref.dirichlet[0])
        if lo is None:
            assert host.dirichlet[0] is None
        else:
            npt.assert_array_equal(np.broadcast_to(lo, (blocks,)), host.dirichlet[0])


def test_mixed_for_vector_matches_oracle_bitwise(fd):
    import fd_adi_amd.BandedOperatorGPU as BOG
    G, O = engine_pair(fd)
    host = BOG.mixed_for_vector(G.spots, G.vars).immigrate()
    npt.assert_array_equal(host.D.toarray(), O.simple_operators[(0, 1)].dense())


@pytest.mark.parametrize("barrier", [None, 170.0])
def test_scale_and_combine_matches_oracle(fd, barrier):
    G, O = engine_pair(fd, barrier=barrier)
    G.set_zero_derivative()
    G.scale_and_combine_operators()
    for k in (0, 1, (0, 1)):
        host = G.operators[k].immigrate()
        ref = O.operators[k]
        npt.assert_array_equal(host.offsets, ref.offsets)
        npt.assert_allclose(host.data, ref.data, rtol=1e-14, atol=0)
        if k != (0, 1):
            npt.assert_allclose(host.R, ref.R, rtol=1e-14, atol=0)


# -- apply / solve ------------------------------------------------------------------------
@pytest.mark.parametrize("k", [0, 1])
def test_apply_solve_match_oracle(fd, k):
    G, O = engine_pair(fd)
    G.set_zero_derivative()
    G.scale_and_combine_operators()
    rng = np.random.default_rng(1)
    V = rng.random(G.shape) * 10
    L = G.operators[k]
    Lo = O.operators[k]
    want_apply = Lo.apply(V)
    # FMA contraction in the stencil sum: agreement to a few ulps of the terms
    npt.assert_allclose(L.apply(V), want_apply, rtol=1e-14, atol=1e-14 * np.max(np.abs(want_apply)))
    # solve I - 0.01 A, folded and unfolded
    Li = L.mul_scalar_from_host(-0.01).add(1, inplace=True)
    Lio = (Lo * -0.01).add(1)
    want = Lio.solve(V)
    npt.assert_allclose(Li.solve(V), want, rtol=1e-12, atol=1e-12)
    if Li.is_foldable():
        Li.diagonalize()
        assert Li.is_folded()
        npt.assert_allclose(Li.solve(V), want, rtol=1e-12, atol=1e-12)
        npt.assert_allclose(Li.apply(V), Lio.apply(V), rtol=1e-12, atol=1e-12)
        Li.undiagonalize()
        assert Li.is_foldable() and not Li.is_folded()


def test_cross_apply_matches_oracle(fd):
    G, O = engine_pair(fd, rho=-0.7)
    G.set_zero_derivative()
    G.scale_and_combine_operators()
    V = np.random.default_rng(2).random(G.shape)
    npt.assert_array_almost_equal(G.operators[(0, 1)].apply(V), O.operators[(0, 1)].apply(V),
                                  decimal=13)


def test_emigrate_immigrate_roundtrip(fd):
    import fd_adi_amd.BandedOperatorGPU as BOG
    G, O = engine_pair(fd)
    for k in (0, 1, (0, 1)):
        ref = O.operators[k]
        B = BOG.BandedOperator(ref)
        back = B.immigrate()
        npt.assert_array_equal(back.D.toarray(), ref.D.toarray())
        if k != (0, 1):
            npt.assert_array_equal(back.R, ref.R)
            assert back.bottom_fold_status == ref.bottom_fold_status
        V = np.random.default_rng(3).random(G.shape)
        npt.assert_array_almost_equal(B.apply(V), ref.apply(V), decimal=12)


def test_add_and_errors(fd):
    import fd_adi_amd.BandedOperatorGPU as BOG
    G, O = engine_pair(fd)
    a = BOG.for_vector(G.spots, G.vars.size, 1, 0, False)
    b = BOG.for_vector(G.spots, G.vars.size, 2, 0, False)
    c = a + b
    ref = O.simple_operators[(0,)].add(O.simple_operators[(0, 0)])
    host = c.immigrate()
    npt.assert_array_equal(host.data, ref.data)
    npt.assert_array_equal(host.R, ref.R)
    v1 = BOG.for_vector(G.vars, G.spots.size, 2, 1, False)
    with pytest.raises(ValueError):
        a.add_operator(v1)  # different axis
    v1.diagonalize()
    v2 = BOG.for_vector(G.vars, G.spots.size, 1, 1, False)
    with pytest.raises(NotImplementedError):
        v2.add_operator(v1)  # folded operand
    m = BOG.mixed_for_vector(G.spots, G.vars)
    with pytest.raises(NotImplementedError):
        m.add_operator(a)


def test_derivative_of_exp_and_its_integral(fd):
    """test_BandedOperatorGPU.py:275-289: d/dx e^x by apply, rtol 1e-4."""
    import fd_adi_amd.BandedOperatorGPU as BOG
    x = np.linspace(0, 1, 500)
    B = OB.for_vector(x)
    B.data[1, :] += 1
    B.data[1, 0] = B.data[1, -1] = 2
    BG = BOG.BandedOperator(B)
    ref = np.e ** x
    tst = BG.apply(ref) / 2
    npt.assert_allclose(tst, ref, rtol=1e-4, atol=1e-6)
    tst = BG.solve(ref) * 2
    npt.assert_allclose(ref, tst, rtol=1e-4, atol=1e-6)


def test_coefficient_fields_bitwise(fd):
    import ctypes
    from fd_adi_amd import _lib
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    G, O = engine_pair(fd, rho=-0.6)
    p = np.array([0.03, 1.0, 0.12, 0.3, -0.6])
    out = SizedArrayPtr().alloc(G.spots.size * G.vars.size)
    m0, m1 = SizedArrayPtr(G.spots), SizedArrayPtr(G.vars)
    for which, d, dim in [(0, (0,), 0), (1, (0, 0), 0), (2, (1,), 1), (3, (1, 1), 1),
                          (4, (0, 1), 1)]:
        _lib.call("fdadi_heston_coefficients", which, m0.p, m0.size, m1.p, m1.size,
                  _lib.ptr(p), out.p, _lib.default_stream())
        npt.assert_array_equal(out.to_numpy(), O.coefficient_vector(d, dim))
