"""Surrogate MLP kernels on the GPU: the MFMA GEMM against NumPy and the
gradients against the NumPy reference (bf16 rounding emulated), at the
training batch of the surrogate among others.  Training on ADI labels,
labels and serving are in tests/test_gpu_surrogate.py."""

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu

from oracle import mlp_ref as R


@pytest.fixture(scope="module")
def fd(gpu):
    import fd_adi_amd
    return fd_adi_amd


def test_gemm_nt_matches_numpy(fd):
    from fd_adi_amd import _lib
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    rng = np.random.default_rng(0)
    M, N, K = 256, 128, 384
    A = rng.normal(0, 1, (M, K))
    B = rng.normal(0, 1, (N, K))
    Ab, Bb = R.bf16_bits(A), R.bf16_bits(B)
    Ar, Br = R.bf16_round(A), R.bf16_round(B)
    dA = SizedArrayPtr(Ab.view(np.float64).copy())
    dB = SizedArrayPtr(Bb.view(np.float64).copy())
    C = SizedArrayPtr().alloc(M * N // 2)  # fp32 output in a double allocation
    _lib.call("fdadi_mlp_gemm_nt", dA.p, dB.p, M, N, K, C.p, _lib.default_stream())
    got = C.to_numpy().view(np.float32).reshape(M, N).astype(np.float64)
    want = Ar.dot(Br.T)
    npt.assert_allclose(got, want, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("B,nin", [(256, 8), (2048, 8), (8192, 8), (1000, 13)])
def test_gradients_match_reference(fd, B, nin):
    """B = 2048 exercises the split-K weight-gradient GEMMs (4 splits); 8192
    is the training batch of the surrogate, 16 splits of 512 rows; 1000 rows
    of 13 features (the error model's input) pad the batch to 1024 rows and
    fill a second 8-column chunk of the input."""
    from fd_adi_amd.surrogate import MLP, init_params
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    rng = np.random.default_rng(1)
    lo, hi = -np.ones(nin), 2 * np.ones(nin)
    p = init_params(3, nin)
    P = rng.uniform(-1, 2, (B, nin))
    t = np.sin(P.sum(axis=1))
    m = MLP(lo, hi, 0.0, 1.0, params=p, max_rows=(B + 127) // 128 * 128)
    loss = m.grad(SizedArrayPtr(P), SizedArrayPtr(t), None, B)
    g = m.grads().astype(np.float64)
    X = R.features(P, lo, hi)
    ref_loss, ref = R.loss_and_grad(p.astype(np.float64), X, t, bf16=True)
    npt.assert_allclose(loss, ref_loss, rtol=1e-3)
    offs, o4, ob4, n = R.layout()
    blocks = [(ow, ob) for ow, ob in offs] + [(ob, ob + R.N_H) for _, ob in offs] + [(o4, ob4)]
    for a, b in blocks:
        rel = np.linalg.norm(g[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        assert rel < 2e-2, (a, b, rel)
