"""fdadi_mlp_predict_noalloc (activations in a static device array) against
fdadi_mlp_predict (activations in the model's buffers): bit-identical
prices on two calls in a row, the second smaller; sizes beyond the maximum
are refused."""

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(m, P, out):
    from fd_adi_amd import _lib
    _lib.call("fdadi_mlp_predict_noalloc", m._h, P.p, P.size // 8, out.p, m.stream)


def test_noalloc_predict_matches_predict(gpu):
    from fd_adi_amd import surrogate as S
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    m = S.MLP(S.BOX_LO, S.BOX_HI, 1.5, 2.0, seed=3, max_rows=1024)
    Pt = S.draw_params(2048, 9)
    m.fit(Pt, np.sin(Pt.sum(axis=1)), steps=20, batch=1024)
    for n, seed in ((4096, 1), (700, 2)):
        P = S.draw_params(n, seed)
        want = m.predict(P)
        dP, out = SizedArrayPtr(P), SizedArrayPtr().alloc(n)
        _noalloc(m, dP, out)
        got = out.to_numpy()
        npt.assert_array_equal(got, want)
        print("n = %d: %d prices equal" % (n, n))
    P = S.draw_params(4097, 3)
    with pytest.raises(ValueError):
        _noalloc(m, SizedArrayPtr(P), SizedArrayPtr().alloc(4097))
