"""uvlm_direct_noalloc (static device staging, no allocation) against uvlm_direct.

Both run the same kernel on the same staged values, so the results must be identical bit for
bit: two calls in a row, the second with fewer targets (the static array keeps the first call's
contents) and without the J output, plus the size limit.
"""
import numpy as np
import pytest

from uvlm_amd import FMMCalculator, workloads as W
from uvlm_amd._capi import DIRECT_NOALLOC_MAX, UVLMError

pytestmark = pytest.mark.gpu


def test_direct_noalloc_matches_direct():
    X, A, S, _, _, _ = W.blobs(n=20000, seed=11, n_probe=3, n_check=1)
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    rng = np.random.default_rng(17)
    for m, want_J in ((5000, True), (1200, False)):
        xt = X[rng.choice(X.shape[0], m, replace=False)] + rng.normal(scale=0.01, size=(m, 3))
        xt[:50] = X[:50]  # coincident with sources: skipped pairs
        at = rng.normal(size=(m, 3))
        print(f"[noalloc] m={m} want_J={want_J}", flush=True)
        u0, J0, D0 = g.direct(xt, at, want_J=want_J)
        u1, J1, D1 = g.direct(xt, at, want_J=want_J, noalloc=True)
        assert np.array_equal(u0, u1) and np.array_equal(D0, D1)
        assert (J0 is None and J1 is None) or np.array_equal(J0, J1)
        assert np.all(np.isfinite(u1)) and np.linalg.norm(u1) > 0
    with pytest.raises(UVLMError):
        g.direct(np.zeros((DIRECT_NOALLOC_MAX + 1, 3)), noalloc=True)
    # the largest accepted size still works after the rejected call
    m = DIRECT_NOALLOC_MAX
    xt = rng.uniform(-1, 1, size=(m, 3))
    u0, _, _ = g.direct(xt, want_J=False)
    u1, _, _ = g.direct(xt, want_J=False, noalloc=True)
    assert np.array_equal(u0, u1)
    g.close()
