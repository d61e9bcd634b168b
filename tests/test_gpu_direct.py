"""The FP64 direct sum on the GPU (uvlm_direct, the facade's exafmm::direct) against the oracle.

uvlm_direct is the yardstick of every FMM-vs-direct error this repository reports, so it is
pinned on its own: u and J against the oracle's device-formula direct sum
(multipoleKernels.hpp:223-301 over all pairs), and the classic stretching J.alpha_target
against the oracle's restatement of the reference's host formula (multipoleKernels.hpp:174-221,
the loop of exafmm::direct, traverse.hpp:88-98). Targets include points that coincide with
sources (those pairs are skipped, multipoleKernels.hpp:242). Bar: 1e-12 relative L2.
"""
import numpy as np
import pytest

from uvlm_amd import FMMCalculator, workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("case", ["blobs", "cube"])
def test_direct_matches_oracle(orc, case):
    n, m = 50000, 2000
    if case == "blobs":
        X, A, S, _, _, _ = W.blobs(n=n, seed=1004, n_probe=3, n_check=1)
    else:
        X, A, S = W.scaled("cube", n)
    rng = np.random.default_rng(29)
    xt = X[rng.choice(n, m, replace=False)].copy()
    xt[: m // 2] += rng.normal(scale=0.02, size=(m // 2, 3))  # the other half sit on sources
    at = rng.normal(size=(m, 3))
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    u, J, D = g.direct(xt, at)
    ru, rJ = orc.direct(xt, X, A, S)
    cu, cD = orc.direct_classic(xt, at, X, A, S)
    eu, eJ, ecu, eD = rel(u, ru), rel(J, rJ), rel(u, cu), rel(D, cD)
    print(f"[direct/{case}] {m} targets x {n} sources: u {eu:.2e}  J {eJ:.2e}  u(host formula) {ecu:.2e}"
          f"  classic dadt {eD:.2e}", flush=True)
    assert eu < 1e-12 and eJ < 1e-12 and ecu < 1e-12 and eD < 1e-12
    # the classic stretching is J.alpha_target of the same J
    assert rel(D, np.einsum("mab,mb->ma", J.reshape(m, 3, 3), at)) < 1e-13
    g.close()


def test_direct_single_pair_closed_form():
    """One source, targets at known offsets: u = -(g(R/sigma) / (4 pi R^3)) (r x alpha) with
    g = erf(s) - 2 s e^{-s^2} / sqrt(pi), s = R / (sqrt(2) sigma); the coincident target gets 0."""
    from math import erf, exp, pi, sqrt
    xs, a, sg = np.array([[0.1, -0.2, 0.3]]), np.array([[0.3, -1.0, 0.5]]), np.array([0.2])
    g = FMMCalculator(P=4, theta=0.5, ncrit=64)
    g.set_arrays(xs, a, sg)
    offs = np.array([[0, 0, 0], [0.05, 0, 0], [0.1, 0.2, -0.1], [1.0, -2.0, 0.5], [0, 0, 3.0]])
    u, J, _ = g.direct(xs + offs)
    for k, r in enumerate(offs):
        R = np.linalg.norm(r)
        if R == 0:
            assert np.all(u[k] == 0) and np.all(J[k] == 0)
            continue
        s = R / (sqrt(2) * sg[0])
        gg = erf(s) - 2 * s * exp(-s * s) / sqrt(pi)
        want = -gg / (4 * pi * R ** 3) * np.cross(r, a[0])
        assert np.allclose(u[k], want, rtol=1e-13, atol=0), (k, u[k], want)
        assert abs(np.trace(J[k].reshape(3, 3))) < 1e-12 * np.abs(J[k]).max()
    g.close()
