"""Viscous schemes, Gaussian vorticity and RBF re-initialisation on the GPU against the oracle
(fmmCalculator.hpp:328-337, 713-917; multipoleKernels.hpp:343-450) on the same inputs."""
import numpy as np
import pytest

from uvlm_amd import FMMCalculator, make_bodies, workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _lattice(L, h):
    g = (np.arange(-L, L) + 0.5) * h
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


@pytest.mark.parametrize("scheme", ["pse", "csm"])
def test_viscous_rates_match_oracle(orc, scheme):
    X, A, S = W.scaled("ring", 8192)
    nu, dt = 0.005, 0.01
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=64)
    c.init(b_ref)
    c.calculate()
    c.set_viscosity(scheme, nu)
    c.viscosity(dt)
    ref = c.bodies(len(X))
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    b = make_bodies(X, A, S)
    g.init(b)
    g.calculate()
    g.set_viscosity(scheme, nu)
    g.viscosity(dt)
    out = g.copyBodies2Host()
    assert np.array_equal(out[:, 9:12], ref[:, 9:12])
    assert rel(out[:, 18:21], ref[:, 18:21]) < 1e-10
    assert rel(out[:, 34], ref[:, 34]) < 1e-12 or np.allclose(ref[:, 34], 0)


@pytest.mark.parametrize("scheme", ["pse", "csm"])
def test_advance_applies_the_viscous_scheme(orc, scheme):
    """One Euler step (a single rebuild from the same input) with the scheme set on both sides."""
    X, A, S = W.scaled("cube", 8000)
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=64)
    c.set_viscosity(scheme, 0.01)
    c.advance(b_ref, 0.01, "euler")
    g = FMMCalculator(P=8, theta=0.5, ncrit=64, scheme="euler")
    g.set_viscosity(scheme, 0.01)
    b = make_bodies(X, A, S)
    g.advance(b, 0.01)
    for sl in (slice(9, 12), slice(12, 15), 33):
        assert rel(b[:, sl], b_ref[:, sl]) < 1e-10, sl


def test_vorticity_matches_oracle(orc):
    X, A, S, probes, v, _ = W.blobs(n=20000, n_probe=2048, n_check=1)
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=64)
    c.init(b_ref)
    c.calculate()
    s_ref = orc.make_bodies(probes, np.zeros_like(probes), np.full(len(probes), 0.05))
    s_gpu = s_ref.copy()
    c.vorticity(s_ref)
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    g.build()
    g.vorticity(s_gpu)
    assert np.array_equal(s_gpu[:, 9:12], s_ref[:, 9:12])
    assert rel(s_gpu[:, 21:24], s_ref[:, 21:24]) < 1e-10


@pytest.mark.parametrize("case", ["own_points", "shifted_lattice"])
def test_rbf_matches_oracle(orc, case):
    """RBF onto the particles' own points (recovers their strengths) and onto a shifted, wider
    lattice whose outer points carry no vorticity and are dropped."""
    h = 0.1
    X = _lattice(4, h) + np.array([0.004, -0.003, 0.002])
    rng = np.random.default_rng(3)
    A = (1.0 + 0.3 * rng.normal(size=X.shape)) * h ** 3
    S = np.full(len(X), 0.5 * h)
    if case == "own_points":
        P = X.copy()
    else:
        P = _lattice(6, h) + np.array([0.05, 0.04, -0.03])
    pts_ref = orc.make_bodies(P, np.zeros_like(P), np.full(len(P), 0.5 * h))
    pts_gpu = pts_ref.copy()
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=6, theta=0.5, ncrit=64)
    c.init(b_ref)
    c.calculate()
    k_ref, it_ref, res_ref = c.rbf(pts_ref)
    g = FMMCalculator(P=6, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    g.build()
    k, it, res = g.rbf(pts_gpu)
    print(case, "kept", k, k_ref, "iterations", it, it_ref, "residuals", res, res_ref)
    assert k == k_ref and it == it_ref and 0 < k <= len(P)
    if case == "shifted_lattice":
        assert k < len(P)
    assert np.array_equal(pts_gpu[:k, 9:12], pts_ref[:k, 9:12])
    assert rel(pts_gpu[:k, 12:15], pts_ref[:k, 12:15]) < 1e-8
    if case == "own_points":
        arr = g.arrays()
        assert rel(pts_gpu[:k, 12:15], arr["alpha"]) < 1e-5
