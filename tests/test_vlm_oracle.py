"""Pins the oracle's restatement of the reference's vortex-lattice wing (oracle/vlm.py).

  segment kernel  -> the five known-answer cases of testVelocity (main.cpp:82-113, printed to
                     3 decimals) and Gauss-Legendre quadrature of the Biot-Savart integral
  system matrix   -> symmetric on lattices of congruent parallelograms (the reference's wing),
                     so the reference's transposed fill (AIC(j, i)) gives the same circulations
  solve           -> zero normal velocity at every control point
  force           -> Kutta-Joukowski sum evaluated segment by segment
  shedding        -> closed form for uniform circulation (only the tip side segments shed)
"""
import numpy as np
import pytest

from oracle import vlm


def test_segment_known_answers():
    # main.cpp:82-113: line, target point, printed expectation
    cases = [((0, 0, 0), (0, 0, 1), (0, 0.5, 0.5), (-0.225, 0, 0)),
             ((1, 0, 0), (0, 1, 0), (0, 0, 1), (0.038, 0.038, 0.038)),
             ((1, 0, 0), (0, 1, 0), (0.5, 0.5, 0.5), (0.184, 0.184, 0.000)),
             ((1, 0, 1), (0, 0, 1), (0.5, 0.5, 0.5), (0.000, -0.092, -0.092)),
             ((1, 0, 1), (0, 0, 1), (1.5, 1.5, 1.5), (0.000, 0.006, -0.018))]
    for a, b, p, e in cases:
        u = vlm.segment_velocity(a, b, p)
        assert np.all(np.abs(u - np.array(e)) <= 5e-4 + 1e-12), (a, b, p, u, e)
    with pytest.raises(ValueError):
        vlm.segment_velocity((0, 0, 0), (0, 0, 1), (0, 0, 2))  # on the line: the reference exits


def test_segment_matches_quadrature():
    rng = np.random.default_rng(1)
    x, w = np.polynomial.legendre.leggauss(200)
    for _ in range(20):
        a, b, p = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3) * 2
        t = (x + 1) / 2
        pts = a[None, :] + t[:, None] * (b - a)[None, :]
        r = p[None, :] - pts
        dl = (b - a)[None, :]
        integrand = np.cross(dl, r) / np.linalg.norm(r, axis=1)[:, None] ** 3
        u = (w / 2) @ integrand / (4 * np.pi)
        assert np.allclose(vlm.segment_velocity(a, b, p), u, rtol=1e-9, atol=1e-12)


def test_system_matrix_symmetric_on_the_reference_wing():
    V = vlm.reference_wing(npanels=40)
    A = vlm.Wing(V).aic()
    assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()
    V2 = np.stack([V[0], (V[0] + V[1]) / 2, V[1]])
    A2 = vlm.Wing(V2).aic()
    assert np.abs(A2 - A2.T).max() < 1e-13 * np.abs(A2).max()
    # the reference's fill solves A^T gamma = rhs: same circulations here
    rhs = np.random.default_rng(0).normal(size=A.shape[0])
    assert np.allclose(np.linalg.solve(A, rhs), np.linalg.solve(A.T, rhs), rtol=1e-10, atol=1e-12)


def _tapered():
    V = vlm.reference_wing(npanels=12, chord=0.5)
    V = np.stack([V[0], V[0] + 0.45 * (V[1] - V[0]), V[1]])
    V[:, :, 1] *= (1.0 - 0.3 * np.linspace(0, 1, V.shape[1]))[None, :]  # taper: not congruent
    return V


def test_solve_gives_zero_normal_velocity():
    U = np.array([0.0, 1.0, 0.0])
    for V in (vlm.reference_wing(npanels=30), _tapered()):
        w = vlm.Wing(V)
        v = np.random.default_rng(2).normal(size=(w.n, 3)) * 0.1
        g = vlm.solve(w, U, v)
        u = np.einsum("ijd,j->id", w.influence(w.cp), g) + U[None, :] + v
        assert np.abs(np.sum(u * w.normal, axis=1)).max() < 1e-12


def test_force_is_the_segment_by_segment_sum():
    w = vlm.Wing(_tapered())
    rng = np.random.default_rng(3)
    g, U, v = rng.normal(size=w.n), np.array([0.1, 1.0, 0.05]), rng.normal(size=(w.n, 3)) * 0.01
    F = np.zeros(3)
    s, e = w.lines()
    for i in range(w.n):
        u = U + v[i]
        for j in range(w.n):
            for k in range(4):
                u = u + g[j] * vlm.segment_velocity(s[k][j], e[k][j], w.cp[i])
        up = g[i - w.nspan] if i >= w.nspan else 0.0
        F += (g[i] - up) * np.cross(u, w.LE1[i] - w.LE[i])
    assert np.allclose(vlm.force(w, g, U, v), F, rtol=1e-12, atol=1e-14)


def test_shedding_closed_form_for_uniform_circulation():
    V = vlm.reference_wing(npanels=8)
    w = vlm.Wing(V)
    U, dt, g0 = np.array([0.0, 1.0, 0.0]), 0.05, 0.7
    X, A, S = vlm.shed(w, np.full(w.n, g0), U, dt)
    te, te1, le, le1 = w.TE, w.TE1, w.LE, w.LE1
    want = (te - te1) * g0
    want[0] += (le[0] - te[0]) * g0
    want[-1] += (te1[-1] - le1[-1]) * g0
    assert np.allclose(A, want, rtol=0, atol=1e-15)
    assert np.allclose(S, 2.5 * np.linalg.norm(te - te1, axis=1))
    # main.cpp:252-254: x, y offset by U dt, z by U[1] dt as written
    assert np.allclose(X - (te + te1) / 2, [0.0, dt, dt])
    X2, _, _ = vlm.shed(w, np.full(w.n, g0), U, dt, reference_offset=False)
    assert np.allclose(X2 - (te + te1) / 2, [0.0, dt, 0.0])


def test_design_gradient_chain_rule():
    """design_gradient maps a vertex cotangent to (dCL/dalpha, dCL/dtwist): against central
    differences of sum(vbar . V(alpha, twist)) for a random vbar, at a twisted design."""
    from uvlm_amd.vlm import design_gradient, design_wing
    rng = np.random.default_rng(4)
    nspan, nchord = 7, 3
    vb = rng.normal(size=(nchord + 1, nspan + 1, 3))
    alpha, tw = 0.09, 0.05 * rng.normal(size=nspan)
    f = lambda a, t: float(np.sum(vb * design_wing(a, t, span=4.0, chord=1.3, nspan=nspan, nchord=nchord)))
    ga, gt = design_gradient(vb, alpha, tw, chord=1.3)
    h = 1e-6
    assert abs(ga - (f(alpha + h, tw) - f(alpha - h, tw)) / (2 * h)) < 1e-8 * max(1, abs(ga))
    for c in range(nspan):
        e = np.zeros(nspan)
        e[c] = h
        assert abs(gt[c] - (f(alpha, tw + e) - f(alpha, tw - e)) / (2 * h)) < 1e-8 * max(1, abs(gt[c]))
