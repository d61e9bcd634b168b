"""The vortex-lattice wing (uvlm_wing_*) and particle appending on the GPU against the oracle's
restatement of the reference's loop (oracle/vlm.py, main.cpp:168-295) on the same inputs.

Bars: geometry 1e-14; circulations and force of one solve 1e-11 (LU with partial pivoting vs
LAPACK); the time loop (solve, shed, RK2 advance of the wake, wing moved by -U dt) 1e-10 per
step while every particle sits in one leaf (the FMM is then the direct sum on both sides), 1e-8
on a multi-level tree (400 panels); appending = setting the concatenated set, bit for bit.
"""
import numpy as np
import pytest

from uvlm_amd import FMMCalculator, make_bodies
from uvlm_amd.vlm import Wing, reference_wing

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _tapered():
    V = reference_wing(npanels=12, chord=0.5)
    V = np.stack([V[0], V[0] + 0.45 * (V[1] - V[0]), V[1]])
    V[:, :, 1] *= (1.0 - 0.3 * np.linspace(0, 1, V.shape[1]))[None, :]
    return V


@pytest.mark.parametrize("which", ["reference40", "tapered12x2"])
def test_wing_geometry_and_solve(orc, which):
    from oracle import vlm
    V = reference_wing(npanels=40) if which == "reference40" else _tapered()
    U = np.array([0.0, 1.0, 0.0])
    w, ow = Wing(V), vlm.Wing(V)
    cp, nrm, area, _ = w.geometry()
    assert np.abs(cp - ow.cp).max() < 1e-14 and np.abs(nrm - ow.normal).max() < 1e-14
    assert np.abs(area - ow.area).max() < 1e-14 * ow.area.max()
    g, F = w.solve(None, U)
    eg = vlm.solve(ow, U, np.zeros((ow.n, 3)))
    eF = vlm.force(ow, eg, U, np.zeros((ow.n, 3)))
    print(f"[{which}] gamma {rel(g, eg):.3e} force {rel(F, eF):.3e}", flush=True)
    assert rel(g, eg) < 1e-11 and rel(F, eF) < 1e-11


def test_wing_rejects_a_control_point_on_a_vortex_line():
    V = reference_wing(npanels=4)
    V[1, 2] = V[0, 2]  # a degenerate panel: zero chord at one station
    with pytest.raises(Exception):
        Wing(V)


def test_append_equals_setting_the_concatenation():
    rng = np.random.default_rng(4)
    X, A, S = rng.normal(size=(900, 3)), rng.normal(size=(900, 3)), np.full(900, 0.2)
    g1 = FMMCalculator(P=6, theta=0.5, ncrit=64)
    g1.set_arrays(X[:500], A[:500], S[:500])
    g1.append_arrays(X[500:], A[500:], S[500:])
    g1.build()
    g1.calculate()
    g2 = FMMCalculator(P=6, theta=0.5, ncrit=64)
    g2.set_arrays(X, A, S)
    g2.build()
    g2.calculate()
    a1, a2 = g1.arrays(), g2.arrays()
    assert np.array_equal(g1.permutation(), g2.permutation())
    for k in ("X", "alpha", "velocity", "J", "dadt"):
        assert np.array_equal(a1[k], a2[k]), k


@pytest.mark.parametrize("reference_offset", [True, False])
def test_wing_loop_matches_oracle_one_leaf(orc, reference_offset):
    from oracle import vlm
    V = reference_wing(npanels=6, chord=0.4)
    V = np.stack([V[0], (V[0] + V[1]) / 2, V[1]])
    U, dt, steps = np.array([0.0, 1.0, 0.0]), 0.05, 8
    hist, bodies = vlm.run(V, U, dt, steps, P=6, ncrit=256, scheme="rk2", reference_offset=reference_offset)
    w = Wing(V)
    p = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2")
    for t in range(steps):
        g, F = w.step(p, U, dt, reference_offset)
        eg, eF = hist[t]
        assert rel(g, eg) < 1e-10 and rel(F, eF) < 1e-10, (t, rel(g, eg), rel(F, eF))
    arr = p.arrays()
    assert np.array_equal(p.permutation(), np.arange(p.n))
    assert rel(arr["X"], bodies[:, 9:12]) < 1e-10 and rel(arr["alpha"], bodies[:, 12:15]) < 1e-10
    print(f"[loop one leaf, offset {reference_offset}] n={p.n} last force {F}", flush=True)


def test_wing_loop_400_panels_matches_oracle(orc):
    """40 x 10 panels (the configuration's lattice), 5 steps, multi-level particle trees."""
    from oracle import vlm
    V = reference_wing(npanels=40, span=4.0, chord=1.0, angle_deg=5.0, nchord=10)
    U, dt, steps = np.array([0.0, 1.0, 0.0]), 0.02, 5
    hist, _ = vlm.run(V, U, dt, steps, P=8, theta=0.5, ncrit=16, scheme="rk2")
    w = Wing(V)
    p = FMMCalculator(P=8, theta=0.5, ncrit=16, scheme="rk2")
    for t in range(steps):
        g, F = w.step(p, U, dt)
        eg, eF = hist[t]
        print(f"  step {t}: gamma {rel(g, eg):.3e} force {rel(F, eF):.3e} leaves {p.timings()['n_leaves']}", flush=True)
        assert rel(g, eg) < 1e-8 and rel(F, eF) < 1e-8


def test_lift_adjoint_finite_differences():
    """dC_L(final) / d(wake state at the window start) through 4 loop steps (solve, shed, RK2
    advance), against central differences of the same loop run from perturbed states (every
    particle in one leaf, so the FMM is the direct sum and the loop is smooth)."""
    V = reference_wing(npanels=6, chord=0.4)
    V = np.stack([V[0], (V[0] + V[1]) / 2, V[1]])
    U, dt, warm, win = np.array([0.0, 1.0, 0.0]), 0.05, 3, 4
    w = Wing(V)
    p = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2")
    for _ in range(warm):
        w.step(p, U, dt)
    arr, perm = p.arrays(), p.permutation()
    n = p.n
    X0, A0, S0 = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    X0[perm], A0[perm], S0[perm] = arr["X"], arr["alpha"], arr["sigma"]
    area = w.geometry()[2].sum()
    Vstart = V - warm * dt * U

    def objective(X, A):
        w2 = Wing(Vstart)
        p2 = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2")
        p2.set_arrays(X, A, S0)
        for _ in range(win):
            w2.step(p2, U, dt)
        _, F = w2.solve(p2, U)
        return F[2] / (0.5 * U @ U * area)

    CL, xb, ab = w.lift_adjoint(p, U, dt, win)
    assert abs(CL - objective(X0, A0)) < 1e-12 * abs(CL)
    h = 1e-6
    fd, an = [], []
    for j in (0, 5, n // 2, n - 1):
        for d in range(3):
            for which in ("x", "a"):
                Xp, Ap, Xm, Am = X0.copy(), A0.copy(), X0.copy(), A0.copy()
                (Xp if which == "x" else Ap)[j, d] += h
                (Xm if which == "x" else Am)[j, d] -= h
                fd.append((objective(Xp, Ap) - objective(Xm, Am)) / (2 * h))
                an.append((xb if which == "x" else ab)[j, d])
    fd, an = np.array(fd), np.array(an)
    print(f"[lift adjoint] CL {CL:.6f} n={n} vs FD {rel(an, fd):.3e} |grad| {np.abs(fd).max():.3e}", flush=True)
    assert rel(an, fd) < 1e-6


def test_cpp_wing_facade_runs_the_reference_loop(orc, tmp_path):
    """tests/cpp/wing_demo.cpp: the driver's loop written against include/uvlm/vlm_wing.hpp and
    the FMMCalculator facade (RK2, P = 6, ncrit 256) gives the oracle's forces and wake."""
    import os
    import subprocess
    from oracle import vlm
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "uvlm_amd", "lib",
                       "uvlm_wing_demo")
    out = tmp_path / "wing.bin"
    steps, npanels = 6, 6
    r = subprocess.run([exe, str(npanels), str(steps), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    forces = raw[:3 * steps].reshape(steps, 3)
    bodies = raw[3 * steps:].reshape(-1, 37)
    V = reference_wing(npanels=npanels, chord=0.4)
    hist, eb = vlm.run(V, [0.0, 1.0, 0.0], 0.05, steps, P=6, ncrit=256, scheme="rk2")
    assert rel(forces, np.array([h[1] for h in hist])) < 1e-10
    assert bodies.shape == eb.shape and rel(bodies[:, 9:15], eb[:, 9:15]) < 1e-10


def test_wing_single_panel_and_empty_window(orc):
    """Edge cases: one spanwise panel (both tip terms on the same shed particle) against the
    oracle's solve and shedding, and a lift adjoint window started with no particles (no gradient
    entries, CL as the loop gives it).  A single particle is not run through the oracle's FMM:
    it sits on its own cell centre, where the reference's cart2sph divides by zero."""
    from oracle import vlm
    U, dt = np.array([0.0, 1.0, 0.0]), 0.05
    V1 = reference_wing(npanels=1, chord=0.5)
    w, ow = Wing(V1), vlm.Wing(V1)
    g, F = w.solve(None, U)
    eg = vlm.solve(ow, U, np.zeros((1, 3)))
    assert rel(g, eg) < 1e-14 and rel(F, vlm.force(ow, eg, U, np.zeros((1, 3)))) < 1e-14
    p = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2")
    w.shed(p, U, dt)
    X, A, S = vlm.shed(ow, eg, U, dt)
    arr = p.arrays()
    assert p.n == 1 and rel(arr["X"], X) < 1e-15 and rel(arr["alpha"], A) < 1e-14 and rel(arr["sigma"], S) < 1e-15
    # an even panel count: no particle lands on the root centre of the oracle's trees
    V = reference_wing(npanels=4, chord=0.5)
    hist, _ = vlm.run(V, U, dt, 4, P=6, ncrit=256, scheme="rk2")
    w2 = Wing(V)
    p2 = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2")
    CL, xb, ab = w2.lift_adjoint(p2, U, dt, 3)
    area = w2.geometry()[2].sum()
    assert xb.shape == (0, 3) and ab.shape == (0, 3) and p2.n == 12
    assert abs(CL - hist[3][1][2] / (0.5 * area)) < 1e-11 * abs(CL)


def test_lift_adjoint_mixed_matches_fp64():
    """The lift adjoint with FP32 pair arithmetic against the FP64 one (same loop, one leaf)."""
    V = reference_wing(npanels=6, chord=0.4)
    V = np.stack([V[0], (V[0] + V[1]) / 2, V[1]])
    U, dt = np.array([0.0, 1.0, 0.0]), 0.05
    out = {}
    for prec in ("fp64", "mixed"):
        w = Wing(V)
        p = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2", precision=prec)
        for _ in range(3):
            w.step(p, U, dt)
        out[prec] = w.lift_adjoint(p, U, dt, 4)
    CLd, xd, ad = out["fp64"]
    CLm, xm, am = out["mixed"]
    print(f"[lift adjoint mixed] CL {abs(CLm - CLd) / abs(CLd):.2e} x {rel(xm, xd):.2e} a {rel(am, ad):.2e}", flush=True)
    assert abs(CLm - CLd) < 1e-6 * abs(CLd) and rel(xm, xd) < 1e-5 and rel(am, ad) < 1e-5


def test_wing_rejects_bad_inputs():
    with pytest.raises(Exception):
        Wing(np.zeros((2, 1, 3)))  # no panel
    V = reference_wing(npanels=6)
    w = Wing(V)
    p = FMMCalculator(P=6, theta=0.5, ncrit=64, formulation="rvpm")
    w.step(p, [0.0, 1.0, 0.0], 0.05)  # the forward loop runs with rVPM ...
    with pytest.raises(Exception):
        w.lift_adjoint(p, [0.0, 1.0, 0.0], 0.05, 2)  # ... its adjoint covers cVPM only


@pytest.mark.parametrize("ncrit,every,twisted", [(1024, 2, False), (1024, 1, True), (1024, 4, True)])
def test_wing_design_gradient_finite_differences(ncrit, every, twisted):
    """dC_L/d(angle of attack) and dC_L/d(strip twist) through a window of the loop
    (uvlm_wing_window_adjoint, checkpoints every `every` steps, the wake at the window start held
    fixed) against central differences of the loop run with a new wing built from the perturbed
    design (design_wing).  ncrit 1024: every particle in one leaf (the FMM is the direct sum);
    ncrit 8: multi-level trees at P = 12."""
    from uvlm_amd.vlm import design_gradient, design_wing
    alpha, nspan, nchord, span = np.deg2rad(5.0), 6, 2, 3.0
    twist = 0.04 * np.sin(np.arange(nspan)) if twisted else None
    U, dt, warm, win = np.array([0.0, 1.0, 0.0]), 0.05, 4, 6
    P = 6 if ncrit > 100 else 12
    w = Wing(design_wing(alpha, twist, span=span, nspan=nspan, nchord=nchord))
    p = FMMCalculator(P=P, theta=0.5, ncrit=ncrit, scheme="rk2")
    for _ in range(warm):
        w.step(p, U, dt)
    arr, perm = p.arrays(), p.permutation()
    n = p.n
    X0, A0, S0 = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    X0[perm], A0[perm], S0[perm] = arr["X"], arr["alpha"], arr["sigma"]
    shift = warm * dt * U

    def run(a, tw, adjoint=False):
        w2 = Wing(design_wing(a, tw, span=span, nspan=nspan, nchord=nchord) - shift)
        p2 = FMMCalculator(P=P, theta=0.5, ncrit=ncrit, scheme="rk2")
        p2.set_arrays(X0, A0, S0)
        if adjoint:
            return w2.window_adjoint(p2, U, dt, win, every=every)
        for _ in range(win):
            w2.step(p2, U, dt)
        _, F = w2.solve(p2, U)
        return F[2] / (0.5 * U @ U * w2.geometry()[2].sum())

    CL, xb, ab, vb = run(alpha, twist, adjoint=True)
    CL0 = run(alpha, twist)
    assert abs(CL - CL0) < 1e-12 * abs(CL)
    ga, gt = design_gradient(vb, alpha, twist)
    h = 1e-6
    tw0 = np.zeros(nspan) if twist is None else twist
    fa = (run(alpha + h, twist) - run(alpha - h, twist)) / (2 * h)
    an, fd = [ga], [fa]
    for c in (0, 2, nspan - 1):
        e = np.zeros(nspan)
        e[c] = h
        an.append(gt[c])
        fd.append((run(alpha, tw0 + e) - run(alpha, tw0 - e)) / (2 * h))
    an, fd = np.array(an), np.array(fd)
    print(f"[design ncrit {ncrit} every {every} twisted {twisted}] CL {CL:.6f} dCL/dalpha {ga:.6e} (FD {fa:.6e}); "
          f"twist gradient {gt}; vs FD {rel(an, fd):.3e}", flush=True)
    assert rel(an, fd) < 1e-6


def test_cpp_wing_design_gradient(tmp_path):
    """tests/cpp/wing_demo.cpp design: uvlm::WingDesign and Wing::windowAdjoint from C++ give the
    design gradient that the same window gives through the Python binding (design_gradient)."""
    import os
    import subprocess
    from uvlm_amd.vlm import design_gradient, design_wing
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "uvlm_amd", "lib",
                       "uvlm_wing_demo")
    out = tmp_path / "design.bin"
    r = subprocess.run([exe, "design", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got = np.fromfile(out)
    alpha, U, dt = np.deg2rad(5.0), np.array([0.0, 1.0, 0.0]), 0.05
    w = Wing(design_wing(alpha, None, span=3.0, nspan=6, nchord=2))
    p = FMMCalculator(P=6, theta=0.5, ncrit=1024, scheme="rk2")
    for _ in range(4):
        w.step(p, U, dt)
    CL, _, _, vb = w.window_adjoint(p, U, dt, 6, every=2)
    ga, gt = design_gradient(vb, alpha, None)
    want = np.concatenate([[CL, ga], gt])
    print(f"[design C++] {got[:2]} python {want[:2]}", flush=True)
    assert got.shape == want.shape and rel(got, want) < 1e-13


def test_wing_window_adjoint_checkpoints_multilevel():
    """Multi-level trees (ncrit 4, P 12): the window adjoint's forward CL equals the loop's, and
    its gradients (particles and vertices) do not depend on the checkpoint interval (every step,
    every 3 steps, one segment), which changes only what is recomputed.  Central differences
    through the loop are not used here: a perturbation that moves a pair across the acceptance
    test switches it between the regularised and the singular kernel, a jump the differences
    cannot see past."""
    from uvlm_amd.vlm import design_wing
    alpha, nspan, nchord, span = np.deg2rad(5.0), 6, 2, 3.0
    U, dt, warm, win = np.array([0.0, 1.0, 0.0]), 0.05, 10, 6
    V = design_wing(alpha, None, span=span, nspan=nspan, nchord=nchord)
    w = Wing(V)
    p = FMMCalculator(P=12, theta=0.5, ncrit=4, scheme="rk2")
    for _ in range(warm):
        w.step(p, U, dt)
    arr, perm = p.arrays(), p.permutation()
    n = p.n
    X0, A0, S0 = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    X0[perm], A0[perm], S0[perm] = arr["X"], arr["alpha"], arr["sigma"]
    pc = FMMCalculator(P=12, theta=0.5, ncrit=4, scheme="rk2")
    pc.set_arrays(X0, A0, S0)
    pc.build()
    assert pc.timings()["n_levels"] > 2 and pc.timings()["n_m2l_pairs"] > 0
    outs = {}
    for every in (1, 3, 6):
        w2 = Wing(V - warm * dt * U)
        p2 = FMMCalculator(P=12, theta=0.5, ncrit=4, scheme="rk2")
        p2.set_arrays(X0, A0, S0)
        outs[every] = w2.window_adjoint(p2, U, dt, win, every=every)
    w3 = Wing(V - warm * dt * U)
    p3 = FMMCalculator(P=12, theta=0.5, ncrit=4, scheme="rk2")
    p3.set_arrays(X0, A0, S0)
    for _ in range(win):
        w3.step(p3, U, dt)
    _, F = w3.solve(p3, U)
    CL = F[2] / (0.5 * U @ U * w3.geometry()[2].sum())
    for every, (cl, xb, ab, vb) in outs.items():
        ex, ea, ev = rel(xb, outs[6][1]), rel(ab, outs[6][2]), rel(vb, outs[6][3])
        print(f"[window every {every}] CL {cl:.9f} (loop {CL:.9f}) vs one segment: xbar {ex:.1e} alphabar {ea:.1e} "
              f"vertices {ev:.1e}", flush=True)
        assert abs(cl - CL) < 1e-12 * abs(CL)
        assert ex < 1e-9 and ea < 1e-9 and ev < 1e-9
