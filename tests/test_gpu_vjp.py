"""Vector-Jacobian products of the rates, of one time step and of a checkpointed trajectory.

  * near field only (one leaf holds every particle): the GPU VJP against the oracle's exact
    direct-sum VJP, itself pinned by central differences, both stretching forms, bar 1e-10;
  * far field included (multi-level tree, P = 8): against central differences (h = 1e-6) of the
    FP64 GPU forward with the tree and lists frozen, on 64 sampled particles, so the expansions'
    own truncation is part of the differentiated function; bar 1e-6;
  * time steps (Euler, RK2, RK3, RK4) and a checkpointed RK2 trajectory on one leaf, against
    central differences of uvlm_advance, bar 1e-7.
"""
import numpy as np
import pytest

from uvlm_amd import FMMCalculator, workloads as W
from uvlm_amd._capi import UVLMError

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _blobs(n, seed=7):
    X, A, S, _, _, _ = W.blobs(n=n, seed=seed, n_probe=3, n_check=1)
    return X, A, S


def _objective(g, ub, ab):
    g.calculate()
    arr = g.arrays()
    return float(np.sum(ub * arr["velocity"]) + np.sum(ab * arr["dadt"]))


@pytest.mark.parametrize("stretching", ["transposed", "classic"])
def test_eval_vjp_near_field_matches_oracle(orc, stretching):
    X, A, S = _blobs(600)
    rng = np.random.default_rng(3)
    ub, ab = rng.normal(size=(600, 3)), rng.normal(size=(600, 3))
    g = FMMCalculator(P=6, theta=0.5, ncrit=1024, stretching=stretching)
    g.set_arrays(X, A, S)
    g.build()
    assert g.timings()["n_leaves"] == 1
    xb, alb = g.eval_vjp(ub, ab)
    arr = g.arrays()
    exb, ealb = orc.direct_vjp(arr["X"], arr["alpha"], arr["sigma"], ub, ab, stretching)
    print(f"[near/{stretching}] xbar {rel(xb, exb):.3e} alphabar {rel(alb, ealb):.3e}", flush=True)
    assert rel(xb, exb) < 1e-10 and rel(alb, ealb) < 1e-10
    # mixed precision: FP32 pair arithmetic (FP64 sums) in the near-field adjoint
    gm = FMMCalculator(P=6, theta=0.5, ncrit=1024, stretching=stretching, precision="mixed")
    gm.set_arrays(X, A, S)
    gm.build()
    mxb, malb = gm.eval_vjp(ub, ab)
    print(f"  mixed: xbar {rel(mxb, exb):.3e} alphabar {rel(malb, ealb):.3e}", flush=True)
    assert rel(mxb, exb) < 1e-5 and rel(malb, ealb) < 1e-5


@pytest.mark.parametrize("stretching", ["transposed", "classic"])
def test_eval_vjp_frozen_tree_finite_differences(orc, stretching):
    n = 6000
    X, A, S = _blobs(n)
    rng = np.random.default_rng(5)
    ub, ab = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    g = FMMCalculator(P=8, theta=0.5, ncrit=64, stretching=stretching)
    g.set_arrays(X, A, S)
    g.build()
    tm = g.timings()
    assert tm["n_m2l_pairs"] > 0 and tm["n_levels"] > 3
    xb, alb = g.eval_vjp(ub, ab)
    print(f"[far/{stretching}] vjp {g.timings()['vjp_ms']:.3f} ms", flush=True)
    arr = g.arrays()
    X0, A0 = arr["X"].copy(), arr["alpha"].copy()
    h = 1e-6
    fd, an = {"x": [], "a": []}, {"x": [], "a": []}
    for j in rng.choice(n, 64, replace=False):
        for d in range(3):
            for which, base, grad in (("x", X0, xb), ("a", A0, alb)):
                p, m = base.copy(), base.copy()
                p[j, d] += h
                m[j, d] -= h
                g.update_arrays(**({"X": p} if which == "x" else {"alpha": p}))
                sp = _objective(g, ub, ab)
                g.update_arrays(**({"X": m} if which == "x" else {"alpha": m}))
                sm = _objective(g, ub, ab)
                g.update_arrays(X=X0, alpha=A0)
                fd[which].append((sp - sm) / (2 * h))
                an[which].append(grad[j, d])
    ex, ea = rel(an["x"], fd["x"]), rel(an["a"], fd["a"])
    print(f"  vs frozen-tree FD: xbar {ex:.3e} alphabar {ea:.3e}", flush=True)
    # the direct-sum VJP differs by the far-field approximation, as the forward rates do
    exb, ealb = orc.direct_vjp(X0, A0, arr["sigma"], ub, ab, stretching)
    u, J = orc.direct(X0, X0, A0, arr["sigma"])
    print(f"  vs direct: xbar {rel(xb, exb):.3e} alphabar {rel(alb, ealb):.3e}; forward u {rel(arr['velocity'], u):.3e}"
          f" J {rel(arr['J'], J):.3e}", flush=True)
    assert ex < 1e-6 and ea < 1e-6
    assert rel(xb, exb) < 0.2 and rel(alb, ealb) < 0.2


@pytest.mark.parametrize("P", [4, 10, 13])
def test_eval_vjp_orders_frozen_tree_fd(P):
    """The rate VJP at other expansion orders: P = 10 runs the fixed-order instances of the adjoint
    L2P kernels, P = 4 and 13 the runtime-order ones.  Against frozen-tree central differences of
    the FP64 forward on 16 sampled particles; bar 1e-6."""
    n = 6000
    X, A, S = _blobs(n, seed=9)
    rng = np.random.default_rng(40 + P)
    ub, ab = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    g = FMMCalculator(P=P, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    g.build()
    assert g.timings()["n_m2l_pairs"] > 0
    xb, alb = g.eval_vjp(ub, ab)
    arr = g.arrays()
    X0, A0 = arr["X"].copy(), arr["alpha"].copy()
    h = 1e-6
    fd, an = {"x": [], "a": []}, {"x": [], "a": []}
    for j in rng.choice(n, 16, replace=False):
        for d in range(3):
            for which, base, grad in (("x", X0, xb), ("a", A0, alb)):
                p, m = base.copy(), base.copy()
                p[j, d] += h
                m[j, d] -= h
                g.update_arrays(**({"X": p} if which == "x" else {"alpha": p}))
                sp = _objective(g, ub, ab)
                g.update_arrays(**({"X": m} if which == "x" else {"alpha": m}))
                sm = _objective(g, ub, ab)
                g.update_arrays(X=X0, alpha=A0)
                fd[which].append((sp - sm) / (2 * h))
                an[which].append(grad[j, d])
    ex, ea = rel(an["x"], fd["x"]), rel(an["a"], fd["a"])
    print(f"[vjp P={P}] vs frozen-tree FD: xbar {ex:.3e} alphabar {ea:.3e}", flush=True)
    assert ex < 1e-6 and ea < 1e-6


def test_eval_vjp_mixed_matches_fp64():
    """Mixed precision (FP32 pair arithmetic) against the FP64 adjoint on a multi-level tree,
    leaves of every size (ncrit 64 and 200), with the far field included."""
    n = 20000
    X, A, S = _blobs(n, seed=21)
    rng = np.random.default_rng(6)
    ub, ab = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    for ncrit in (64, 200):
        out = {}
        for prec in ("fp64", "mixed"):
            g = FMMCalculator(P=8, theta=0.5, ncrit=ncrit, precision=prec)
            g.set_arrays(X, A, S)
            g.build()
            out[prec] = g.eval_vjp(ub, ab)
            out[prec + "_ms"] = g.timings()["vjp_ms"]
        ex, ea = rel(out["mixed"][0], out["fp64"][0]), rel(out["mixed"][1], out["fp64"][1])
        print(f"[mixed vs fp64, ncrit {ncrit}] xbar {ex:.3e} alphabar {ea:.3e} "
              f"(vjp {out['fp64_ms']:.2f} / {out['mixed_ms']:.2f} ms)", flush=True)
        assert ex < 1e-5 and ea < 1e-5


def _advance_orig(g, X, A, S, dt, steps=1):
    """State after `steps` advances from (X, A, S) given in the original order, in that order."""
    g.set_arrays(X, A, S)
    for _ in range(steps):
        g.step(dt)
    arr, p = g.arrays(), g.permutation()
    Xo, Ao = np.empty_like(X), np.empty_like(A)
    Xo[p], Ao[p] = arr["X"], arr["alpha"]
    return Xo, Ao


def _fd_state(g, X, A, S, dt, cx, ca, steps, idx, h=1e-6):
    fx, fa = [], []
    for j in idx:
        for d in range(3):
            for which in ("x", "a"):
                vals = []
                for sgn in (1, -1):
                    Xp, Ap = X.copy(), A.copy()
                    (Xp if which == "x" else Ap)[j, d] += sgn * h
                    Xo, Ao = _advance_orig(g, Xp, Ap, S, dt, steps)
                    vals.append(np.sum(cx * Xo) + np.sum(ca * Ao))
                (fx if which == "x" else fa).append((vals[0] - vals[1]) / (2 * h))
    return np.array(fx), np.array(fa)


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk3", "rk4"])
def test_step_vjp_finite_differences(scheme):
    n, dt = 160, 0.02
    X, A, S = _blobs(n, seed=11)
    A = A * 0.05
    rng = np.random.default_rng(8)
    cx, ca = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    g = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme=scheme)
    g.set_arrays(X, A, S)
    xb, ab = g.step_vjp(dt, cx, ca)
    # the state is left at the step's start
    arr, p = g.arrays(), g.permutation()
    assert np.array_equal(p, np.arange(n)) and np.array_equal(arr["X"], X) and np.array_equal(arr["alpha"], A)
    idx = rng.choice(n, 5, replace=False)
    fx, fa = _fd_state(g, X, A, S, dt, cx, ca, 1, idx)
    ax = np.array([xb[j, d] for j in idx for d in range(3)])
    aa = np.array([ab[j, d] for j in idx for d in range(3)])
    print(f"[{scheme}] step vjp vs FD: xbar {rel(ax, fx):.3e} alphabar {rel(aa, fa):.3e}", flush=True)
    assert rel(ax, fx) < 1e-7 and rel(aa, fa) < 1e-7


def test_trajectory_vjp_finite_differences():
    n, dt, steps = 160, 0.02, 5
    X, A, S = _blobs(n, seed=12)
    A = A * 0.05
    rng = np.random.default_rng(9)
    cx, ca = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    g = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme="rk2")
    g.set_arrays(X, A, S)
    g.trajectory_forward(steps, dt, every=2)
    fin, p = g.arrays(), g.permutation()
    xb, ab = g.trajectory_vjp(cx, ca)
    # the particles end at the final state again
    arr2, p2 = g.arrays(), g.permutation()
    Xf, Xf2 = np.empty_like(X), np.empty_like(X)
    Xf[p], Xf2[p2] = fin["X"], arr2["X"]
    assert np.array_equal(Xf, Xf2)
    # composed single-step VJPs give the same cotangent
    states = [(X, A)]
    for _ in range(steps - 1):
        states.append(_advance_orig(g, *states[-1], S, dt))
    lx, la = cx, ca
    for Xk, Ak in reversed(states):
        g.set_arrays(Xk, Ak, S)
        lx, la = g.step_vjp(dt, lx, la)
    assert rel(xb, lx) < 1e-13 and rel(ab, la) < 1e-13
    idx = rng.choice(n, 4, replace=False)
    fx, fa = _fd_state(g, X, A, S, dt, cx, ca, steps, idx)
    ax = np.array([xb[j, d] for j in idx for d in range(3)])
    aa = np.array([ab[j, d] for j in idx for d in range(3)])
    print(f"[trajectory] vs FD: xbar {rel(ax, fx):.3e} alphabar {rel(aa, fa):.3e}", flush=True)
    assert rel(ax, fx) < 1e-7 and rel(aa, fa) < 1e-7


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
def test_trajectory_vjp_other_schemes(scheme):
    """Checkpointed trajectories of RK3 and RK4 (every = 3 over 7 steps: a short last segment)
    equal the composed step VJPs."""
    n, dt, steps = 120, 0.02, 7
    X, A, S = _blobs(n, seed=13)
    A = A * 0.05
    rng = np.random.default_rng(10)
    cx, ca = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    g = FMMCalculator(P=6, theta=0.5, ncrit=256, scheme=scheme)
    g.set_arrays(X, A, S)
    g.trajectory_forward(steps, dt, every=3)
    xb, ab = g.trajectory_vjp(cx, ca)
    states = [(X, A)]
    for _ in range(steps - 1):
        states.append(_advance_orig(g, *states[-1], S, dt))
    lx, la = cx, ca
    for Xk, Ak in reversed(states):
        g.set_arrays(Xk, Ak, S)
        lx, la = g.step_vjp(dt, lx, la)
    assert rel(xb, lx) < 1e-13 and rel(ab, la) < 1e-13


def test_vjp_rejects_unsupported_configurations():
    X, A, S = _blobs(200)
    g = FMMCalculator(P=4, theta=0.5, ncrit=64, formulation="rvpm")
    g.set_arrays(X, A, S)
    g.build()
    with pytest.raises(UVLMError):
        g.eval_vjp(np.ones((200, 3)), np.ones((200, 3)))
    g2 = FMMCalculator(P=4, theta=0.5, ncrit=64)
    g2.set_arrays(X, A, S)
    with pytest.raises(UVLMError):
        g2.eval_vjp(np.ones((200, 3)), np.ones((200, 3)))  # no build yet
    with pytest.raises(UVLMError):
        g2.trajectory_vjp(np.ones((200, 3)), np.ones((200, 3)))  # no forward trajectory
