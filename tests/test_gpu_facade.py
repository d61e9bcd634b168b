"""The C++ derivative API and exafmm::getList, called from C++ (tests/cpp/adjoint_demo.cpp).

The demo drives the FMMCalculator method pairs (rateVJP, stepVJP, forwardTrajectory /
adjointTrajectory, probeAdjoint), the autograd-style functions of include/uvlm/autograd.hpp
(forward, backward, and backward again after the calculator was reused, which recomputes the
forward) and exafmm::getList on exafmm::buildTree cells.  Its results must equal those of the same
entry points called through the C ABI from Python (whose gradients the other GPU tests check
against finite differences and the oracle): the same kernels on the same inputs, so the bar is
1e-13, and what this pins is the marshalling (body order, cotangent order, sizes).  getList's pair
sets must equal the oracle's traversal (traverse.hpp:32-85) exactly.
"""
import os
import subprocess

import numpy as np
import pytest

from uvlm_amd import FMMCalculator, workloads as W

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "uvlm_amd", "lib", "uvlm_adjoint_demo")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def to_caller(perm, a):
    out = np.empty_like(a)
    out[perm] = a
    return out


def test_cpp_derivative_api_and_getlist(orc, tmp_path):
    n, m = 6000, 256
    X, A, S, probes, _, _ = W.blobs(n=n, seed=31, n_probe=m, n_check=1)
    rng = np.random.default_rng(8)
    ub, ab, w = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(m, 3))
    fin, fout, fpairs, ftx = (tmp_path / f for f in ("in.bin", "out.bin", "pairs.bin", "tree_x.bin"))
    np.concatenate([[n, m], X.ravel(), A.ravel(), S, ub.ravel(), ab.ravel(), probes.ravel(), w.ravel()]).tofile(fin)
    print("[facade] running the C++ derivative demo", flush=True)
    r = subprocess.run([EXE, str(fin), str(fout), str(fpairs), str(ftx)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = np.fromfile(fout)
    pos = 0

    def take(k, shape=(n, 3)):
        nonlocal pos
        v = out[pos:pos + k]
        pos += k
        return v.reshape(shape) if shape else v

    c_u, c_dadt, c_xb, c_ab, c_xb2, c_ab2 = (take(3 * n) for _ in range(6))
    c_xbt, c_abt = take(3 * n), take(3 * n)
    c_perm = take(n, (n,)).astype(np.int64)
    c_sx, c_sa = take(3 * n), take(3 * n)
    c_Xf, c_Af, c_tx, c_ta, c_tx2, c_ta2 = (take(3 * n) for _ in range(6))
    c_J = take(1, None)[0]
    c_px, c_pa = take(3 * n), take(3 * n)
    assert pos == out.size

    g = FMMCalculator(P=8, theta=0.5, ncrit=64, scheme="rk2")
    # rates and their VJP
    g.set_arrays(X, A, S)
    g.build()
    g.calculate()
    arr, perm = g.arrays(), g.permutation()
    assert np.array_equal(perm, c_perm)
    xb, alb = g.eval_vjp(ub[perm], ab[perm])
    checks = {
        "rates u": (c_u, to_caller(perm, arr["velocity"])),
        "rates dadt": (c_dadt, to_caller(perm, arr["dadt"])),
        "rates xbar": (c_xb, to_caller(perm, xb)),
        "rates alphabar": (c_ab, to_caller(perm, alb)),
        "rates xbar (recomputed forward)": (c_xb2, to_caller(perm, xb)),
        "rates alphabar (recomputed forward)": (c_ab2, to_caller(perm, alb)),
        "rateVJP xbar (tree order)": (c_xbt, xb),
        "rateVJP alphabar (tree order)": (c_abt, alb),
    }
    # one step backwards
    g.set_arrays(X, A, S)
    sx, sa = g.step_vjp(0.01, ub, ab)
    checks["stepVJP xbar"] = (c_sx, sx)
    checks["stepVJP alphabar"] = (c_sa, sa)
    # trajectory
    g.set_arrays(X, A, S)
    g.trajectory_forward(3, 0.01, every=2)
    fin_arr, fperm = g.arrays(), g.permutation()
    tx, ta = g.trajectory_vjp(ub, ab)
    checks["trajectory X"] = (c_Xf, to_caller(fperm, fin_arr["X"]))
    checks["trajectory alpha"] = (c_Af, to_caller(fperm, fin_arr["alpha"]))
    for k, (a, b) in {"trajectory xbar": (c_tx, tx), "trajectory alphabar": (c_ta, ta),
                      "trajectory xbar (recomputed forward)": (c_tx2, tx),
                      "trajectory alphabar (recomputed forward)": (c_ta2, ta)}.items():
        checks[k] = (a, b)
    # probe objective
    g.set_arrays(X, A, S)
    g.build()
    g.calculate()
    J = g.probe_objective(probes, w)
    ga, gx = g.probe_adjoint(probes, w)
    pperm = g.permutation()
    checks["probe xbar"] = (c_px, 2.0 * to_caller(pperm, gx))
    checks["probe alphabar"] = (c_pa, 2.0 * to_caller(pperm, ga))
    for k, (a, b) in checks.items():
        e = rel(a, b)
        print(f"  {k}: {e:.2e}", flush=True)
        assert e < 1e-13, k
    print(f"  probe J: C++ {c_J:.15e} python {J:.15e}", flush=True)
    assert abs(c_J - J) <= 1e-12 * abs(J)
    # FD sanity of the rates VJP through the C++ function (caller order): one coordinate
    assert np.linalg.norm(c_xb) > 0 and np.linalg.norm(c_tx) > 0

    # getList against the oracle's traversal of the same tree
    pr = np.fromfile(fpairs, dtype=np.int64)
    np2p, nm2l = int(pr[0]), int(pr[1])
    gp = pr[2:2 + 2 * np2p].reshape(-1, 2)
    gm = pr[2 + 2 * np2p:].reshape(-1, 2)
    assert gm.shape[0] == nm2l
    b = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=4, theta=0.5, ncrit=32)
    c.init(b)
    assert np.array_equal(np.fromfile(ftx).reshape(n, 3), b[:, 9:12])
    rp, rm = c.pairs()
    print(f"  getList: {np2p} P2P, {nm2l} M2L pairs", flush=True)
    assert np.array_equal(np.unique(rp, axis=0), gp) and np.array_equal(np.unique(rm, axis=0), gm)
