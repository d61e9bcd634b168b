"""More pins of the oracle against what the mathematics or the reference's own rules fix.

Each test checks an oracle function against something that does not retype its formula, and
each is chosen so that a plausible mistake fails it:
  M2L truncation   -> P2M -> M2L -> L2P of one source equals the curl of the Legendre series of
                      1/|R + d| cut at total degree P-1 (multipoleKernels.hpp:518-552 with the
                      degree < P singular harmonics of 137-172); the full-order variant fails it
  rVPM             -> the correction's geometric identity (fmmCalculator.hpp:407-418)
  sensors          -> probes on the particles reproduce the particle velocity; convergence to the
                      direct sum with P (fmmCalculator.hpp:469-491)
  bounds           -> the closed form of tree.hpp:7-26 bit for bit, including the 1.00001 leeway
  traversal        -> every pair descends from (root, root) by splits that follow traverse.hpp:74
                      through pairs that failed the MAC; the other tie rule is rejected
  leaf order       -> bodies inside a leaf keep their input order (tree.hpp:70-78)
  RK stages        -> the Butcher form of each scheme (fmmCalculator.hpp:493-711) on the stage
                      rates, including the 0.5 dt third stage of RK4 as written
"""
import numpy as np
import pytest

from uvlm_amd import workloads as W

F = {"J": slice(0, 9), "X": slice(9, 12), "alpha": slice(12, 15), "velocity": slice(15, 18), "dadt": slice(18, 21)}
SIG, DRDT = 33, 34


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


# ------------------------------------------------------------------- M2L truncation
def legendre_series(R, d, K):
    """Phi_K(d) = sum_{k<K} |d|^k P_k(c) / |R|^(k+1), c = -R.d/(|R||d|) (the first K terms of the
    expansion of 1/|R + d|) and its gradient in d.  The solid form Q_k = |d|^k P_k(c) is a
    polynomial in d: Q_0 = 1, Q_1 = t, (k+1) Q_{k+1} = (2k+1) t Q_k - k |d|^2 Q_{k-1}, t = -R.d/|R|
    (Bonnet's recurrence), so complex d gives complex-step derivatives."""
    Rn = np.sqrt(R @ R)
    Rh = R / Rn
    t, s2 = -(Rh @ d), d @ d
    dt = np.result_type(d, float)
    Q, G = [1.0 + 0 * t, t], [np.zeros(3, dtype=dt), -Rh.astype(dt)]
    for k in range(1, K):
        Q.append(((2 * k + 1) * t * Q[k] - k * s2 * Q[k - 1]) / (k + 1))
        G.append(((2 * k + 1) * (-Rh * Q[k] + t * G[k]) - k * (2 * d * Q[k - 1] + s2 * G[k - 1])) / (k + 1))
    return sum(Q[k] / Rn ** (k + 1) for k in range(K)), sum(G[k] / Rn ** (k + 1) for k in range(K))


def series_velocity(R, d, a, K):
    """u = grad Phi_K x a / 4pi (the stream function a Phi / 4pi of a unit vortex) and
    J[:, b] = d u / d x_b by a complex step of the gradient (exact: Phi is a polynomial in d)."""
    _, g = legendre_series(R, d, K)
    u = np.cross(g, a) / (4 * np.pi)
    J = np.zeros((3, 3))
    h = 1e-30
    for b in range(3):
        db = d.astype(complex)
        db[b] += 1j * h
        _, gb = legendre_series(R, db, K)
        J[:, b] = np.cross(gb.imag / h, a) / (4 * np.pi)
    return u, J


@pytest.mark.parametrize("P", [4, 8, 10, 12])
def test_m2l_truncation_is_the_legendre_series_to_degree_P_minus_1(orc, P):
    rng = np.random.default_rng(100 + P)
    for _ in range(4):
        Xs = rng.uniform(-1, 1, 3)
        R = rng.normal(size=3)
        R *= rng.uniform(1.1, 1.6) / np.linalg.norm(R)
        Xt = Xs + R
        rs, rt = rng.uniform(-0.25, 0.25, 3), rng.uniform(-0.25, 0.25, 3)
        a = rng.normal(size=3)
        M = orc.p2m(P, Xs, (Xs + rs)[None], a[None])
        u_ref, J_ref = series_velocity(R, rt - rs, a, P)
        u_full, _ = series_velocity(R, rt - rs, a, 2 * P - 1)
        for order, must_match in ((None, True), (2 * P - 1, False)):
            L = orc.m2l(P, Xt, Xs, M, harmonic_order=order)
            u, J = orc.l2p(P, Xt, L, (Xt + rt)[None])
            eu, eJ = rel(u[0], u_ref), rel(J[0], J_ref.ravel())
            if must_match:
                assert eu < 1e-12 and eJ < 1e-12, (P, eu, eJ)
            else:  # the planted full-order truncation is told apart from the reference's
                assert eu > 1e-10, (P, eu)
        # the series to degree 2P-2 differs from the degree P-1 one by far more than the bar
        assert rel(u_full, u_ref) > 1e-10


# ------------------------------------------------------------------- rVPM
def test_rvpm_correction_identity(orc):
    """After the rVPM terms the stretching's component along alpha is 2/5 of the classic VPM one,
    its perpendicular part is unchanged, and drdt = -(sigma / (5 |alpha|)) (dadt_cvpm . alpha_hat)."""
    X, A, S = W.scaled("blobs", 3000)
    out = {}
    for form in ("cvpm", "rvpm"):
        b = orc.make_bodies(X, A, S)
        c = orc.Calculator(P=6, theta=0.5, ncrit=32, formulation=form)
        c.init(b)
        c.calculate()
        out[form] = c.bodies(len(X))
    cv, rv = out["cvpm"], out["rvpm"]
    assert np.array_equal(cv[:, F["X"]], rv[:, F["X"]])
    a = cv[:, F["alpha"]]
    ah = a / np.linalg.norm(a, axis=1)[:, None]
    d0, d1 = cv[:, F["dadt"]], rv[:, F["dadt"]]
    p0, p1 = np.sum(d0 * ah, axis=1), np.sum(d1 * ah, axis=1)
    scale = np.linalg.norm(d0, axis=1)
    assert np.all(np.abs(p1 - 0.4 * p0) <= 1e-12 * scale)
    assert np.all(np.linalg.norm((d1 - p1[:, None] * ah) - (d0 - p0[:, None] * ah), axis=1) <= 1e-12 * scale)
    want = -cv[:, SIG] / (5 * np.linalg.norm(a, axis=1)) * p0
    assert np.all(cv[:, DRDT] == 0)
    assert np.all(np.abs(rv[:, DRDT] - want) <= 1e-12 * np.abs(want).max())
    assert np.allclose(rv[:, F["velocity"]], cv[:, F["velocity"]], rtol=0, atol=0)


# ------------------------------------------------------------------- sensors
def test_sensors_on_the_particles_reproduce_the_particle_velocity(orc):
    """Probes at the particle positions (in tree order) build the same tree and the same lists
    against the particle tree, so the sensor velocity must equal the particle velocity of the
    last calculate(): near pairs by the same P2P, far pairs by M2L from the same multipoles."""
    X, A, S = W.scaled("ring", 6000)
    b = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=32)
    c.init(b)
    c.calculate()
    out = c.bodies(len(X))
    sen = orc.make_bodies(out[:, F["X"]].copy(), np.zeros((len(X), 3)), np.zeros(len(X)))
    c.sensor_data(sen)
    assert np.array_equal(sen[:, F["X"]], out[:, F["X"]])
    e = rel(sen[:, F["velocity"]], out[:, F["velocity"]])
    assert e < 1e-14, e


def test_sensor_velocity_converges_to_direct_sum(orc):
    """With sigma -> 0 the far field's singular kernel is exact, so the probe velocity error
    against the direct sum is the expansion truncation alone: it falls with every P."""
    X, A, S, probes, _, _ = W.blobs(n=12000, seed=5, n_probe=400, n_check=4)
    S = np.full(len(X), 1e-5)
    errs = []
    for P in (4, 6, 8, 10, 12):
        b = orc.make_bodies(X, A, S)
        c = orc.Calculator(P=P, theta=0.5, ncrit=32)
        c.init(b)
        c.calculate()
        sen = orc.make_bodies(probes, np.zeros_like(probes), np.zeros(len(probes)))
        c.sensor_data(sen)
        ud = orc.direct(sen[:, F["X"]], b[:, F["X"]], b[:, F["alpha"]], b[:, SIG], want_J=False)
        errs.append(rel(sen[:, F["velocity"]], ud))
    assert all(errs[i + 1] < errs[i] for i in range(len(errs) - 1)), errs
    assert errs[-1] < errs[0] / 20 and errs[-1] < 2e-4, errs


# ------------------------------------------------------------------- bounds and tree order
@pytest.mark.parametrize("name", ["ring", "cube", "blobs"])
def test_bounds_closed_form_bitexact(orc, name):
    X, A, S = W.scaled(name, 5000)
    b = orc.make_bodies(X, A, S)
    ints, reals, X0, R0 = orc.build_tree(b, 32)
    lo, hi = X.min(axis=0), X.max(axis=0)
    c = (hi + lo) / 2
    r = 0.0
    for d in range(3):
        r = max(r, float(c[d] - lo[d]))
        r = max(r, float(hi[d] - c[d]))
    assert np.array_equal(X0, c) and R0 == r * 1.00001 and R0 != r
    assert np.array_equal(reals[0, :3], c) and reals[0, 3] == R0


@pytest.mark.parametrize("name", ["ring", "blobs"])
def test_leaf_bodies_keep_input_order(orc, name):
    X, A, S = W.scaled(name, 5000)
    idx = np.arange(len(X), dtype=float)
    b = orc.make_bodies(X, A, S)
    b[:, 36] = idx  # drdt_old carries the input index through the tree build
    ints, reals, _, _ = orc.build_tree(b, 32)
    src = b[:, 36].astype(np.int64)
    assert np.array_equal(np.sort(src), np.arange(len(X)))
    assert np.array_equal(b[:, F["X"]], X[src])
    nchild, nbody, _, body, _ = ints.T
    for c in np.nonzero(nchild == 0)[0]:
        seg = src[body[c]:body[c] + nbody[c]]
        assert np.all(np.diff(seg) > 0), c


# ------------------------------------------------------------------- traversal split rule
def _descends_by_rule(ints, reals, theta, pairs, ge_ties=True):
    """True if every pair has exactly one parent pair from which the traversal could have split
    it: (parent(t), s) split on the target or (t, parent(s)) split on the source, as the rule
    decides for that parent pair, with the parent itself failing the MAC (or being the root
    pair) and descending the same way."""
    nchild, _, child, _, _ = ints.T
    Xc, R = reals[:, :3], reals[:, 3]
    parent = np.full(len(ints), -1)
    for q in range(len(ints)):
        parent[child[q]:child[q] + nchild[q]] = q

    def near(t, s):
        dx, dy, dz = (float(Xc[t, k] - Xc[s, k]) for k in range(3))
        return (dx * dx + dy * dy + dz * dz) * theta * theta < (float(R[t]) + float(R[s])) ** 2

    def splits_target(t, s):
        bigger = R[t] >= R[s] if ge_ties else R[t] > R[s]
        return nchild[s] == 0 or (bigger and nchild[t] != 0)

    memo = {(0, 0): True}

    def visited(t, s):
        if (t, s) in memo:
            return memo[(t, s)]
        ok = False
        if parent[t] >= 0:
            p = (int(parent[t]), s)
            ok |= splits_target(*p) and (p == (0, 0) or near(*p)) and visited(*p)
        if parent[s] >= 0:
            p = (t, int(parent[s]))
            ok |= (not splits_target(*p)) and (p == (0, 0) or near(*p)) and visited(*p)
        memo[(t, s)] = ok
        return ok

    def n_parents(t, s):
        n = 0
        if parent[t] >= 0:
            p = (int(parent[t]), s)
            n += splits_target(*p) and (p == (0, 0) or near(*p)) and visited(*p)
        if parent[s] >= 0:
            p = (t, int(parent[s]))
            n += (not splits_target(*p)) and (p == (0, 0) or near(*p)) and visited(*p)
        return n

    return all(n_parents(int(t), int(s)) == 1 for t, s in pairs)


@pytest.mark.parametrize("name,theta", [("ring", 0.5), ("cube", 0.4), ("blobs", 0.5)])
def test_traversal_follows_the_split_rule(orc, name, theta):
    X, A, S = W.scaled(name, 4000)
    b = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=4, theta=theta, ncrit=24)
    c.init(b)
    ints, reals = c.cells()
    p2p, m2l = c.pairs()
    pairs = np.concatenate([p2p, m2l])
    assert _descends_by_rule(ints, reals, theta, pairs, ge_ties=True)
    # the tie rule matters: with R_t > R_s (split the source on equal radii) the same lists
    # do not descend from the root
    assert not _descends_by_rule(ints, reals, theta, pairs, ge_ties=False)


# ------------------------------------------------------------------- RK stages
BUTCHER = {
    "euler": ([[]], [1.0]),
    "rk2": ([[], [1.0]], [0.5, 0.5]),
    "rk3": ([[], [1 / 3], [-3 / 16, 15 / 16]], [1 / 6, 3 / 10, 8 / 15]),
    # RK4 as written: every stage point at y0 + dt/2 k (fmmCalculator.hpp:683-689)
    "rk4": ([[], [0.5], [0.0, 0.5], [0.0, 0.0, 0.5]], [1 / 6, 1 / 3, 1 / 3, 1 / 6]),
}


@pytest.mark.parametrize("scheme", list(BUTCHER))
def test_rk_stages_follow_the_butcher_form(orc, scheme):
    """One leaf holds every body, so the order is kept and each stage's rates are exact. The
    state after stage s must be y0 + dt sum_r a[s+1][r] k_r and after the last stage
    y0 + dt sum_r b_r k_r, with k_r the rates the stage evaluated (velocity, dadt, drdt)."""
    rng = np.random.default_rng(11)
    n, dt = 40, 0.05
    X = rng.normal(size=(n, 3)) * 0.4
    A = rng.normal(size=(n, 3)) * 0.1
    S = rng.uniform(0.2, 0.3, n)
    b = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=4, theta=0.5, ncrit=1000, formulation="rvpm")  # rVPM: drdt != 0
    a_rows, weights = BUTCHER[scheme]
    y0 = np.concatenate([X, A, S[:, None]], axis=1)
    ks = []
    for s in range(len(weights)):
        c.stage(b, dt, scheme, s)
        ks.append(np.concatenate([b[:, F["velocity"]], b[:, F["dadt"]], b[:, [DRDT]]], axis=1))
        y = np.concatenate([b[:, F["X"]], b[:, F["alpha"]], b[:, [SIG]]], axis=1)
        coef = a_rows[s + 1] if s + 1 < len(weights) else weights
        want = y0 + dt * sum(w * k for w, k in zip(coef, ks))
        tol = 1e-14 * (np.abs(y0) + dt * sum(np.abs(k) for k in ks))
        assert np.all(np.abs(y - want) <= tol + 1e-300), (scheme, s, np.max(np.abs(y - want) / (tol + 1e-300)))
    assert np.abs(ks[-1][:, 6]).max() > 0  # the radius equation was exercised


# ------------------------------------------------------------------- near-field option
def test_near_field_option_keeps_the_lists_complete_and_removes_the_kernel_mismatch(orc):
    """With the near-field option (pairs whose bodies can come closer than c * sigma_max are near),
    every M2L pair's bodies stay at least c * sigma_max apart, the lists still cover every source
    body exactly once per target leaf, c = 0 gives the reference's lists, and the FMM error against
    the direct sum falls from the regularisation mismatch (~1e-2) to the expansion truncation."""
    X, A, S = W.scaled("cube", 8000)
    errs = {}
    for c in (0.0, 6.0):
        b = orc.make_bodies(X, A, S)
        calc = orc.Calculator(P=12, theta=0.4, ncrit=32)
        calc.set_near_field(c)
        calc.init(b)
        calc.calculate()
        out = calc.bodies(len(X))
        ints, reals = calc.cells()
        p2p, m2l = calc.pairs()
        nchild, nbody, child, body, level = ints.T
        Xc, R = reals[:, :3], reals[:, 3]
        gap = c * S.max()
        d = np.linalg.norm(Xc[m2l[:, 0]] - Xc[m2l[:, 1]], axis=1)
        if c > 0:
            assert np.all(d >= gap + np.sqrt(3) * (R[m2l[:, 0]] + R[m2l[:, 1]]) - 1e-12)
            # bodies of every M2L pair are at least gap apart (checked on a sample of pairs)
            for t, s in m2l[:: max(1, len(m2l) // 200)]:
                xt, xs = out[body[t]:body[t] + nbody[t], 9:12], out[body[s]:body[s] + nbody[s], 9:12]
                dmin = np.min(np.linalg.norm(xt[:, None, :] - xs[None, :, :], axis=2))
                assert dmin >= gap, (t, s, dmin, gap)
        parent = np.full(len(ints), -1)
        for q in range(len(ints)):
            parent[child[q]:child[q] + nchild[q]] = q
        by_t = {}
        for t, s in np.concatenate([p2p, m2l]):
            by_t.setdefault(t, []).append(s)
        for t in np.nonzero(nchild == 0)[0][::5]:
            cnt = np.zeros(len(X), dtype=int)
            a = t
            while a >= 0:
                for s in by_t.get(a, []):
                    cnt[body[s]:body[s] + nbody[s]] += 1
                a = parent[a]
            assert (cnt == 1).all(), (c, t)
        ud, Jd = orc.direct(out[:500, 9:12], out[:, 9:12], out[:, 12:15], out[:, 33])
        errs[c] = np.linalg.norm(out[:500, 15:18] - ud) / np.linalg.norm(ud)
    assert errs[0.0] > 1e-3 and errs[6.0] < 2e-5, errs
