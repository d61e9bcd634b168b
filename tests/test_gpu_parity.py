"""HIP path vs the oracle (CPU restatement of the reference) on the same seeded inputs.

Bars: tree, permutation and interaction lists bit-exact; FP64 FMM within 1e-10 relative
L2 (u, J, dadt, expansions); mixed precision within 1e-6 of the FP64 FMM.
"""
import numpy as np
import pytest

from uvlm_amd import FMMCalculator, make_bodies, workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    "ring": lambda: W.ring(),                       # 8,192 particles, config of the single ring
    "cube": lambda: W.scaled("cube", 32768),
    "blobs": lambda: W.scaled("blobs", 24000),
    "leapfrog": lambda: W.scaled("leapfrog", 32768),
}
PARAMS = {"ring": (8, 0.5, 64), "cube": (10, 0.4, 64), "blobs": (8, 0.5, 64), "leapfrog": (10, 0.5, 64)}


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def run_pair(orc, name, precision="fp64", stretching="transposed", formulation="cvpm"):
    X, A, S = CASES[name]()
    P, th, nc = PARAMS[name]
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=P, theta=th, ncrit=nc, stretching=stretching, formulation=formulation)
    c.init(b_ref)
    c.calculate()
    ref = c.bodies(b_ref.shape[0])
    g = FMMCalculator(P=P, theta=th, ncrit=nc, precision=precision, stretching=stretching, formulation=formulation)
    b_gpu = make_bodies(X, A, S)
    g.init(b_gpu)
    g.calculate()
    out = g.copyBodies2Host()
    print(f"[{name}/{precision}] N={X.shape[0]} timings={g.timings()}", flush=True)
    return c, g, ref, out, b_ref, b_gpu


@pytest.mark.parametrize("name", list(CASES))
def test_tree_and_lists_bitexact(orc, name):
    X, A, S = CASES[name]()
    P, th, nc = PARAMS[name]
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=P, theta=th, ncrit=nc)
    c.init(b_ref)
    g = FMMCalculator(P=P, theta=th, ncrit=nc)
    b_gpu = make_bodies(X, A, S)
    g.init(b_gpu)
    assert np.array_equal(b_gpu, b_ref), "permuted bodies differ"
    ri, rr = c.cells()
    gi, gr = g.cells()
    assert ri.shape == gi.shape, (ri.shape, gi.shape)
    assert np.array_equal(ri, gi)
    assert np.array_equal(rr, gr)
    rp, rm = c.pairs()
    gp, gm = g.lists()
    assert np.array_equal(np.unique(rp, axis=0), gp)
    assert np.array_equal(np.unique(rm, axis=0), gm)
    assert len(np.unique(rp, axis=0)) == len(rp) and len(np.unique(rm, axis=0)) == len(rm)
    # the counts the bench's interactions/s rest on: pairs, and bodies(t) * bodies(s) over P2P pairs
    t = g.timings()
    nbody = ri[:, 1]
    assert t["n_p2p_pairs"] == len(rp) and t["n_m2l_pairs"] == len(rm)
    assert t["n_interactions"] == int(np.sum(nbody[rp[:, 0]] * nbody[rp[:, 1]])), t["n_interactions"]


def test_build_defers_the_output_permutation():
    """uvlm_build permutes the fields an evaluation overwrites (J, velocity, dadt, drdt) only when
    something reads them before the next evaluation. After a rebuild on moved bodies the last
    evaluation's outputs must follow their bodies exactly, through the SoA and AoS readers, also
    after two rebuilds in a row with nothing in between, and an evaluation after a deferred
    build must match one after a full build."""
    X, A, S = W.scaled("cube", 6000)
    g = FMMCalculator(P=6, theta=0.5, ncrit=32)
    g.set_arrays(X, A, S)
    g.build()
    g.calculate()
    a1, o1 = g.arrays(), g.permutation()
    pos1 = np.empty_like(o1)
    pos1[o1] = np.arange(len(o1))
    rng = np.random.default_rng(3)
    X2 = a1["X"] + rng.normal(scale=0.05, size=a1["X"].shape)
    g.update_arrays(X=X2)
    g.build()
    g.build()  # the second build applies the first one's pending permutation first
    o2 = g.permutation()
    assert not np.array_equal(o1, o2)
    src = pos1[o2]  # tree position before the rebuilds of the body now at each position
    a2 = g.arrays()
    for k in ("velocity", "J", "dadt", "drdt"):
        assert np.array_equal(a2[k], a1[k][src]), k
    assert np.array_equal(a2["X"], X2[src])
    g.build()
    b = g.copyBodies2Host()
    o3 = g.permutation()
    pos2 = np.empty_like(o2)
    pos2[o2] = np.arange(len(o2))
    src3 = src[pos2[o3]]
    assert np.array_equal(b[:, 15:18], a1["velocity"][src3]) and np.array_equal(b[:, 0:9], a1["J"][src3])
    # an evaluation after a deferred build equals one after a full build (set_bodies resets)
    g.calculate()
    u_def = g.arrays()["velocity"]
    h = FMMCalculator(P=6, theta=0.5, ncrit=32)
    h.init(b.copy())
    h.calculate()
    assert np.array_equal(u_def, h.arrays()["velocity"])


def test_tree_rebuilds_on_one_handle_stay_bitexact(orc):
    """A handle sorts only the key digits its previous tree used (plus one level); a much deeper
    tree on the next build (uniform cube -> clustered blobs) must fall back to the full sort, and
    every rebuild must still match the oracle's tree, permutation and lists bit for bit."""
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    for name in ("cube", "blobs", "cube", "leapfrog", "blobs"):
        X, A, S = CASES[name]()
        b_ref = orc.make_bodies(X, A, S)
        c = orc.Calculator(P=8, theta=0.5, ncrit=64)
        c.init(b_ref)
        b_gpu = make_bodies(X, A, S)
        g.init(b_gpu)
        ri, rr = c.cells()
        gi, gr = g.cells()
        print(f"[rebuild {name}] cells={len(ri)} levels={int(ri[:, 4].max()) + 1}", flush=True)
        assert np.array_equal(b_gpu, b_ref), name
        assert np.array_equal(ri, gi) and np.array_equal(rr, gr), name
        rp, rm = c.pairs()
        gp, gm = g.lists()
        assert np.array_equal(np.unique(rp, axis=0), gp) and np.array_equal(np.unique(rm, axis=0), gm), name


@pytest.mark.parametrize("ncrit", [100, 200, 700])
def test_large_leaves_match_oracle(orc, ncrit):
    """Leaves of more than 64 bodies take the 256-lane P2P, whose leaves of more than 128 bodies
    run several passes of fewer targets (and more than 256 bodies several passes of 256): FP64 against
    the oracle to 1e-10 and mixed against FP64 to 1e-6."""
    X, A, S = CASES["leapfrog"]()
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=ncrit)
    c.init(b_ref)
    c.calculate()
    ref = c.bodies(b_ref.shape[0])
    outs = {}
    for precision in ("fp64", "mixed"):
        g = FMMCalculator(P=8, theta=0.5, ncrit=ncrit, precision=precision)
        b = make_bodies(X, A, S)
        g.init(b)
        g.calculate()
        outs[precision] = g.copyBodies2Host()
    for f, sl in (("velocity", slice(15, 18)), ("J", slice(0, 9)), ("dadt", slice(18, 21))):
        e64 = rel(outs["fp64"][:, sl], ref[:, sl])
        emx = rel(outs["mixed"][:, sl], outs["fp64"][:, sl])
        print(f"ncrit {ncrit} {f}: fp64 vs oracle {e64:.2e}, mixed vs fp64 {emx:.2e}", flush=True)
        assert e64 < 1e-10 and emx < 1e-6, (f, e64, emx)


def test_lists_regrow_from_tiny_buffers(orc, monkeypatch):
    """The traversal runs its steps in batches without the host and, when a frontier or list
    buffer overflows, grows it and traverses again: with 64-entry first buffers (several regrowths)
    the lists and the FMM result still match the oracle exactly."""
    monkeypatch.setenv("UVLM_LIST_CAPACITY", "64")
    X, A, S = CASES["blobs"]()
    P, th, nc = PARAMS["blobs"]
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=P, theta=th, ncrit=nc)
    c.init(b_ref)
    c.calculate()
    ref = c.bodies(b_ref.shape[0])
    g = FMMCalculator(P=P, theta=th, ncrit=nc)
    b_gpu = make_bodies(X, A, S)
    g.init(b_gpu)
    rp, rm = c.pairs()
    gp, gm = g.lists()
    assert np.array_equal(np.unique(rp, axis=0), gp) and np.array_equal(np.unique(rm, axis=0), gm)
    g.calculate()
    out = g.copyBodies2Host()
    assert rel(out[:, 15:18], ref[:, 15:18]) < 1e-10


@pytest.mark.parametrize("name", list(CASES))
def test_fmm_fp64_parity(orc, name):
    c, g, ref, out, _, _ = run_pair(orc, name)
    for f, sl in (("velocity", slice(15, 18)), ("J", slice(0, 9)), ("dadt", slice(18, 21))):
        e = rel(out[:, sl], ref[:, sl])
        print(f"{name} {f} rel={e:.3e}")
        assert e < 1e-10, (f, e)
    M0, L0 = c.expansions()
    M1, L1 = g.expansions()
    assert rel(M1, M0) < 1e-10
    assert rel(L1, L0) < 1e-10


@pytest.mark.parametrize("P,S,Wv", [(2, 0, 1), (3, 0, 2), (4, 0, 2), (6, 0, 1), (12, 0, 2), (13, 0, 2), (16, 0, 2),
                                    (10, 4, 2), (10, 5, 2), (10, 6, 2), (10, 12, 2), (10, 6, 1), (12, 0, 1),
                                    (16, 0, 1)])
def test_fmm_fp64_parity_orders_and_plans(orc, P, S, Wv, monkeypatch):
    """Expansion orders away from the default (P = 13 and 16 need a third coefficient per lane
    in P2M/M2M/L2L and take the M2M without its term table; P = 2 and 3 leave the L2P's table of
    paired +-mu terms empty or with one entry; P = 4 runs with a single home slot), the M2L plan
    sizes 4, 5, 6 and 12, and one or two wavefronts per M2L workgroup (each with its own degrees)."""
    if S:
        monkeypatch.setenv("UVLM_M2L_SLOTS", str(S))
    monkeypatch.setenv("UVLM_M2L_WAVES", str(Wv))
    X, A, S_ = W.scaled("cube", 6000)
    b_ref = orc.make_bodies(X, A, S_)
    c = orc.Calculator(P=P, theta=0.5, ncrit=32)
    c.init(b_ref)
    c.calculate()
    ref = c.bodies(b_ref.shape[0])
    g = FMMCalculator(P=P, theta=0.5, ncrit=32)
    b_gpu = make_bodies(X, A, S_)
    g.init(b_gpu)
    g.calculate()
    out = g.copyBodies2Host()
    for sl in (slice(15, 18), slice(0, 9), slice(18, 21)):
        assert rel(out[:, sl], ref[:, sl]) < 1e-10, (P, S, Wv, sl)
    M0, L0 = c.expansions()
    M1, L1 = g.expansions()
    assert rel(M1, M0) < 1e-10 and rel(L1, L0) < 1e-10


@pytest.mark.parametrize("name", ["ring", "leapfrog", "cube"])
def test_mixed_vs_fp64(orc, name):
    _, _, ref, out, _, _ = run_pair(orc, name, precision="mixed")
    eu = rel(out[:, 15:18], ref[:, 15:18])
    eJ = rel(out[:, 0:9], ref[:, 0:9])
    ed = rel(out[:, 18:21], ref[:, 18:21])
    print(f"{name} mixed: u {eu:.3e} J {eJ:.3e} dadt {ed:.3e}")
    assert eu < 1e-6 and eJ < 1e-6 and ed < 1e-6


@pytest.mark.parametrize("variant", [("classic", "cvpm"), ("transposed", "rvpm")])
def test_stretching_variants(orc, variant):
    st, fo = variant
    _, _, ref, out, _, _ = run_pair(orc, "ring", stretching=st, formulation=fo)
    assert rel(out[:, 18:21], ref[:, 18:21]) < 1e-10
    assert rel(out[:, 34], ref[:, 34]) < 1e-10 or np.allclose(ref[:, 34], 0)


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk3", "rk4"])
def test_advance_matches_oracle(orc, scheme):
    """Every stage of two time steps, each started from the same full body state on both
    sides, agrees to 1e-10.  A trajectory over several rebuilds is compared only loosely:
    a position one ulp apart can fall on the other side of an octant plane or of the
    multipole acceptance test at a later rebuild, which moves the answer by the FMM
    truncation error (the oracle alone moves by ~3e-7 when its input is perturbed by
    1e-15 for this ring), so the trajectory bar is 1e-5."""
    X, A, S = W.scaled("ring", 4096)
    stages = {"euler": 1, "rk2": 2, "rk3": 3, "rk4": 4}[scheme]
    c = orc.Calculator(P=8, theta=0.5, ncrit=64)
    g = FMMCalculator(P=8, theta=0.5, ncrit=64, scheme=scheme)
    b = make_bodies(X, A, S)
    for _ in range(2):
        for k in range(stages):
            ref = b.copy()
            c.stage(ref, 0.01, scheme, k)
            g.init(b)
            g.calculate()
            g.stage(k, 0.01)
            g.copyBodies2Host(b)
            for sl in (slice(9, 12), slice(12, 15), slice(15, 18), slice(18, 21), slice(21, 24),
                       slice(24, 27), slice(27, 30), slice(30, 33), 33, 35):
                assert rel(b[:, sl], ref[:, sl]) < 1e-10, (k, sl)
    b_ref = orc.make_bodies(X, A, S)
    b_gpu = make_bodies(X, A, S)
    for _ in range(2):
        c.advance(b_ref, 0.01, scheme)
        g.advance(b_gpu, 0.01)
    for sl in (slice(9, 12), slice(12, 15)):
        assert rel(b_gpu[:, sl], b_ref[:, sl]) < 1e-5, sl


def test_sensors_match_oracle(orc):
    X, A, S, probes, v, _ = W.blobs(n=20000, n_probe=1024, n_check=4)
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=64)
    c.init(b_ref)
    c.calculate()
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    b_gpu = make_bodies(X, A, S)
    g.init(b_gpu)
    g.calculate()
    s_ref = orc.make_bodies(probes, np.zeros_like(probes), np.full(len(probes), 0.05))
    s_gpu = s_ref.copy()
    c.sensor_data(s_ref)
    g.getSensorData(s_gpu)
    assert np.array_equal(s_gpu[:, 9:12], s_ref[:, 9:12])
    assert rel(s_gpu[:, 15:18], s_ref[:, 15:18]) < 1e-10
    assert rel(s_gpu[:, 0:9], s_ref[:, 0:9]) < 1e-10


def test_single_leaf_and_tiny_sets(orc):
    for n in (1, 2, 17, 64):
        X, A, S = W.scaled("cube", max(n, 8))
        X, A, S = X[:n], A[:n], S[:n]
        b_ref = orc.make_bodies(X, A, S)
        c = orc.Calculator(P=6, theta=0.5, ncrit=64)
        c.init(b_ref)
        c.calculate()
        ref = c.bodies(n)
        g = FMMCalculator(P=6, theta=0.5, ncrit=64)
        b = make_bodies(X, A, S)
        g.init(b)
        g.calculate()
        out = g.copyBodies2Host()
        assert np.array_equal(b, b_ref)
        if n > 1:
            assert rel(out[:, 15:18], ref[:, 15:18]) < 1e-12
            assert rel(out[:, 0:9], ref[:, 0:9]) < 1e-12


def test_depth_error_is_reported():
    from uvlm_amd import UVLMError
    X = np.zeros((100, 3))
    X[:, 0] = np.arange(100) * 0.0  # 100 coincident bodies > ncrit: the octree cannot split them
    g = FMMCalculator(P=4, theta=0.5, ncrit=16)
    g.set_bodies(make_bodies(X, np.ones_like(X), np.ones(100)))
    with pytest.raises(UVLMError) as e:
        g.build()
    assert e.value.code == 2


def test_cpp_facade_drop_in(orc, tmp_path):
    """A C++ caller written against the reference's Vortex::FMMCalculator spelling
    (tests/cpp/facade_demo.cpp) gives the oracle's results."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "uvlm_amd", "lib",
                       "uvlm_facade_demo")
    X, A, S = W.scaled("ring", 4096)
    b = orc.make_bodies(X, A, S)
    fin, fev, fad = tmp_path / "in.bin", tmp_path / "eval.bin", tmp_path / "adv.bin"
    b.tofile(fin)
    r = subprocess.run([exe, str(fin), str(fev), str(fad), "2"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got_eval = np.fromfile(fev).reshape(-1, 37)
    got_adv = np.fromfile(fad).reshape(-1, 37)
    ref = b.copy()
    c = orc.Calculator(P=8, theta=0.5, ncrit=64)
    c.init(ref)
    c.calculate()
    want = c.bodies(len(ref))
    assert np.array_equal(got_eval[:, 9:12], want[:, 9:12])
    assert rel(got_eval[:, 15:18], want[:, 15:18]) < 1e-10 and rel(got_eval[:, 0:9], want[:, 0:9]) < 1e-10
    adv = want.copy()
    c2 = orc.Calculator(P=8, theta=0.5, ncrit=64)
    for _ in range(2):
        c2.advance(adv, 0.01, "rk3")
    assert rel(got_adv[:, 9:12], adv[:, 9:12]) < 1e-10 and rel(got_adv[:, 12:15], adv[:, 12:15]) < 1e-10


def test_probe_adjoint(orc):
    """Adjoint of J = sum_p v . u(p) (probes from the blob mixture): vs central finite
    differences (h = 1e-6) of the FP64 FMM forward objective with the tree and lists frozen
    (bar 1e-6; the far-field passes use the forward lists transposed, so the adjoint is the
    discrete one), and vs the exact direct-sum adjoint (which differs by the FMM's own
    far-field approximation: singular kernel beyond the MAC)."""
    X, A, S, probes, v, chk = W.blobs(n=30000, n_probe=512, n_check=8)
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    g.build()
    arr = g.arrays()
    ga, gx = g.probe_adjoint(probes, v)
    ea, ex = orc.probe_adjoint(probes, v, arr["X"], arr["alpha"], arr["sigma"])
    ra, rx = rel(ga, ea), rel(gx, ex)
    print(f"adjoint vs exact direct: dJ/dalpha {ra:.3e}  dJ/dx {rx:.3e}  t={g.timings()['adjoint_ms']:.3f} ms")
    assert ra < 2e-2 and rx < 2e-2
    # frozen-tree central differences of the FMM objective on sampled particles
    g.calculate()
    h = 1e-6
    Xt, At = arr["X"].copy(), arr["alpha"].copy()
    fd_x, fd_a, an_x, an_a = [], [], [], []
    for j in chk[:6]:
        for d in range(3):
            for which, base in (("x", Xt), ("a", At)):
                p, m = base.copy(), base.copy()
                p[j, d] += h
                m[j, d] -= h
                g.update_arrays(**({"X": p} if which == "x" else {"alpha": p}))
                g.calculate()
                jp = g.probe_objective(probes, v)
                g.update_arrays(**({"X": m} if which == "x" else {"alpha": m}))
                g.calculate()
                jm = g.probe_objective(probes, v)
                g.update_arrays(X=Xt, alpha=At)
                (fd_x if which == "x" else fd_a).append((jp - jm) / (2 * h))
                (an_x if which == "x" else an_a).append((gx if which == "x" else ga)[j, d])
    print("fd dJ/dx", np.array(fd_x)[:6], "adjoint", np.array(an_x)[:6])
    ex_, ea_ = rel(np.array(an_x), np.array(fd_x)), rel(np.array(an_a), np.array(fd_a))
    print(f"adjoint vs FD of the FMM objective: dJ/dx {ex_:.3e}  dJ/dalpha {ea_:.3e}")
    assert ex_ < 1e-6 and ea_ < 1e-6


def test_relaxation_and_diagnostics():
    """Relaxation invariants (f=0 identity; f=1 aligns alpha with curl u and keeps |alpha|)
    and diagnostics against their definitions on the downloaded state."""
    X, A, S = W.scaled("leapfrog", 16384)
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    g.build()
    g.calculate()
    arr = g.arrays()
    dg = g.diagnostics()
    x, a, u = arr["X"], arr["alpha"], arr["velocity"]
    c = np.cross(x, a)
    I_ref, A_ref = 0.5 * c.sum(0), np.cross(x, c).sum(0) / 3
    # sums of 16k terms in a different order: agreement to 1e-9 of the vector norm
    assert np.abs(dg["impulse"] - I_ref).max() < 1e-9 * np.linalg.norm(I_ref)
    assert np.abs(dg["angular_impulse"] - A_ref).max() < 1e-9 * np.linalg.norm(A_ref)
    assert abs(dg["energy"] - np.einsum("ij,ij->", u, c)) < 1e-12 * abs(dg["energy"])
    g.relax(0.0)
    assert np.array_equal(g.arrays()["alpha"], a)
    g.relax(1.0)
    a1 = g.arrays()["alpha"]
    Jm = arr["J"].reshape(-1, 3, 3)
    w = np.stack([Jm[:, 2, 1] - Jm[:, 1, 2], Jm[:, 0, 2] - Jm[:, 2, 0], Jm[:, 1, 0] - Jm[:, 0, 1]], 1)
    assert np.allclose(np.linalg.norm(a1, axis=1), np.linalg.norm(a, axis=1), rtol=1e-12)
    cosang = np.einsum("ij,ij->i", a1, w) / (np.linalg.norm(a1, axis=1) * np.linalg.norm(w, axis=1))
    assert np.allclose(cosang, 1.0, atol=1e-12)
    # a partial blend (the leapfrog workload relaxes with f = 0.3): relaxation.hpp:8-18's
    # alpha <- (1 - f) alpha + f |alpha| w / |w| with w the curl of the same evaluation
    g.update_arrays(alpha=a)
    g.calculate()
    Jm = g.arrays()["J"].reshape(-1, 3, 3)
    w = np.stack([Jm[:, 2, 1] - Jm[:, 1, 2], Jm[:, 0, 2] - Jm[:, 2, 0], Jm[:, 1, 0] - Jm[:, 0, 1]], 1)
    g.relax(0.3)
    a3 = g.arrays()["alpha"]
    want = 0.7 * a + 0.3 * np.linalg.norm(a, axis=1)[:, None] * w / np.linalg.norm(w, axis=1)[:, None]
    err = np.abs(a3 - want).max() / np.abs(a).max()
    print(f"relax f=0.3: max abs err / max |alpha| = {err:.2e}", flush=True)
    assert err < 1e-13


def test_autotune_keeps_a_candidate_and_the_results_stay_exact(orc):
    """uvlm_autotune picks one of the candidate leaf sizes; a build + eval with it matches the
    oracle run at that ncrit to the FP64 bar."""
    X, A, S = W.scaled("cube", 20000)
    g = FMMCalculator(P=8, theta=0.5, ncrit=64)
    g.set_arrays(X, A, S)
    best, ms = g.autotune(ncrits=(32, 64, 128), reps=1)
    assert best in (32, 64, 128) and ms > 0
    b_gpu = make_bodies(X, A, S)
    g.init(b_gpu)
    g.calculate()
    t = g.timings()
    assert t["ncrit"] == best and t["m2l_slots"] in (5, 6, 8)
    out = g.copyBodies2Host()
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=8, theta=0.5, ncrit=best)
    c.init(b_ref)
    c.calculate()
    ref = c.bodies(len(X))
    assert rel(out[:, 15:18], ref[:, 15:18]) < 1e-10 and rel(out[:, 0:9], ref[:, 0:9]) < 1e-10



@pytest.mark.parametrize("name", ["cube", "blobs"])
def test_near_field_option_matches_oracle(orc, name):
    """The near-field option (uvlm_set_near_field, c = 5): lists bit-exact against the oracle's
    traversal with the same option, FP64 results within 1e-10, mixed within 1e-6 of FP64, and the
    probes' lists and velocities too; c = 0 afterwards gives the reference's lists again."""
    X, A, S = CASES[name]()
    P, th, nc = PARAMS[name]
    b_ref = orc.make_bodies(X, A, S)
    c = orc.Calculator(P=P, theta=th, ncrit=nc)
    c.set_near_field(5.0)
    c.init(b_ref)
    c.calculate()
    ref = c.bodies(len(X))
    out = {}
    for prec in ("fp64", "mixed"):
        g = FMMCalculator(P=P, theta=th, ncrit=nc, precision=prec)
        g.set_near_field(5.0)
        b = make_bodies(X, A, S)
        g.init(b)
        if prec == "fp64":
            assert np.array_equal(b, b_ref)
            rp, rm = c.pairs()
            gp, gm = g.lists()
            assert np.array_equal(np.unique(rp, axis=0), gp) and np.array_equal(np.unique(rm, axis=0), gm)
            print(f"[near field {name}] {len(gp)} P2P, {len(gm)} M2L pairs", flush=True)
        g.calculate()
        out[prec] = g.copyBodies2Host()
        if prec == "fp64":
            probes = X[::97] + 0.003
            s_ref = orc.make_bodies(probes, np.zeros_like(probes), np.zeros(len(probes)))
            s_gpu = s_ref.copy()
            c.sensor_data(s_ref)
            g.getSensorData(s_gpu)
            assert np.array_equal(s_gpu[:, 9:12], s_ref[:, 9:12])
            assert rel(s_gpu[:, 15:18], s_ref[:, 15:18]) < 1e-10
            g.set_near_field(0.0)
            g.init(make_bodies(X, A, S))
            c0 = orc.Calculator(P=P, theta=th, ncrit=nc)
            c0.init(orc.make_bodies(X, A, S))
            assert np.array_equal(np.unique(c0.pairs()[1], axis=0), g.lists()[1])
    for f, sl in (("velocity", slice(15, 18)), ("J", slice(0, 9)), ("dadt", slice(18, 21))):
        e64, emx = rel(out["fp64"][:, sl], ref[:, sl]), rel(out["mixed"][:, sl], out["fp64"][:, sl])
        print(f"  {f}: fp64 vs oracle {e64:.2e}, mixed vs fp64 {emx:.2e}", flush=True)
        assert e64 < 1e-10 and emx < 1e-6
