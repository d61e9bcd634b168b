"""Host-side checks that need no GPU: the C ABI library loads and exports every entry
point of include/uvlm_fmm.h, argument errors are reported before any device call, and
the seeded workload generators produce the configured shapes."""
import ctypes
import os
import re

import numpy as np
import pytest

from uvlm_amd import _capi, workloads as W
from uvlm_amd.fmm import BODY_DOUBLES, BODY_FIELDS, make_bodies

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "uvlm_fmm.h")).read()
    declared = set(re.findall(r"\b(uvlm_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(_capi.EXPORTS)
    L = _capi.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.uvlm_abi_version() == 1


@pytest.mark.parametrize("field,value", [("P", 1), ("P", 17), ("theta", 0.0), ("ncrit", 0), ("precision", 7),
                                         ("scheme", 9), ("stretching", 3), ("formulation", 2)])
def test_create_rejects_bad_params_without_touching_the_device(field, value):
    L = _capi.lib()
    prm = _capi.Params(P=8, theta=0.5, ncrit=64, precision=0, stretching=0, formulation=0, scheme=2, device=0)
    setattr(prm, field, value)
    h = ctypes.c_void_p()
    rc = L.uvlm_create(ctypes.byref(prm), ctypes.byref(h))
    assert rc == 1 and not h.value
    assert L.uvlm_last_error().decode()


def test_null_arguments_are_errors():
    L = _capi.lib()
    assert L.uvlm_create(None, None) == 1
    assert L.uvlm_timings(None, None) == 1
    assert L.uvlm_destroy(None) == 0
    assert L.uvlm_build(None) == 1
    assert L.uvlm_eval(None) == 1
    assert L.uvlm_autotune(None, None, 0, 1, None, None) == 1


def test_body_layout_matches_exafmm_body():
    assert BODY_DOUBLES * 8 == 296
    cols = []
    for v in BODY_FIELDS.values():
        cols += list(range(v.start, v.stop)) if isinstance(v, slice) else [v]
    assert sorted(cols) == list(range(37))
    b = make_bodies(np.ones((3, 3)), 2 * np.ones((3, 3)), 3 * np.ones(3))
    assert (b[:, 9:12] == 1).all() and (b[:, 12:15] == 2).all() and (b[:, 33] == 3).all()


def test_generators_are_seeded_and_shaped():
    X, A, S = W.ring()
    assert X.shape == (8192, 3) and np.allclose(S, 0.1)
    assert np.allclose(np.linalg.norm(A, axis=1), 2 * np.pi / 64 / 128)
    X2, _, _ = W.ring()
    assert np.array_equal(X, X2)
    Xc, Ac, Sc, idx = W.cube(n=4096, n_sample=100)
    h = 2 / 4096 ** (1 / 3)
    assert Xc.min() >= -1 and Xc.max() <= 1 and np.allclose(Sc, 2 * h) and len(set(idx.tolist())) == 100
    Xl, Al, Sl = W.leapfrog(n_stations=64, samples=8)
    assert Xl.shape == (1024, 3) and np.allclose(Sl, 0.08)
    assert abs(Xl[:512, 2].mean()) < 0.05 and abs(Xl[512:, 2].mean() - 0.5) < 0.05
    Xb, Ab, Sb, probes, v, chk = W.blobs(n=3000, n_probe=30, n_check=5)
    assert Xb.shape == (3000, 3) and probes.shape == (30, 3) and abs(np.linalg.norm(v) - 1) < 1e-12


def test_splitmix_reference_values():
    # splitmix64 seeded with 0: the first outputs of the published generator
    r = W.SplitMix64(0)
    z = r._next_u64(2)
    assert int(z[0]) == 0xE220A8397B1DCDAF and int(z[1]) == 0x6E789E6AA1B965F4


def _g_series(x, der):
    """d^der/dx^der of G(x) = (2/sqrt(pi)) sum (-1)^i 2(i+1) / ((2i+3)(i+1)!) x^i (x <= 2)."""
    import math
    tot = 0.0
    for i in range(der, 60):
        c = 2 / math.sqrt(math.pi) * (-1) ** i * 2 * (i + 1) / ((2 * i + 3) * math.factorial(i + 1))
        tot += c * math.factorial(i) / math.factorial(i - der) * x ** (i - der)
    return tot


def test_p2p_polynomial_matches_erf_form():
    """The FP32 P2P's polynomial for G(x) = g(s)/s^3 (x = s^2, g = erf(s) - 2 s e^{-s^2}/sqrt(pi))
    and its first two derivatives, read from pair_kernel.hpp, against the erf form evaluated in
    FP64 and (G'') the term-wise second derivative of G's power series."""
    import math
    src = open(os.path.join(ROOT, "uvlm_amd", "csrc", "pair_kernel.hpp")).read()
    body = src[src.index("c_gpoly[NPOLY] = {") + 18: src.index("};", src.index("c_gpoly[NPOLY] = {"))]
    coef = [float(v.strip().rstrip("f")) for v in body.split(",")]
    assert len(coef) == 10
    for x in np.linspace(0.05, 2.0, 60):
        s = math.sqrt(x)
        g = math.erf(s) - 2 * s * math.exp(-x) / math.sqrt(math.pi)
        G = g / s ** 3
        # derivative: G'(x) = (s g'(s) - 3 g) / (2 s^5) with g'(s) = 4 s^2 e^{-s^2} / sqrt(pi)
        Gp = (s * 4 * x * math.exp(-x) / math.sqrt(math.pi) - 3 * g) / (2 * s ** 5)
        t = x - 1.0
        p = sum(c * t ** i for i, c in enumerate(coef))
        pp = sum(i * c * t ** (i - 1) for i, c in enumerate(coef) if i)
        ppp = sum(i * (i - 1) * c * t ** (i - 2) for i, c in enumerate(coef) if i > 1)
        assert abs(p - G) < 1e-9 * abs(G) + 1e-12
        assert abs(pp - Gp) < 2e-8 * 0.3
        assert abs(_g_series(x, 1) - Gp) < 1e-12
        assert abs(ppp - _g_series(x, 2)) < 1e-7


def test_p2p_near_polynomials_match_erf_form():
    """The forward P2P's degree-7 polynomials for G and G' (c_gnear and c_gpnear in
    pair_kernel.hpp), evaluated as the kernel does (Horner in FP32, each chain started after its
    first step), against the erf form in FP64 and (small x, where the erf form cancels) G's power
    series; and G' against the derivative of G's polynomial, so that a swapped or mistyped
    coefficient in either table fails."""
    import math
    src = open(os.path.join(ROOT, "uvlm_amd", "csrc", "pair_kernel.hpp")).read()

    def table(name):
        m = re.search(re.escape(name) + r"\[NNEAR\]\s*=\s*\{([^}]*)\}", src)
        body = m.group(1)
        return np.array([float(v.strip().rstrip("f")) for v in body.split(",")], dtype=np.float32)

    cg, cp = table("c_gnear"), table("c_gpnear")
    assert len(cg) == 8 and len(cp) == 8
    f32 = np.float32
    for x in np.linspace(0.0, 2.0, 201):
        t = f32(f32(x) - f32(1.0))
        g = f32(cg[7] * t + cg[6])
        p = f32(cp[7] * t + cp[6])
        for k in range(5, -1, -1):
            g = f32(g * t + cg[k])
            p = f32(p * t + cp[k])
        if x < 0.05:
            G, Gp = _g_series(x, 0), _g_series(x, 1)
        else:
            s = math.sqrt(x)
            gr = math.erf(s) - 2 * s * math.exp(-x) / math.sqrt(math.pi)
            G = gr / s ** 3
            Gp = (s * 4 * x * math.exp(-x) / math.sqrt(math.pi) - 3 * gr) / (2 * s ** 5)
        assert abs(float(g) - G) < 2e-7 * abs(G)
        assert abs(float(p) - Gp) < 1.2e-7
        # G' against the derivative of the G polynomial (exact arithmetic; the derivative of
        # the degree-7 interpolant is within 7e-7 of G' at the ends of [0, 2])
        dg = sum(i * float(c) * (x - 1.0) ** (i - 1) for i, c in enumerate(cg) if i)
        assert abs(dg - Gp) < 1e-6


def test_p2p_far_form_matches_erf_form():
    """The FP32 P2P's form for x = s^2 >= 2, read from near.hip: G = (1 - e (erfcx(s) +
    2s/sqrt(pi))) w^3 and G' = w^2 (2e/sqrt(pi) - 1.5 G) with w = 1/s, e = e^{-x} and erfcx by
    the c_far polynomial in w, and G'' = -(2/sqrt(pi)) e (w^2 + w^4) - 1.5 w^2 (G' - G w^2), all
    read from pair_kernel.hpp, against the erf form evaluated in FP64 (G'' by differentiating
    G' = (2/sqrt(pi)) e / x - 1.5 G / x once more, and by the power series at x = 2)."""
    import math
    src = open(os.path.join(ROOT, "uvlm_amd", "csrc", "pair_kernel.hpp")).read()
    body = src[src.index("c_far[NFAR] = {") + 15: src.index("};", src.index("c_far[NFAR] = {"))]
    coef = [float(v.strip().rstrip("f")) for v in body.split(",")]
    wc = float(re.search(r"FAR_WC\s*=\s*([0-9.eE+-]+)f", src).group(1))
    wmin = float(re.search(r"FAR_WMIN\s*=\s*([0-9.eE+-]+)f", src).group(1))
    assert len(coef) == 8
    for x in list(np.linspace(2.0, 30.0, 300)) + [50.0, 200.0, 1e4]:
        s = math.sqrt(x)
        w = 1 / s
        e = math.exp(-x)
        t = max(w, wmin) - wc
        p = sum(c * t ** i for i, c in enumerate(coef))
        g = 1 - e * (p + 2 * s / math.sqrt(math.pi))
        G = g * w ** 3
        Gp = w * w * (2 * e / math.sqrt(math.pi) - 1.5 * G)
        gr = math.erf(s) - 2 * s * e / math.sqrt(math.pi)
        Gr = gr / s ** 3
        Gpr = (s * 4 * x * e / math.sqrt(math.pi) - 3 * gr) / (2 * s ** 5)
        assert abs(G - Gr) < 2e-8 * abs(Gr)
        assert abs(Gp - Gpr) < 5e-8 * abs(Gpr)
        Gpp = -2 / math.sqrt(math.pi) * e * (w ** 2 + w ** 4) - 1.5 * w * w * (Gp - G * w * w)
        Gppr = -2 / math.sqrt(math.pi) * e * (1 / x + 1 / x ** 2) - 1.5 * Gpr / x + 1.5 * Gr / x ** 2
        assert abs(Gpp - Gppr) < 2e-7 * abs(Gppr)
    assert abs(-2 / math.sqrt(math.pi) * math.exp(-2.0) * 0.75 - 1.5 * 0.5 * (_g_series(2.0, 1) - 0.5 * _g_series(2.0, 0))
               - _g_series(2.0, 2)) < 1e-12


def test_m2l_plan_owns_every_term_once(tmp_path):
    """tests/cpp/m2l_plan_check.cpp, compiled with g++ and run: the M2L work plan for every P in
    1..16, group width 16/32 and slot budget owns each truncated M2L term exactly once and keeps
    every index in bounds."""
    import subprocess
    exe = str(tmp_path / "m2l_plan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "cpp", "m2l_plan_check.cpp")],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "all terms owned exactly once, in bounds" in r.stdout


def _unbound_names(path):
    """(function, name, line) for every name a top-level function of the module loads that is
    bound neither in that function (arguments, assignments, imports, nested definitions) nor at
    module level nor as a builtin: a NameError waiting for the line to run."""
    import ast
    import builtins

    tree = ast.parse(open(path).read())

    def bound(nodes):
        b = set()
        for top in nodes:
            for n in ast.walk(top):
                if isinstance(n, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)):
                    b.add(n.name)
                elif isinstance(n, (ast.Import, ast.ImportFrom)):
                    b.update((a.asname or a.name).split(".")[0] for a in n.names)
                elif isinstance(n, ast.Name) and isinstance(n.ctx, (ast.Store, ast.Del)):
                    b.add(n.id)
                elif isinstance(n, ast.arg):
                    b.add(n.arg)
                elif isinstance(n, ast.ExceptHandler) and n.name:
                    b.add(n.name)
                elif isinstance(n, (ast.Global, ast.Nonlocal)):
                    b.update(n.names)
        return b

    module = set(dir(builtins)) | {"__file__", "__name__"}
    for n in tree.body:
        if isinstance(n, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)):
            module.add(n.name)
        else:
            module |= bound([n])
    bad = []
    for f in tree.body:
        if isinstance(f, (ast.FunctionDef, ast.AsyncFunctionDef)):
            local = bound([f])
            bad += [(f.name, n.id, n.lineno) for n in ast.walk(f)
                    if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load) and n.id not in local | module]
    return bad


@pytest.mark.parametrize("rel", ["bench.py", "__graft_entry__.py", "uvlm_amd/fmm.py", "uvlm_amd/vlm.py",
                                 "uvlm_amd/_capi.py", "uvlm_amd/workloads.py", "uvlm_amd/vtk.py"])
def test_no_unbound_names(rel):
    """bench.py and the package run only on the GPU box; a name left over from an edit would
    first fail there.  Every top-level function loads only names it or the module binds."""
    assert _unbound_names(os.path.join(ROOT, rel)) == []
