"""ctypes wrapper of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of kursatyurt/UVLM's FMM vortex-particle path
(uvlm_oracle.cpp cites the reference file:line of every function).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the
product package (uvlm_amd) never does.

Bodies are numpy float64 arrays of shape (n, 37) laid out like exafmm::Body
(src/exafmmTypes.hpp:18-32); BODY_FIELDS gives the column slices.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# exafmm::Body field -> column slice in the (n, 37) float64 view (exafmmTypes.hpp:18-32)
BODY_FIELDS = {
    "J": slice(0, 9), "X": slice(9, 12), "alpha": slice(12, 15), "velocity": slice(15, 18),
    "dadt": slice(18, 21), "velocity_old": slice(21, 24), "dadt_old": slice(24, 27),
    "X_old": slice(27, 30), "alpha_old": slice(30, 33), "radius": 33, "drdt": 34,
    "radius_old": 35, "drdt_old": 36,
}
BODY_DOUBLES = 37

STRETCH = {"transposed": 0, "classic": 1}
FORM = {"cvpm": 0, "rvpm": 1}
VISCOUS = {"inviscid": 0, "pse": 1, "csm": 2}
SCHEME = {"euler": 0, "rk2": 1, "rk3": 2, "rk4": 3}

_lib = None


def build():
    """Compile liboracle.so with its Makefile (g++)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, D, I, L64, VP = (ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_int,
                            ctypes.c_int64, ctypes.c_void_p)
        P64 = ctypes.POINTER(ctypes.c_int64)
        sig = {
            "orc_body_size": (I, []),
            "orc_create": (VP, [I, D, L64, I, I]),
            "orc_destroy": (None, [VP]),
            "orc_init": (I, [VP, P, L64]),
            "orc_calculate": (I, [VP]),
            "orc_get_bodies": (I, [VP, P, L64]),
            "orc_advance": (I, [VP, P, L64, D, I]),
            "orc_stage": (I, [VP, P, L64, D, I, I]),
            "orc_direct_vjp": (None, [L64, P, P, P, P, P, I, P, P]),
            "orc_set_viscosity": (None, [VP, I, D]),
            "orc_viscosity": (I, [VP, D]),
            "orc_vorticity": (I, [VP, P, L64]),
            "orc_rbf": (I, [VP, P, L64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), P]),
            "orc_sensor_data": (I, [VP, P, L64]),
            "orc_num_cells": (L64, [VP]),
            "orc_get_cells": (None, [VP, P64, P]),
            "orc_num_sensor_cells": (L64, [VP]),
            "orc_get_sensor_cells": (None, [VP, P64, P]),
            "orc_num_pairs": (None, [VP, I, P64, P64]),
            "orc_get_pairs": (None, [VP, I, P64, P64]),
            "orc_get_expansions": (None, [VP, P, P]),
            "orc_build_tree": (L64, [P, L64, L64, P64, P, L64, P]),
            "orc_eval_multipole": (None, [I, D, D, D, P, P]),
            "orc_eval_local": (None, [I, D, D, D, P]),
            "orc_p2m": (None, [I, P, L64, P, P, P]),
            "orc_m2m": (None, [I, P, P, P, P]),
            "orc_m2l": (None, [I, P, P, P, P]),
            "orc_m2l_order": (None, [I, I, P, P, P, P]),
            "orc_set_near_gap": (None, [VP, D]),
            "orc_l2l": (None, [I, P, P, P, P]),
            "orc_l2p": (None, [I, P, P, L64, P, P, P]),
            "orc_direct": (None, [L64, P, L64, P, P, P, P, P]),
            "orc_direct_classic": (None, [L64, P, P, L64, P, P, P, P, P]),
            "orc_direct_tsigma": (None, [L64, P, P, L64, P, P, P, P]),
            "orc_max_threads": (I, []),
            "orc_set_threads": (None, [I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        assert L.orc_body_size() == 8 * BODY_DOUBLES
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _p64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def nterm(P):
    return 3 * (P * (P + 1) // 2)


def make_bodies(X, alpha, sigma):
    """(n,37) Body array with X, alpha, radius set and every other field zero."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    b = np.zeros((n, BODY_DOUBLES))
    b[:, BODY_FIELDS["X"]] = X
    b[:, BODY_FIELDS["alpha"]] = alpha
    b[:, BODY_FIELDS["radius"]] = sigma
    return b


class Calculator:
    """Mirror of Vortex::FMMCalculator (src/fmmCalculator.hpp:25-1006) on the CPU."""

    def __init__(self, P=4, theta=0.3, ncrit=128, stretching="transposed", formulation="cvpm"):
        self.P, self.theta, self.ncrit = P, theta, ncrit
        self._h = lib().orc_create(P, theta, ncrit, STRETCH[stretching], FORM[formulation])
        if not self._h:
            raise ValueError("bad oracle parameters")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_near_field(self, c):
        """Near-field option: pairs whose bodies can come closer than c * max sigma are near (P2P);
        0 is the reference's acceptance test alone.  Applies from the next init."""
        lib().orc_set_near_gap(self._h, float(c))

    def init(self, bodies):
        """Builds tree + lists; permutes `bodies` (n,37) in place like FMMCalculator::init."""
        assert bodies.flags.c_contiguous and bodies.dtype == np.float64
        st = lib().orc_init(self._h, _p(bodies), bodies.shape[0])
        if st:
            raise RuntimeError(f"orc_init failed ({st})")

    def calculate(self):
        st = lib().orc_calculate(self._h)
        if st:
            raise RuntimeError(f"orc_calculate failed ({st})")

    def bodies(self, n):
        out = np.zeros((n, BODY_DOUBLES))
        st = lib().orc_get_bodies(self._h, _p(out), n)
        if st:
            raise RuntimeError("size mismatch")
        return out

    def advance(self, bodies, dt, scheme="rk3"):
        st = lib().orc_advance(self._h, _p(bodies), bodies.shape[0], dt, SCHEME[scheme])
        if st:
            raise RuntimeError(f"orc_advance failed ({st})")

    def stage(self, bodies, dt, scheme, k):
        """Stage k of advance() from the full body state in `bodies` (tree order out)."""
        st = lib().orc_stage(self._h, _p(bodies), bodies.shape[0], dt, SCHEME[scheme], k)
        if st:
            raise RuntimeError(f"orc_stage failed ({st})")

    def set_viscosity(self, scheme, nu):
        """scheme: "inviscid", "pse" or "csm"; applied after calculate() in advance/stage."""
        lib().orc_set_viscosity(self._h, VISCOUS[scheme], float(nu))

    def viscosity(self, dt):
        """calculateViscosity on the current bodies (after calculate())."""
        st = lib().orc_viscosity(self._h, float(dt))
        if st:
            raise RuntimeError(f"orc_viscosity failed ({st})")

    def vorticity(self, sensors):
        """Gaussian vorticity (negated, into velocity_old) at the probes, near pairs only."""
        st = lib().orc_vorticity(self._h, _p(sensors), sensors.shape[0])
        if st:
            raise RuntimeError(f"orc_vorticity failed ({st})")

    def rbf(self, points):
        """RBF re-initialisation onto `points` (n, 37); returns (kept, iterations, residuals):
        points[:kept] hold the kept points in tree order with the fitted alpha."""
        kept = ctypes.c_int64()
        it = ctypes.c_int()
        res = np.zeros(3)
        st = lib().orc_rbf(self._h, _p(points), points.shape[0], ctypes.byref(kept), ctypes.byref(it), _p(res))
        if st:
            raise RuntimeError(f"orc_rbf failed ({st})")
        return kept.value, it.value, res

    def sensor_data(self, sensors):
        st = lib().orc_sensor_data(self._h, _p(sensors), sensors.shape[0])
        if st:
            raise RuntimeError(f"orc_sensor_data failed ({st})")

    def cells(self, sensor=False):
        L = lib()
        C = L.orc_num_sensor_cells(self._h) if sensor else L.orc_num_cells(self._h)
        ints = np.zeros((C, 5), dtype=np.int64)
        reals = np.zeros((C, 4))
        (L.orc_get_sensor_cells if sensor else L.orc_get_cells)(self._h, _p64(ints), _p(reals))
        return ints, reals

    def pairs(self, sensor=False):
        L = lib()
        a, b = ctypes.c_int64(), ctypes.c_int64()
        L.orc_num_pairs(self._h, int(sensor), ctypes.byref(a), ctypes.byref(b))
        p2p = np.zeros((a.value, 2), dtype=np.int64)
        m2l = np.zeros((b.value, 2), dtype=np.int64)
        L.orc_get_pairs(self._h, int(sensor), _p64(p2p), _p64(m2l))
        return p2p, m2l

    def expansions(self):
        C = lib().orc_num_cells(self._h)
        M = np.zeros((C, nterm(self.P)), dtype=np.complex128)
        Lc = np.zeros((C, nterm(self.P)), dtype=np.complex128)
        lib().orc_get_expansions(self._h, _p(M.view(np.float64)), _p(Lc.view(np.float64)))
        return M, Lc


def build_tree(bodies, ncrit):
    """Standalone tree build; permutes `bodies` in place.  Returns (ints, reals, X0, R0)."""
    L = lib()
    n = bodies.shape[0]
    cap = max(16, 4 * n // max(1, ncrit) + 64)
    while True:
        ints = np.zeros((cap, 5), dtype=np.int64)
        reals = np.zeros((cap, 4))
        x0r0 = np.zeros(4)
        work = bodies.copy()
        C = L.orc_build_tree(_p(work), n, ncrit, _p64(ints), _p(reals), cap, _p(x0r0))
        if C < 0:
            raise RuntimeError(f"orc_build_tree failed ({-C})")
        if C <= cap:
            bodies[:] = work
            return ints[:C], reals[:C], x0r0[:3].copy(), float(x0r0[3])
        cap = int(C)


def eval_multipole(P, rho, alpha, beta):
    Y = np.zeros(P * P, dtype=np.complex128)
    Yt = np.zeros(P * P, dtype=np.complex128)
    lib().orc_eval_multipole(P, rho, alpha, beta, _p(Y.view(np.float64)), _p(Yt.view(np.float64)))
    return Y, Yt


def eval_local(P, rho, alpha, beta):
    Y = np.zeros(P * P, dtype=np.complex128)
    lib().orc_eval_local(P, rho, alpha, beta, _p(Y.view(np.float64)))
    return Y


def p2m(P, Xc, X, alpha):
    M = np.zeros(nterm(P), dtype=np.complex128)
    lib().orc_p2m(P, _p(_c(Xc)), X.shape[0], _p(_c(X)), _p(_c(alpha)), _p(M.view(np.float64)))
    return M


def m2m(P, Xp, Xc, Mc):
    Mp = np.zeros(nterm(P), dtype=np.complex128)
    lib().orc_m2m(P, _p(_c(Xp)), _p(_c(Xc)), _p(_c(Mc, np.complex128).view(np.float64)), _p(Mp.view(np.float64)))
    return Mp


def m2l(P, Xt, Xs, Ms, harmonic_order=None):
    """M2L of one pair.  harmonic_order (tests only) fills the singular harmonics up to degree
    harmonic_order-1 instead of P-1: a deliberately different truncation."""
    Lt = np.zeros(nterm(P), dtype=np.complex128)
    if harmonic_order is None:
        lib().orc_m2l(P, _p(_c(Xt)), _p(_c(Xs)), _p(_c(Ms, np.complex128).view(np.float64)), _p(Lt.view(np.float64)))
    else:
        lib().orc_m2l_order(P, int(harmonic_order), _p(_c(Xt)), _p(_c(Xs)),
                            _p(_c(Ms, np.complex128).view(np.float64)), _p(Lt.view(np.float64)))
    return Lt


def l2l(P, Xp, Xc, Lp):
    Lc = np.zeros(nterm(P), dtype=np.complex128)
    lib().orc_l2l(P, _p(_c(Xp)), _p(_c(Xc)), _p(_c(Lp, np.complex128).view(np.float64)), _p(Lc.view(np.float64)))
    return Lc


def l2p(P, Xc, L, X):
    X = _c(X)
    n = X.shape[0]
    u = np.zeros((n, 3))
    J = np.zeros((n, 9))
    lib().orc_l2p(P, _p(_c(Xc)), _p(_c(L, np.complex128).view(np.float64)), n, _p(X), _p(u), _p(J))
    return u, J


def direct(xt, xs, alpha_s, sigma_s, want_J=True):
    """FP64 O(M*N) direct sum (device P2P formula): u (m,3) and J (m,9) at targets."""
    xt, xs, a, s = _c(xt), _c(xs), _c(alpha_s), _c(sigma_s)
    m = xt.shape[0]
    u = np.zeros((m, 3))
    J = np.zeros((m, 9)) if want_J else None
    lib().orc_direct(m, _p(xt), xs.shape[0], _p(xs), _p(a), _p(s), _p(u), _p(J) if want_J else None)
    return (u, J) if want_J else u


def direct_classic(xt, at, xs, alpha_s, sigma_s):
    """FP64 direct sum with the reference's host formula: u and classic J.alpha_target."""
    xt, at, xs, a, s = _c(xt), _c(at), _c(xs), _c(alpha_s), _c(sigma_s)
    m = xt.shape[0]
    u = np.zeros((m, 3))
    dadt = np.zeros((m, 3))
    lib().orc_direct_classic(m, _p(xt), _p(at), xs.shape[0], _p(xs), _p(a), _p(s), _p(u), _p(dadt))
    return u, dadt


def direct_vjp(X, alpha, sigma, ubar, abar, stretching="transposed"):
    """Exact VJP of the direct-sum rates (u, stretching dadt = J^T alpha, or J alpha when
    stretching == "classic") with cotangents ubar, abar (n, 3): returns (xbar, alphabar)."""
    X, A, S, ub, ab = _c(X), _c(alpha), _c(sigma), _c(ubar), _c(abar)
    n = X.shape[0]
    xb, alb = np.zeros((n, 3)), np.zeros((n, 3))
    lib().orc_direct_vjp(n, _p(X), _p(A), _p(S), _p(ub), _p(ab), int(stretching == "classic"), _p(xb), _p(alb))
    return xb, alb


def probe_adjoint(probes, weights, X, alpha, sigma):
    """Exact (direct-sum) gradient of J = sum_p weights_p . u(probes_p), u from the particles
    (X, alpha, sigma): returns dJ/dalpha (n,3), dJ/dX (n,3)."""
    X, A, S = _c(X), _c(alpha), _c(sigma)
    Pp = _c(probes)
    W = _c(np.broadcast_to(weights, Pp.shape))
    n = X.shape[0]
    u = np.zeros((n, 3))
    J = np.zeros((n, 9))
    lib().orc_direct_tsigma(n, _p(X), _p(S), Pp.shape[0], _p(Pp), _p(W), _p(u), _p(J))
    gx = np.einsum("iab,ia->ib", J.reshape(-1, 3, 3), A)
    return u, gx


def probe_objective(probes, weights, X, alpha, sigma):
    """J = sum_p weights_p . u(probes_p) by the direct sum (device P2P formula)."""
    u = direct(probes, X, alpha, sigma, want_J=False)
    return float(np.sum(np.broadcast_to(weights, u.shape) * u))


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()
