"""Python mirror of Vortex::FMMCalculator over the C ABI of libuvlm_fmm.so.

The reference interface (src/fmmCalculator.hpp:25-1006) is a C++ class template whose
tags select the integrator, viscous scheme, formulation and stretching; here they are
constructor arguments.  Method names and argument meaning follow the reference:

  init(bodies)          -> fmmCalculator.hpp:282-288 (tree + lists; permutes `bodies`)
  calculate()           -> fmmCalculator.hpp:420-467 (velocity, J, stretching)
  advance(bodies, dt)   -> fmmCalculator.hpp:919-1005 (device-resident RK step)
  getSensorData(probes) -> fmmCalculator.hpp:726-731
  copyBodies2Host(bodies) -> fmmCalculator.hpp:235-243

Bodies are float64 arrays of shape (n, 37) laid out like exafmm::Body
(src/exafmmTypes.hpp:18-32); BODY_FIELDS maps field names to columns.  Errors of the C
ABI raise UVLMError.  The HIP library is required: there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _capi

BODY_FIELDS = {
    "J": slice(0, 9), "X": slice(9, 12), "alpha": slice(12, 15), "velocity": slice(15, 18),
    "dadt": slice(18, 21), "velocity_old": slice(21, 24), "dadt_old": slice(24, 27),
    "X_old": slice(27, 30), "alpha_old": slice(30, 33), "radius": 33, "drdt": 34,
    "radius_old": 35, "drdt_old": 36,
}
BODY_DOUBLES = 37
BODY_BYTES = 8 * BODY_DOUBLES

PRECISION = {"fp64": 0, "mixed": 1}
STRETCHING = {"transposed": 0, "classic": 1}
FORMULATION = {"cvpm": 0, "rvpm": 1}
VISCOUS = {"inviscid": 0, "pse": 1, "csm": 2}
SCHEME = {"euler": 0, "rk2": 1, "rk3": 2, "rk4": 3}

UVLMError = _capi.UVLMError


def make_bodies(X, alpha, sigma):
    """(n, 37) exafmm::Body array with X, alpha and radius set, every other field zero."""
    X = np.asarray(X, dtype=np.float64)
    b = np.zeros((X.shape[0], BODY_DOUBLES))
    b[:, BODY_FIELDS["X"]] = X
    b[:, BODY_FIELDS["alpha"]] = alpha
    b[:, BODY_FIELDS["radius"]] = sigma
    return b


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class FMMCalculator:
    """One device-resident particle set and its FMM (a uvlm_handle)."""

    def __init__(self, P=4, theta=0.3, ncrit=128, precision="fp64", stretching="transposed",
                 formulation="cvpm", scheme="rk3", device=0):
        self.lib = _capi.lib()
        prm = _capi.Params(P=P, theta=theta, ncrit=ncrit, precision=PRECISION[precision],
                           stretching=STRETCHING[stretching], formulation=FORMULATION[formulation],
                           scheme=SCHEME[scheme], device=device)
        self.P, self.theta, self.ncrit = P, theta, ncrit
        self.nterm = 3 * (P * (P + 1) // 2)
        self._h = ctypes.c_void_p()
        _capi.check(self.lib.uvlm_create(ctypes.byref(prm), ctypes.byref(self._h)))
        self.n = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.uvlm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- bodies ---------------------------------------------------------------------
    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies, dtype=np.float64)
        assert b.ndim == 2 and b.shape[1] == BODY_DOUBLES
        _capi.check(self.lib.uvlm_set_bodies_aos(self._h, _dp(b), b.shape[0], BODY_BYTES))
        self.n = b.shape[0]

    def set_arrays(self, X, alpha, sigma):
        X = np.ascontiguousarray(X, dtype=np.float64)
        A = np.ascontiguousarray(alpha, dtype=np.float64)
        S = np.ascontiguousarray(sigma, dtype=np.float64)
        _capi.check(self.lib.uvlm_set_bodies_soa(self._h, X.shape[0], _dp(X), _dp(A), _dp(S)))
        self.n = X.shape[0]

    def append_arrays(self, X, alpha, sigma):
        """Append particles after the current ones (next original indices)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        A = np.ascontiguousarray(alpha, dtype=np.float64)
        S = np.ascontiguousarray(sigma, dtype=np.float64)
        _capi.check(self.lib.uvlm_append_bodies_soa(self._h, X.shape[0], _dp(X), _dp(A), _dp(S)))
        self.n += X.shape[0]

    def copyBodies2Host(self, bodies=None):
        out = bodies if bodies is not None else np.zeros((self.n, BODY_DOUBLES))
        assert out.flags.c_contiguous and out.shape == (self.n, BODY_DOUBLES)
        _capi.check(self.lib.uvlm_get_bodies_aos(self._h, _dp(out), self.n, BODY_BYTES))
        return out

    def arrays(self):
        """Current (tree-order) X, alpha, sigma, velocity, J (n,9), dadt, drdt."""
        n = self.n
        X, A, U, D = (np.zeros((n, 3)) for _ in range(4))
        S, DR = np.zeros(n), np.zeros(n)
        J = np.zeros((n, 9))
        _capi.check(self.lib.uvlm_get_bodies_soa(self._h, n, _dp(X), _dp(A), _dp(S), _dp(U), _dp(J), _dp(D),
                                                 _dp(DR)))
        return {"X": X, "alpha": A, "sigma": S, "velocity": U, "J": J, "dadt": D, "drdt": DR}

    def permutation(self):
        p = np.zeros(self.n, dtype=np.int64)
        _capi.check(self.lib.uvlm_get_permutation(self._h, self.n, _ip(p)))
        return p

    # -- the reference's operations -------------------------------------------------
    def init(self, bodies):
        """Upload, build tree + lists on the device, write the permuted bodies back."""
        self.set_bodies(bodies)
        self.build()
        self.copyBodies2Host(bodies)

    def build(self):
        _capi.check(self.lib.uvlm_build(self._h))

    def calculate(self):
        _capi.check(self.lib.uvlm_eval(self._h))

    def stage(self, k, dt):
        _capi.check(self.lib.uvlm_stage(self._h, int(k), float(dt)))

    def advance(self, bodies, dt):
        self.set_bodies(bodies)
        _capi.check(self.lib.uvlm_advance(self._h, float(dt)))
        self.copyBodies2Host(bodies)

    def step(self, dt):
        """A device-resident time step on the bodies already on the device."""
        _capi.check(self.lib.uvlm_advance(self._h, float(dt)))

    def getSensorData(self, sensors):
        s = sensors
        assert s.flags.c_contiguous and s.dtype == np.float64 and s.shape[1] == BODY_DOUBLES
        _capi.check(self.lib.uvlm_sensors_aos(self._h, _dp(s), s.shape[0], BODY_BYTES))

    def direct(self, xt, at=None, want_J=True, noalloc=False):
        """FP64 direct sum from the current bodies to points xt: (u, J or None, dadt or None).
        noalloc=True calls uvlm_direct_noalloc (static device staging, at most 16,384 targets)."""
        xt = np.ascontiguousarray(xt, dtype=np.float64)
        m = xt.shape[0]
        u = np.zeros((m, 3))
        J = np.zeros((m, 9)) if want_J else None
        D = np.zeros((m, 3)) if at is not None else None
        a = np.ascontiguousarray(at, dtype=np.float64) if at is not None else None
        fn = self.lib.uvlm_direct_noalloc if noalloc else self.lib.uvlm_direct
        _capi.check(fn(self._h, m, _dp(xt), _dp(a) if a is not None else None, _dp(u),
                       _dp(J) if J is not None else None, _dp(D) if D is not None else None))
        return u, J, D

    def update_arrays(self, X=None, alpha=None):
        """Overwrite positions/strengths in the current tree order, keeping tree and lists."""
        X = None if X is None else np.ascontiguousarray(X, dtype=np.float64)
        A = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        _capi.check(self.lib.uvlm_update_bodies_soa(self._h, self.n, _dp(X) if X is not None else None,
                                                    _dp(A) if A is not None else None))

    def probe_adjoint(self, probes, weights):
        """dJ/dalpha, dJ/dx (tree order) of J = sum_p weights_p . u(probes_p)."""
        Pp = np.ascontiguousarray(probes, dtype=np.float64)
        W = np.ascontiguousarray(np.broadcast_to(weights, Pp.shape), dtype=np.float64)
        ga = np.zeros((self.n, 3))
        gx = np.zeros((self.n, 3))
        _capi.check(self.lib.uvlm_probe_adjoint(self._h, Pp.shape[0], _dp(Pp), _dp(W), _dp(ga), _dp(gx)))
        return ga, gx

    def eval_vjp(self, ubar=None, abar=None):
        """(xbar, alphabar), tree order, of S = sum ubar.u + abar.dadt for the current tree held
        fixed (evaluates the rates first)."""
        ub = None if ubar is None else np.ascontiguousarray(ubar, dtype=np.float64)
        ab = None if abar is None else np.ascontiguousarray(abar, dtype=np.float64)
        xb = np.zeros((self.n, 3))
        alb = np.zeros((self.n, 3))
        _capi.check(self.lib.uvlm_eval_vjp(self._h, self.n, _dp(ub) if ub is not None else None,
                                           _dp(ab) if ab is not None else None, _dp(xb), _dp(alb)))
        return xb, alb

    def step_vjp(self, dt, xbar, alphabar):
        """Cotangent of the state before one advance(dt) from that after it (original order)."""
        xb = np.array(xbar, dtype=np.float64, order="C", copy=True)
        ab = np.array(alphabar, dtype=np.float64, order="C", copy=True)
        _capi.check(self.lib.uvlm_step_vjp(self._h, float(dt), self.n, _dp(xb), _dp(ab)))
        return xb, ab

    def trajectory_forward(self, steps, dt, every=1):
        """Advance `steps` steps, keeping checkpoints every `every` steps on the device."""
        _capi.check(self.lib.uvlm_trajectory_forward(self._h, int(steps), float(dt), int(every)))

    def trajectory_vjp(self, xbar, alphabar):
        """Cotangent of the initial state from that of the final one (original order)."""
        xb = np.array(xbar, dtype=np.float64, order="C", copy=True)
        ab = np.array(alphabar, dtype=np.float64, order="C", copy=True)
        _capi.check(self.lib.uvlm_trajectory_vjp(self._h, self.n, _dp(xb), _dp(ab)))
        return xb, ab

    def probe_objective(self, probes, weights):
        """J = sum_p weights_p . u(probes_p) through getSensorData (needs a calculate())."""
        from .fmm import make_bodies as _mk
        s = _mk(probes, np.broadcast_to(weights, np.shape(probes)), np.ones(len(probes)))
        self.getSensorData(s)
        return float(np.sum(s[:, BODY_FIELDS["alpha"]] * s[:, BODY_FIELDS["velocity"]]))

    def relax(self, f):
        """Pedrizzetti relaxation of the strengths toward the last eval's vorticity."""
        _capi.check(self.lib.uvlm_relax(self._h, float(f)))

    def diagnostics(self):
        """Linear impulse (3), angular impulse (3), energy, total strength (3)."""
        out = np.zeros(10)
        _capi.check(self.lib.uvlm_diagnostics(self._h, _dp(out)))
        return {"impulse": out[0:3], "angular_impulse": out[3:6], "energy": out[6], "total_strength": out[7:10]}

    # -- inspection -----------------------------------------------------------------
    def cells(self):
        C = ctypes.c_int64()
        _capi.check(self.lib.uvlm_num_cells(self._h, ctypes.byref(C)))
        ints = np.zeros((C.value, 5), dtype=np.int64)
        reals = np.zeros((C.value, 4))
        _capi.check(self.lib.uvlm_get_tree(self._h, C.value, _ip(ints), _dp(reals)))
        return ints, reals

    def lists(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _capi.check(self.lib.uvlm_get_lists(self._h, ctypes.byref(a), ctypes.byref(b), None, None))
        p2p = np.zeros((a.value, 2), dtype=np.int64)
        m2l = np.zeros((b.value, 2), dtype=np.int64)
        _capi.check(self.lib.uvlm_get_lists(self._h, None, None, _ip(p2p), _ip(m2l)))
        return p2p, m2l

    def expansions(self):
        C = ctypes.c_int64()
        _capi.check(self.lib.uvlm_num_cells(self._h, ctypes.byref(C)))
        M = np.zeros((C.value, self.nterm), dtype=np.complex128)
        L = np.zeros((C.value, self.nterm), dtype=np.complex128)
        _capi.check(self.lib.uvlm_get_expansions(self._h, _dp(M.view(np.float64)), _dp(L.view(np.float64))))
        return M, L

    def timings(self):
        t = _capi.Timing()
        _capi.check(self.lib.uvlm_timings(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_near_field(self, c):
        """Near-field option: cell pairs whose bodies can come closer than c * max sigma are near
        (P2P, regularised kernel); 0 restores the reference's lists.  Applies from the next build."""
        _capi.check(self.lib.uvlm_set_near_field(self._h, float(c)))

    def set_viscosity(self, scheme, nu):
        """"inviscid", "pse" or "csm" with viscosity nu; uvlm_advance applies it after each eval."""
        _capi.check(self.lib.uvlm_set_viscosity(self._h, VISCOUS[scheme], float(nu)))

    def viscosity(self, dt):
        """calculateViscosity on the rates of the last calculate()."""
        _capi.check(self.lib.uvlm_viscosity(self._h, float(dt)))

    def vorticity(self, probes):
        """Gaussian vorticity (negated, in velocity_old) of the particles at probe records
        (m, 37), near pairs only; the probes come back in tree order (in place)."""
        s = probes
        assert s.flags.c_contiguous and s.dtype == np.float64 and s.shape[1] == BODY_DOUBLES
        _capi.check(self.lib.uvlm_vorticity_aos(self._h, _dp(s), s.shape[0], BODY_BYTES))

    def rbf(self, points):
        """RBF re-initialisation onto point records (m, 37); returns (kept, iterations,
        residuals) with points[:kept] the kept points (tree order) carrying the fitted alpha."""
        s = points
        assert s.flags.c_contiguous and s.dtype == np.float64 and s.shape[1] == BODY_DOUBLES
        kept = ctypes.c_int64()
        it = ctypes.c_int()
        res = np.zeros(3)
        _capi.check(self.lib.uvlm_rbf_aos(self._h, _dp(s), s.shape[0], BODY_BYTES, ctypes.byref(kept),
                                          ctypes.byref(it), _dp(res)))
        return kept.value, it.value, res

    def autotune(self, ncrits=(64, 96, 128, 160, 192, 256), reps=2):
        """Time build + eval on the bodies already set for each candidate ncrit and M2L plan
        size; keep the fastest on the handle.  Returns (ncrit, ms).  Needs a build afterwards."""
        c = np.ascontiguousarray(ncrits, dtype=np.int64)
        best = ctypes.c_int64()
        ms = ctypes.c_double()
        _capi.check(self.lib.uvlm_autotune(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(c),
                                           int(reps), ctypes.byref(best), ctypes.byref(ms)))
        return int(best.value), float(ms.value)

    def sync(self):
        _capi.check(self.lib.uvlm_sync(self._h))
