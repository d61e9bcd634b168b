"""Vortex-lattice wing driving the particle wake, over the C ABI (uvlm_wing_*; src/main.cpp:11-295).

Vertices are [nchord+1][nspan+1][3] (row 0 the leading edge); nchord = 1 is the reference's
Wing of leading/trailing-edge vertex couples (main.cpp:27-74).
"""
import ctypes

import numpy as np

from . import _capi


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def reference_wing(npanels=500, span=10.0, chord=0.1, angle_deg=20.0, nchord=1):
    """The reference's wing (main.cpp:172-188): npanels + 1 vertex couples i * span/(npanels+1)
    along x, chord along y, rotated by -angle about x; nchord > 1 splits the chord evenly."""
    dx = span / (npanels + 1)
    a = np.deg2rad(-angle_deg)
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    x = np.arange(npanels + 1) * dx
    rows = [np.stack([x, np.full_like(x, chord * r / nchord), np.zeros_like(x)], 1) @ R.T for r in range(nchord + 1)]
    return np.stack(rows)


def station_twist(twist, nspan):
    """Twist at the nspan + 1 vertex stations from one value per spanwise strip: the mean of the
    two strips at an interior station, the strip's own value at a tip (a linear map T, returned
    with the values so gradients can go back through it)."""
    T = np.zeros((nspan + 1, nspan))
    for k in range(nspan + 1):
        if k == 0:
            T[k, 0] = 1.0
        elif k == nspan:
            T[k, nspan - 1] = 1.0
        else:
            T[k, k - 1] = T[k, k] = 0.5
    tw = np.zeros(nspan) if twist is None else np.asarray(twist, dtype=np.float64)
    return T @ tw, T


def design_wing(alpha, twist=None, span=6.0, chord=1.0, nspan=40, nchord=10):
    """Rectangular lattice of the benchmark wing (aspect ratio span / chord): stations
    x_k = k span / nspan, the chord along y split into nchord rows, every station rotated about
    the x axis through the leading edge by -(alpha + t_k) as the reference rotates its wing
    (main.cpp:172-184); t_k is the station twist of station_twist.  Angles in radians.
    Returns the vertices [nchord+1][nspan+1][3]."""
    t, _ = station_twist(twist, nspan)
    th = alpha + t
    x = np.arange(nspan + 1) * (span / nspan)
    V = np.zeros((nchord + 1, nspan + 1, 3))
    for r in range(nchord + 1):
        y = chord * r / nchord
        V[r, :, 0] = x
        V[r, :, 1] = y * np.cos(th)
        V[r, :, 2] = -y * np.sin(th)
    return V


def design_gradient(vertex_bar, alpha, twist=None, chord=1.0):
    """(dCL/dalpha, dCL/dtwist [nspan]) from the vertex cotangent of uvlm_wing_window_adjoint for
    the lattice of design_wing: dV[r][k]/dtheta_k = (0, -y_r sin theta_k, -y_r cos theta_k) with
    y_r the chord position of row r, theta_k = alpha + t_k."""
    vb = np.asarray(vertex_bar, dtype=np.float64)
    assert vb.ndim == 3 and vb.shape[2] == 3, "vertex_bar is [nchord+1][nspan+1][3]"
    nchord, nspan = vb.shape[0] - 1, vb.shape[1] - 1
    t, T = station_twist(twist, nspan)
    th = alpha + t
    dth = np.zeros(nspan + 1)
    for r in range(nchord + 1):
        y = chord * r / nchord
        dth += vb[r, :, 1] * (-y * np.sin(th)) + vb[r, :, 2] * (-y * np.cos(th))
    return float(dth.sum()), T.T @ dth


class Wing:
    def __init__(self, vertices, device=0):
        V = np.ascontiguousarray(vertices, dtype=np.float64)
        assert V.ndim == 3 and V.shape[2] == 3
        self.nchord, self.nspan = V.shape[0] - 1, V.shape[1] - 1
        self.n = self.nchord * self.nspan
        self.lib = _capi.lib()
        self._w = ctypes.c_void_p()
        _capi.check(self.lib.uvlm_wing_create(int(device), self.nspan, self.nchord, _dp(V), ctypes.byref(self._w)))

    def close(self):
        if getattr(self, "_w", None) and self._w.value:
            self.lib.uvlm_wing_destroy(self._w)
            self._w = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        """Control points, unit normals, areas and the LU factors of the system matrix."""
        n = self.n
        cp, nrm, area, lu = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros((n, n))
        _capi.check(self.lib.uvlm_wing_get(self._w, _dp(cp), _dp(nrm), _dp(area), _dp(lu)))
        return cp, nrm, area, lu

    def translate(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        _capi.check(self.lib.uvlm_wing_translate(self._w, _dp(d)))

    def solve(self, particles, U):
        """(gamma, force) at the wing's position; particles: an FMMCalculator or None."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        g, F = np.zeros(self.n), np.zeros(3)
        h = particles._h if particles is not None else None
        _capi.check(self.lib.uvlm_wing_solve(self._w, h, _dp(U), _dp(g), _dp(F)))
        return g, F

    def shed(self, particles, U, dt, reference_offset=True):
        U = np.ascontiguousarray(U, dtype=np.float64)
        _capi.check(self.lib.uvlm_wing_shed(self._w, particles._h, _dp(U), float(dt), int(bool(reference_offset))))
        particles.n += self.nspan

    def step(self, particles, U, dt, reference_offset=True):
        """One pass of the reference's loop: solve, shed, advance the particles, move the wing
        by -U dt.  Returns (gamma, force) of the solve."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        g, F = np.zeros(self.n), np.zeros(3)
        _capi.check(self.lib.uvlm_wing_step(self._w, particles._h, _dp(U), float(dt), int(bool(reference_offset)),
                                            _dp(g), _dp(F)))
        particles.n += self.nspan
        return g, F

    def window_adjoint(self, particles, U, dt, steps, every=10, e=(0.0, 0.0, 1.0), reference_offset=True):
        """CL after `steps` loop steps and its gradient with respect to the particles now
        (original order) and to the wing's vertices, checkpoints every `every` steps:
        (CL, xbar, alphabar, vertex_bar [nchord+1][nspan+1][3])."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        n = particles.n
        xb, ab = np.zeros((max(n, 1), 3)), np.zeros((max(n, 1), 3))
        vb = np.zeros((self.nchord + 1, self.nspan + 1, 3))
        CL = ctypes.c_double()
        _capi.check(self.lib.uvlm_wing_window_adjoint(self._w, particles._h, _dp(U), float(dt), int(steps), int(every),
                                                      int(bool(reference_offset)), _dp(e),
                                                      ctypes.cast(ctypes.byref(CL), ctypes.POINTER(ctypes.c_double)),
                                                      n, _dp(xb), _dp(ab), _dp(vb)))
        particles.n += steps * self.nspan
        return CL.value, xb[:n], ab[:n], vb

    def lift_adjoint(self, particles, U, dt, steps, e=(0.0, 0.0, 1.0), reference_offset=True):
        """CL at the position after `steps` loop steps and its gradient with respect to the
        particles' positions and strengths now (original order): (CL, xbar, alphabar)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        n = particles.n
        xb, ab = np.zeros((max(n, 1), 3)), np.zeros((max(n, 1), 3))
        CL = ctypes.c_double()
        _capi.check(self.lib.uvlm_wing_lift_adjoint(self._w, particles._h, _dp(U), float(dt), int(steps),
                                                    int(bool(reference_offset)), _dp(e),
                                                    ctypes.cast(ctypes.byref(CL), ctypes.POINTER(ctypes.c_double)),
                                                    n, _dp(xb), _dp(ab)))
        particles.n += steps * self.nspan
        return CL.value, xb[:n], ab[:n]
