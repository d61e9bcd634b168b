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
