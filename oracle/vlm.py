"""CPU restatement of the reference's vortex-lattice wing loop (src/main.cpp) -- TEST
INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's cpu_baseline leg may import it).

The reference's wing (main.cpp:27-74) is one chordwise row of panels between leading-edge and
trailing-edge vertex couples.  This restatement takes a lattice of vertices
V[r][c] (r = 0 .. nchord, row 0 the leading edge; c = 0 .. nspan); nchord = 1 is exactly the
reference's wing, with LE_Vertices[c] = V[0][c] and TE_Vertices[c] = V[1][c].  Panel
p = r * nspan + c has the corners LE_c = V[r][c], LE_c1 = V[r][c+1], TE_c = V[r+1][c],
TE_c1 = V[r+1][c+1].

Dense linear algebra (Eigen's fullPivLu in main.cpp:214) is restated with numpy's LU solve; the
Eigen submodule is empty in the reference and its version is not pinned.
"""
import numpy as np

INV_4PI = 0.25 / np.pi


def segment_velocity(a, b, p):
    """Unit-strength straight vortex segment a -> b at points p (main.cpp:11-25):
    u = (r1 x r2) / |r1 x r2|^2 * r0 . (r1/|r1| - r2/|r2|) / (4 pi), r1 = p - a, r2 = p - b,
    r0 = b - a.  a, b: (..., 3) broadcastable against p: (..., 3).  Raises where the reference
    calls std::exit(1) (|r1|, |r2| or |r1 x r2| below 1e-6)."""
    a, b, p = np.broadcast_arrays(np.asarray(a, float), np.asarray(b, float), np.asarray(p, float))
    r1, r2, r0 = p - a, p - b, b - a
    c = np.cross(r1, r2)
    n1, n2, nc = np.linalg.norm(r1, axis=-1), np.linalg.norm(r2, axis=-1), np.linalg.norm(c, axis=-1)
    if np.any(n1 < 1e-6) or np.any(n2 < 1e-6) or np.any(nc < 1e-6):
        raise ValueError("singular vortex segment (main.cpp:19-21)")
    k = np.sum(r0 * (r1 / n1[..., None] - r2 / n2[..., None]), axis=-1) / (nc * nc)
    return INV_4PI * c * k[..., None]


class Wing:
    def __init__(self, V):
        V = np.asarray(V, dtype=float)
        assert V.ndim == 3 and V.shape[2] == 3 and V.shape[0] >= 2 and V.shape[1] >= 2
        self.V = V
        self.nchord, self.nspan = V.shape[0] - 1, V.shape[1] - 1
        self.n = self.nchord * self.nspan
        LE, LE1, TE, TE1 = V[:-1, :-1], V[:-1, 1:], V[1:, :-1], V[1:, 1:]
        self.LE, self.LE1, self.TE, self.TE1 = (x.reshape(-1, 3) for x in (LE, LE1, TE, TE1))
        # calculateTopology (main.cpp:45-58)
        self.cp = (self.LE + self.LE1 + self.TE + self.TE1) / 4.0
        nrm = np.cross(self.LE1 - self.TE, self.LE - self.TE1)
        self.area = 0.5 * np.linalg.norm(nrm, axis=1)
        self.normal = nrm / np.linalg.norm(nrm, axis=1)[:, None]

    def lines(self):
        """getPanelVortexLine (main.cpp:60-73): (start, end) of lines 0..3 of every panel,
        arrays (4, n, 3): TE_c1->TE_c, TE_c->LE_c, LE_c->LE_c1, LE_c1->TE_c1."""
        s = np.stack([self.TE1, self.TE, self.LE, self.LE1])
        e = np.stack([self.TE, self.LE, self.LE1, self.TE1])
        return s, e

    def influence(self, pts):
        """B[t, j] = velocity at pts[t] of panel j's ring with unit circulation, (m, n, 3)."""
        s, e = self.lines()
        pts = np.asarray(pts, float)
        B = np.zeros((pts.shape[0], self.n, 3))
        for k in range(4):
            B += segment_velocity(s[k][None, :, :], e[k][None, :, :], pts[:, None, :])
        return B

    def aic(self):
        """A[i, j] = (velocity of ring j at control point i) . n_i.  main.cpp:198-205 fills
        AIC(j, i) with this value and solves AIC gamma = rhs; on a lattice of congruent
        parallelograms (the reference's wing) a half turn about the normal through the midpoint
        of two control points swaps ring i at point j with ring j at point i, so A is symmetric
        and the two systems coincide."""
        return np.einsum("ijd,id->ij", self.influence(self.cp), self.normal)

    def net_circulation(self, gamma):
        """Circulation of each panel's leading-edge segment: gamma_p minus that of the panel
        upstream (its trailing-edge segment is the same line reversed); gamma_p on row 0."""
        g = np.asarray(gamma, float).reshape(self.nchord, self.nspan)
        net = g.copy()
        net[1:] -= g[:-1]
        return net.reshape(-1)


def solve(wing, U, v_cp):
    """Panel circulations (main.cpp:207-214): A gamma = -U.n - v.n, v the particle-induced
    velocity at the control points (getSensorData, 196-197)."""
    rhs = -(wing.normal @ np.asarray(U, float)) - np.sum(v_cp * wing.normal, axis=1)
    return np.linalg.solve(wing.aic(), rhs)


def force(wing, gamma, U, v_cp):
    """Kutta-Joukowski force (main.cpp:223-244): sum_i Gamma_i (u_i x (LE_c1 - LE_c)), u_i the
    velocity at control point i of every ring (gamma_j) + U + v_i; Gamma_i is the net
    circulation of the panel's leading-edge segment (gamma_i on the reference's one row)."""
    u = np.einsum("ijd,j->id", wing.influence(wing.cp), gamma) + np.asarray(U, float)[None, :] + v_cp
    dxx = wing.LE1 - wing.LE
    return np.sum(wing.net_circulation(gamma)[:, None] * np.cross(u, dxx), axis=0)


def shed(wing, gamma, U, dt, reference_offset=True):
    """Particles shed from the trailing-edge row (main.cpp:246-289): X at the midpoint of the
    trailing-edge segment plus U dt (the reference offsets z by U[1] dt, line 254, kept when
    reference_offset), alpha = (TE_c - TE_c1) gamma_p plus half the circulation jump across
    each side segment (the full gamma at the tips), sigma = 2.5 |TE_c - TE_c1|.
    Returns X, alpha, sigma of the nspan particles."""
    ns = wing.nspan
    g = np.asarray(gamma, float).reshape(wing.nchord, ns)[-1]
    TE, TE1 = wing.TE[-ns:], wing.TE1[-ns:]
    LE, LE1 = wing.LE[-ns:], wing.LE1[-ns:]
    U = np.asarray(U, float)
    off = np.array([U[0], U[1], U[1] if reference_offset else U[2]]) * dt
    X = (TE1 + TE) / 2.0 + off[None, :]
    dxx = TE - TE1
    A = dxx * g[:, None]
    left, right = LE - TE, TE1 - LE1
    for p in range(ns):
        A[p] += left[p] * ((g[p] - g[p - 1]) * 0.5 if p > 0 else g[p])
        A[p] += right[p] * ((g[p] - g[p + 1]) * 0.5 if p < ns - 1 else g[p])
    return X, A, 2.5 * np.linalg.norm(dxx, axis=1)


def reference_wing(npanels=500, span=10.0, chord=0.1, angle_deg=20.0):
    """The reference's wing (main.cpp:172-188): npanels + 1 vertex couples i dx along x, chord
    along y, rotated by -angle about x."""
    dx = span / (npanels + 1)
    a = np.deg2rad(-angle_deg)
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    le = np.stack([np.arange(npanels + 1) * dx, np.zeros(npanels + 1), np.zeros(npanels + 1)], 1)
    te = le + np.array([0.0, chord, 0.0])
    return np.stack([le @ R.T, te @ R.T])


def run(V0, U, dt, steps, P=6, theta=0.5, ncrit=64, scheme="rk2", reference_offset=True):
    """The reference's time loop (main.cpp:168-295) for `steps` steps from no particles: at step t
    the wing sits at V0 - dt t U; the particle velocity at its control points comes from the
    oracle FMM at the particles' current state (init + calculate + getSensorData); circulations,
    force, shedding; then advance(particles, dt).  Returns the per-step (gamma, force) and the
    final particle records (n, 37), in the oracle's body order."""
    from . import oracle as O

    U = np.asarray(U, float)
    calc = O.Calculator(P=P, theta=theta, ncrit=ncrit)
    bodies = np.zeros((0, O.BODY_DOUBLES))
    hist = []
    for t in range(steps):
        wing = Wing(np.asarray(V0, float) - dt * t * U)
        v = np.zeros((wing.n, 3))
        if bodies.shape[0]:
            calc.init(bodies)
            calc.calculate()
            s = O.make_bodies(wing.cp, np.zeros((wing.n, 3)), np.ones(wing.n))
            calc.sensor_data(s)
            where = {tuple(x): i for i, x in enumerate(wing.cp)}
            for row in s:
                v[where[tuple(row[O.BODY_FIELDS["X"]])]] = row[O.BODY_FIELDS["velocity"]]
        gamma = solve(wing, U, v)
        F = force(wing, gamma, U, v)
        X, A, S = shed(wing, gamma, U, dt, reference_offset)
        bodies = np.ascontiguousarray(np.concatenate([bodies, O.make_bodies(X, A, S)]))
        calc.advance(bodies, dt, scheme)
        hist.append((gamma, F))
    return hist, bodies
