"""Legacy-ASCII VTK POLYDATA writer for particle sets: the per-step file of the reference's
driver (src/main.cpp:127-164) with the same sections and the same number formatting (C++
ostream defaults, i.e. %g with 6 significant digits), so files from either side compare
byte for byte.  The C++ facade has the same function as exafmm::writeVTK."""
import numpy as np

from .fmm import BODY_FIELDS


def _g(v):
    return "%g" % v


def write_vtk(path, X, alpha, velocity, radius):
    """X, alpha, velocity: (n, 3); radius: (n,)."""
    X, alpha, velocity = (np.asarray(a, dtype=np.float64).reshape(-1, 3) for a in (X, alpha, velocity))
    radius = np.asarray(radius, dtype=np.float64).reshape(-1)
    n = X.shape[0]
    if not (alpha.shape[0] == velocity.shape[0] == radius.shape[0] == n):
        raise ValueError("X, alpha, velocity and radius must have the same length")
    out = ["# vtk DataFile Version 3.0", "vtk output", "ASCII", "DATASET POLYDATA", f"POINTS {n} double"]
    out += [" ".join(_g(v) for v in row) for row in X]
    out.append(f"VERTICES {n} {2 * n}")
    out += [f"1 {i}" for i in range(n)]
    out.append(f"POINT_DATA {n}")
    out.append("VECTORS alpha double")
    out += [" ".join(_g(v) for v in row) for row in alpha]
    out.append("VECTORS velocity double")
    out += [" ".join(_g(v) for v in row) for row in velocity]
    out += ["SCALARS radius double", "LOOKUP_TABLE default"]
    out += [_g(v) for v in radius]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_bodies_vtk(path, bodies):
    """Same file from exafmm::Body records, (n, 37) float64."""
    b = np.asarray(bodies)
    write_vtk(path, b[:, BODY_FIELDS["X"]], b[:, BODY_FIELDS["alpha"]], b[:, BODY_FIELDS["velocity"]],
              b[:, BODY_FIELDS["radius"]])
