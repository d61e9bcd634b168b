"""VTK output (the reference driver's per-step file, src/main.cpp:127-164): the C++ facade's
exafmm::writeVTK and the Python writer produce the same bytes, and a two-body file matches
the layout written out by hand."""
import os
import subprocess

import numpy as np
import pytest

from uvlm_amd.fmm import make_bodies
from uvlm_amd.vtk import write_bodies_vtk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CPP = r'''
#include <cstdio>
#include "uvlm/fmm_calculator.hpp"
int main(int argc, char **argv)
{
  FILE *f = std::fopen(argv[1], "rb");
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  exafmm::Bodies b(bytes / sizeof(exafmm::Body));
  if (std::fread(b.data(), sizeof(exafmm::Body), b.size(), f) != b.size()) return 3;
  std::fclose(f);
  return exafmm::writeVTK(b, argv[2]) ? 0 : 4;
}
'''


def test_two_body_layout(tmp_path):
    b = make_bodies(np.array([[0.5, -1.0, 2.0], [1e-5, 3.25, 0.0]]), np.array([[1.0, 0.0, -2.5], [0.125, 1e6, 7.0]]),
                    np.array([0.1, 0.2]))
    b[:, 15:18] = [[0.0, 1.5, -3.0], [2.0, 0.0, 1.0]]
    p = tmp_path / "two.vtk"
    write_bodies_vtk(p, b)
    want = ("# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET POLYDATA\nPOINTS 2 double\n"
            "0.5 -1 2\n1e-05 3.25 0\nVERTICES 2 4\n1 0\n1 1\nPOINT_DATA 2\n"
            "VECTORS alpha double\n1 0 -2.5\n0.125 1e+06 7\n"
            "VECTORS velocity double\n0 1.5 -3\n2 0 1\n"
            "SCALARS radius double\nLOOKUP_TABLE default\n0.1\n0.2\n")
    assert p.read_text() == want


def test_cpp_and_python_writers_agree(tmp_path):
    exe = tmp_path / "vtk_writer"
    src = tmp_path / "vtk_writer.cpp"
    src.write_text(CPP)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    rng = np.random.default_rng(5)
    n = 500
    b = make_bodies(rng.normal(size=(n, 3)) * 10 ** rng.uniform(-6, 6, size=(n, 1)), rng.normal(size=(n, 3)),
                    rng.uniform(0.01, 1, size=n))
    b[:, 15:18] = rng.normal(size=(n, 3))
    fin = tmp_path / "in.bin"
    b.tofile(fin)
    r = subprocess.run([str(exe), str(fin), str(tmp_path / "cpp.vtk")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    write_bodies_vtk(tmp_path / "py.vtk", b)
    assert (tmp_path / "cpp.vtk").read_bytes() == (tmp_path / "py.vtk").read_bytes()
