"""uvlm_amd -- MI355X-native FMM vortex-particle solver (the hot path of kursatyurt/UVLM).

The compute path is libuvlm_fmm.so (HIP kernels for gfx950 behind the C ABI declared in
include/uvlm_fmm.h); this package is its Python host layer:

  FMMCalculator  -- mirror of Vortex::FMMCalculator (src/fmmCalculator.hpp:25-1006)
  workloads      -- seeded synthetic particle sets
"""
from .fmm import (BODY_DOUBLES, BODY_FIELDS, FMMCalculator, UVLMError, make_bodies)  # noqa: F401
from . import workloads  # noqa: F401
