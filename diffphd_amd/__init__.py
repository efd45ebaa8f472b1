"""MI355X-native projective-dynamics time step (DiffPhD `pd_eigen_alg_phd`).

The compute path is libphd.so: HIP kernels for gfx950 behind the C ABI of
include/phd.h. `py_diff_pd_core` mirrors the reference's SWIG module names.
"""
from . import py_diff_pd_core  # noqa: F401
