"""exageostatcpp_amd: ExaGeoStatCPP's maximum-likelihood path (Matern covariance
generation, tiled Cholesky with log-determinant and solve, kriging) written for
MI355X (gfx950) behind a C ABI (include/exageostat_amd.h).

Python here is a thin ctypes layer used by tests and benchmarks; the C++ API
mirror of ExaGeoStat<T> lives in exageostatcpp_amd/api/ and the kernels in
exageostatcpp_amd/csrc/.
"""
from .capi import (Context, EGSError, NotPositiveDefinite, Problem, generate_locations, lib, srand,  # noqa: F401
                   KERNEL_BIVARIATE, KERNEL_UNIVARIATE, PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED_BF16,
                   EXPORTS)
