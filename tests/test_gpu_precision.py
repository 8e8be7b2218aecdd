"""Precision study at 256^3: mixed FP32 stencil + FP64 reductions (the
solver's default) vs the all-FP64 GPU build of the same kernels vs the FP64
oracle (bit-identical to the reference built with FLOAT_SIZE=64).

Workload: 256^3, spacing (0.8, 0.8, 1.0), target exp(g) with g i.i.d.
N(0, 0.08^2) smoothed by a Gaussian of 6 voxels, clipped to [0.7, 1.4]; mask =
two balls of radii 40 and 60 centred at (x, y, z) = (96, 128, 128) and
(160, 128, 128); zero init; gradient; it_max 2000.

For each result the masked max and RMS of |J - T| (J of the returned field by
the FP64 oracle's stencils) are printed; the bar is the solver parity bar:
the same termination reason as the FP64 oracle and a masked max error within
1e-4 of its.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ERR_TOL = 1e-4


def _masked(oracle64, field, T, M, sp):
    J = oracle64.jacobian(field.astype(np.float64), sp)
    d = np.abs(J - T.astype(np.float64))[M]
    return float(d.max()), float(np.sqrt(np.mean(d * d)))


def test_precision_256(gpu_lib, oracle64):
    from disptools_amd.workloads import precision_256
    T, M, sp = precision_256()
    it_max = 2000
    rows = {}
    print("gpu fp32 stencil + fp64 sums", flush=True)
    f32, i32, _ = gpu_lib.solve_host("gradient", T, M, sp, it_max=it_max)
    old = os.environ.get("DISPTOOLS_STENCIL_PRECISION")
    os.environ["DISPTOOLS_STENCIL_PRECISION"] = "64"
    try:
        print("gpu all fp64", flush=True)
        f64, i64, _ = gpu_lib.solve_host("gradient", T, M, sp, it_max=it_max)
    finally:
        if old is None:
            del os.environ["DISPTOOLS_STENCIL_PRECISION"]
        else:
            os.environ["DISPTOOLS_STENCIL_PRECISION"] = old
    oracle64.set_threads(min(16, os.cpu_count() or 1))
    try:
        print("fp64 oracle (CPU)", flush=True)
        fo, io = oracle64.solve("gradient", T.astype(np.float64), M, sp, it_max=it_max)
    finally:
        oracle64.set_threads(1)
    rows["gpu_fp32_stencil_fp64_sums"] = (i32["reason"], i32["iterations"], *_masked(oracle64, f32, T, M, sp))
    rows["gpu_all_fp64"] = (i64["reason"], i64["iterations"], *_masked(oracle64, f64, T, M, sp))
    rows["fp64_oracle"] = (io["reason"], io["iterations"], *_masked(oracle64, fo, T, M, sp))
    du32 = float(np.abs(f32.astype(np.float64) - fo).max())
    du64 = float(np.abs(f64.astype(np.float64) - fo).max())
    print("\n%-28s %6s %6s %14s %14s" % ("solve", "reason", "iters", "max|J-T| mask", "rms|J-T| mask"))
    for k, (r, it, mx, rms) in rows.items():
        print("%-28s %6d %6d %14.9f %14.9f" % (k, r, it, mx, rms))
    print("max |u - u_oracle64|: fp32 stencil %.3g, all fp64 %.3g" % (du32, du64))
    ro = rows["fp64_oracle"]
    for k in ("gpu_fp32_stencil_fp64_sums", "gpu_all_fp64"):
        assert rows[k][0] == ro[0], (k, rows[k][0], ro[0])
        assert abs(rows[k][2] - ro[2]) <= MAX_ERR_TOL, (k, rows[k][2], ro[2])
