"""HIP `matching` solver (src/VolumeMatching3D.c on the GPU) against the oracle.

Parity bar: the same number of accepted iterations, the same stop reason and
the same step lengths as the FP64 oracle; cost, max voxel error and field
within the drift that FP32 rounding grows into over the run (the oracle's
FP32 build drifts from the reference by the same amounts: up to a few percent
relative on the cost after 60 steps, under 1e-3 on the max voxel error, under
5e-4 on the field); the first log lines to 1e-4 relative.  The conversion of
the field to node positions and back rounds as the reference (FP32), so with
it_max = 0 the field matches the FP32 oracle bit for bit.
"""

import ctypes

import numpy as np
import pytest

from conftest import sphere

pytestmark = pytest.mark.gpu


def _ragged():
    from disptools_amd.workloads import volume_case
    return volume_case((13, 21, 34), 5)


def _config1():
    from disptools_amd.workloads import sphere_case
    return sphere_case(40)


CASES = {
    "sphere16": lambda: (*sphere(16, 5), (1.0, 1.0, 1.0), 60),
    "sphere24_aniso": lambda: (*sphere(24, 7, inside=0.9), (1.0, 1.0, 1.5), 60),
    "ragged_13x21x34": lambda: (*_ragged(), (1.0, 1.2, 0.9), 200),
    "config1_40": lambda: (*_config1(), 40),
}


@pytest.mark.parametrize("case", list(CASES))
def test_matching_parity(gpu_lib, oracle64, case):
    T, M, sp, it_max = CASES[case]()
    print("matching", case, T.shape, "it_max", it_max, flush=True)
    fo, io = oracle64.matching(T.astype(np.float64), sp, it_max=it_max)
    fg, ig, trace = gpu_lib.matching_host(T, sp, it_max=it_max, trace_cap=4096)
    print("iterations gpu %d oracle %d; reason %d %d; final max error gpu %.7g oracle %.7g; cost %.7g %.7g" %
          (ig["iterations"], io["iterations"], ig["reason"], io["reason"], ig["final_max_error"],
           io["final_max_error"], ig["final_cost"], io["final_cost"]))
    assert ig["iterations"] == io["iterations"]
    assert ig["reason"] == io["reason"]
    assert len(trace) == len(io["trace"])
    np.testing.assert_allclose(trace[:, 3], io["trace"][:, 3], rtol=1e-6)
    np.testing.assert_allclose(trace[:3, 1], io["trace"][:3, 1], rtol=1e-4)
    np.testing.assert_allclose(trace[:, 1], io["trace"][:, 1], rtol=5e-2)
    np.testing.assert_allclose(trace[:, 2], io["trace"][:, 2], rtol=0, atol=1e-3)
    assert abs(ig["final_max_error"] - io["final_max_error"]) <= 1e-3
    assert np.abs(fg - fo).max() < 5e-4
    assert ig["gradient_passes"] == io["gradient_evaluations"]
    assert ig["cost_passes"] == io["cost_evaluations"]


def test_matching_converging_case_tight(gpu_lib, oracle64):
    """A run that stops on the tolerance: final max voxel error within 1e-4."""
    T, M = _ragged()
    fo, io = oracle64.matching(T.astype(np.float64), (1.0, 1.0, 1.0), it_max=500)
    fg, ig, _ = gpu_lib.matching_host(T, (1.0, 1.0, 1.0), it_max=500)
    assert io["reason"] == ig["reason"] == 5
    assert abs(ig["final_max_error"] - io["final_max_error"]) <= 1e-4
    assert np.abs(fg - fo).max() < 1e-4


def test_matching_it_max_zero_roundtrip(gpu_lib, oracle32):
    T, M = sphere(12, 4)
    guess = np.random.default_rng(4).normal(0, 0.3, (3, 12, 12, 12)).astype(np.float32)
    sp = (1.0, 1.3, 0.7)
    fo, io = oracle32.matching(T, sp, field=guess, it_max=0)
    fg, ig, _ = gpu_lib.matching_host(T, sp, field=guess, it_max=0)
    assert ig["iterations"] == 0 and ig["reason"] == 6
    assert np.array_equal(fg, fo)


def test_matching_initial_guess(gpu_lib, oracle64):
    T, M = sphere(16, 5)
    guess = np.random.default_rng(9).normal(0, 0.02, (3, 16, 16, 16)).astype(np.float32)
    fo, io = oracle64.matching(T.astype(np.float64), field=guess.astype(np.float64), it_max=40)
    fg, ig, trace = gpu_lib.matching_host(T, field=guess, it_max=40, trace_cap=100)
    assert ig["iterations"] == io["iterations"] and ig["reason"] == io["reason"]
    np.testing.assert_allclose(trace[:, 3], io["trace"][:, 3], rtol=1e-6)
    assert np.abs(fg - fo).max() < 5e-4


def test_matching_all_ones_target(gpu_lib):
    ones = np.ones((8, 9, 10), np.float32)
    f, info, _ = gpu_lib.matching_host(ones, it_max=10)
    assert info["reason"] == 5 and info["iterations"] == 0 and info["final_max_error"] == 0
    assert np.abs(f).max() == 0


def test_matching_deterministic(gpu_lib):
    T, M = sphere(20, 6)
    a, ia, _ = gpu_lib.matching_host(T, it_max=30)
    b, ib, _ = gpu_lib.matching_host(T, it_max=30)
    assert np.array_equal(a, b) and ia["final_cost"] == ib["final_cost"]


def test_reference_entry_point_and_extension(gpu_lib):
    """volume_matching_3d (the reference's C entry point) and
    _disptools.displacement(..., "matching", gpu_id) run the same solve."""
    from disptools_amd import load_extension
    T, M = sphere(16, 5)
    sp = (1.0, 1.0, 1.5)
    fh, ih, _ = gpu_lib.matching_host(T, sp, it_max=25, epsilon=9.99e-4, delta=1e-3, zeta=10.0)
    f = np.zeros((3,) + T.shape, np.float32)
    fl = ctypes.c_float
    gpu_lib.lib.volume_matching_3d(T.shape[2], T.shape[1], T.shape[0], *map(fl, sp),
                                   T.ctypes.data_as(ctypes.c_void_p), M.ctypes.data_as(ctypes.c_void_p),
                                   fl(9.99e-4), fl(0.2), fl(0.1), fl(0.4), fl(1.2), fl(0.5), fl(0.1),
                                   fl(1e-3), fl(10.0), fl(1e-6), fl(1e-9), False, 25,
                                   f.ctypes.data_as(ctypes.c_void_p))
    assert not gpu_lib.lib.disptools_has_error()
    assert np.array_equal(f, fh)
    ext = load_extension()
    g = np.zeros((3,) + T.shape, np.float32)
    ext.displacement(sp, T, M, 9.99e-4, 0.2, 0.1, 0.4, 1.2, 0.5, 0.1, 1e-3, 10.0, 1e-6, 1e-9, False, 25, g,
                     "matching", 0)
    assert np.array_equal(g, fh)


def test_matching_device_buffers(gpu_lib):
    import torch
    T, M = sphere(16, 5)
    fh, ih, _ = gpu_lib.matching_host(T, it_max=20)
    dev = torch.device("cuda:0")
    dT = torch.from_numpy(T).to(dev)
    dF = torch.zeros((3,) + T.shape, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev)
    info, _ = gpu_lib.matching_device(T.shape[2], T.shape[1], T.shape[0], (1.0, 1.0, 1.0), dT.data_ptr(),
                                      dF.data_ptr(), stream=s.cuda_stream, it_max=20)
    assert info["iterations"] == ih["iterations"]
    assert np.array_equal(dF.cpu().numpy(), fh)
    assert info["kernel_seconds"] > 0


def test_python_displacement_matching(gpu_lib, oracle64):
    """displacement(..., algorithm='matching') pads, solves on the GPU and
    crops, like the other algorithms (reference displacements.py:443-483)."""
    from disptools_amd import displacements
    from disptools_amd.workloads import sphere_case
    T, M, sp = sphere_case(20, radius=5)
    out = displacements.displacement(T, mask=M, it_max=40, algorithm="matching", gpu_id=0)
    assert out.shape == (20, 20, 20, 3)
    Tp = np.pad(T, 1, mode="edge")
    fo, io = oracle64.matching(Tp.astype(np.float64), it_max=40)
    fo = np.moveaxis(fo, 0, -1)[1:-1, 1:-1, 1:-1]
    # FP32 drift over 40 steps (the field reaches ~1e-2 here)
    assert np.abs(out - fo).max() < 1e-3
    # two levels: the pyramid path runs matching on device buffers
    out2 = displacements.displacement(T, mask=M, it_max=[20, 20], levels=2, algorithm="matching", gpu_id=0)
    assert out2.shape == (20, 20, 20, 3) and np.isfinite(out2).all()
