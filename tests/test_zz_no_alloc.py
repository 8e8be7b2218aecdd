"""dt_solve_device_noalloc (the entry point that allocates nothing) against
dt_solve_device: the same fields and records, bit for bit, on two calls in a
row (the second with a smaller volume, which finds the first call's contents
in the shared scratch array), for the gradient and the greedy solver; a
volume whose state does not fit the scratch array is refused before any
device work."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("iterations", "evaluations", "rejections", "reason", "final_error", "final_max_error", "field_io")


def _dev(T, M):
    import torch
    d = torch.device("cuda:0")
    return (torch.from_numpy(T).to(d), torch.from_numpy(M.astype(np.uint8)).to(d),
            torch.zeros((3,) + T.shape, dtype=torch.float32, device=d))


@pytest.mark.parametrize("alg", ["gradient", "greedy"])
def test_noalloc_equals_solve_device_on_two_calls(gpu_lib, alg):
    import torch
    from disptools_amd.workloads import volume_case
    cases = [volume_case((40, 37, 33), 21), volume_case((19, 23, 17), 22)]   # second call smaller
    for k, (T, M) in enumerate(cases):
        nz, ny, nx = T.shape
        sp = (1.0, 1.1, 0.9)
        t, m, f_ref = _dev(T, M)
        _, _, f_na = _dev(T, M)
        torch.cuda.synchronize()
        print("call", k, alg, T.shape, flush=True)
        ref, _ = gpu_lib.solve_device(alg, nx, ny, nz, sp, t.data_ptr(), m.data_ptr(), f_ref.data_ptr(), it_max=300)
        torch.cuda.synchronize()
        got = gpu_lib.solve_device_noalloc(alg, nx, ny, nz, sp, t.data_ptr(), m.data_ptr(), f_na.data_ptr(),
                                           it_max=300)
        torch.cuda.synchronize()
        for key in KEYS:
            assert got[key] == ref[key], (k, key, got[key], ref[key])
        np.testing.assert_array_equal(f_na.cpu().numpy(), f_ref.cpu().numpy())


def test_noalloc_refuses_a_volume_beyond_its_scratch(gpu_lib):
    import ctypes
    info = gpu_lib.SolveInfo()
    p = gpu_lib.make_params(it_max=10)
    # 200^3 gradient: 96 MB of solver state > 80 MiB; refused before any device access
    rc = gpu_lib.lib.dt_solve_device_noalloc(0, 200, 200, 200, 1.0, 1.0, 1.0, 1, 1, ctypes.byref(p), 1,
                                             ctypes.byref(info), None)
    assert rc == -1 and "too large" in gpu_lib.last_error()
