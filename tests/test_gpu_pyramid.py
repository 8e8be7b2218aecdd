"""Image pyramid on the device (dt_scale_volume_device, dt_displacement_levels)
against the pyramid oracle (oracle/pyramid.py: Deriche's recursive Gaussian,
which SimpleITK's SmoothingRecursiveGaussian documents, and ITK's resampling
rules, pinned by mathematics in test_pyramid_oracle.py).

The HIP filter runs the same FP64 recursions and stores float: agreement with
the FP64 oracle within 2e-6 of the values' scale.  The multi-level driver is
checked level by level against the same composition of oracle scaling and
lone device solves, and for the solver's own properties.
"""

import numpy as np
import pytest

from oracle import pyramid as P

pytestmark = pytest.mark.gpu


def _scale_device(gpu_lib, v, spacing, size, nearest, sigma=2.0):
    import torch
    d = torch.from_numpy(np.ascontiguousarray(v, np.float32)).to("cuda:0")
    nc, nz, ny, nx = v.shape
    mx, my, mz = size
    out = torch.empty((nc, mz, my, mx), dtype=torch.float32, device="cuda:0")
    gpu_lib.scale_volume_device(nx, ny, nz, nc, spacing, d.data_ptr(), size, sigma, nearest, out.data_ptr(),
                                torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("nearest", [False, True])
@pytest.mark.parametrize("shape,size,spacing", [((2, 17, 13, 11), (5, 6, 8), (1.0, 1.5, 0.8)),
                                                ((3, 20, 20, 20), (10, 10, 10), (1.0, 1.0, 1.0)),
                                                ((1, 9, 10, 12), (24, 20, 18), (2.0, 1.0, 1.0)),
                                                ((1, 40, 33, 70), (35, 16, 20), (1.0, 1.0, 1.5))])
def test_scale_volume_matches_oracle(gpu_lib, shape, size, spacing, nearest):
    rng = np.random.default_rng(1)
    v = rng.normal(size=shape).astype(np.float32)
    out = _scale_device(gpu_lib, v, spacing, size, nearest)
    ref, _ = P.scale_image(v, spacing, size, nearest)
    if nearest:
        # the same cell is picked; the smoothed values agree to float rounding
        np.testing.assert_allclose(out, ref, rtol=0, atol=2e-6)
    else:
        np.testing.assert_allclose(out, ref, rtol=0, atol=2e-6)


def test_scale_volume_impulse_is_gaussian(gpu_lib):
    """An impulse smoothed on the device (no resampling: same grid) is the
    separable recursive Gaussian: its profile along each axis matches the
    oracle's 1-D impulse response, which approximates the sampled Gaussian
    to 5e-4 of its peak."""
    n = 41
    v = np.zeros((1, n, n, n), np.float32)
    v[0, n // 2, n // 2, n // 2] = 1.0
    out = _scale_device(gpu_lib, v, (1.0, 1.0, 1.0), (n, n, n), False)
    line = np.zeros(n)
    line[n // 2] = 1.0
    r = P.recursive_gaussian_1d(line, 2.0)
    ref = r[:, None, None] * r[None, :, None] * r[None, None, :]
    np.testing.assert_allclose(out[0], ref, rtol=0, atol=1e-8)
    assert abs(float(out.sum()) - 1.0) < 1e-5


def test_levels_match_oracle_composition(gpu_lib, oracle32):
    """dt_displacement_levels with levels = 2 equals, bit for bit, the same
    steps made one by one: the coarse target and mask from the device
    scaling (which matches the oracle above), the coarse solve, its field
    scaled onto the fine grid, the fine solve from it."""
    from disptools_amd import native
    from disptools_amd.workloads import sphere_case
    T, M, sp = sphere_case(30, radius=8.0)
    Mf = M.astype(np.float32)
    field, info = native.displacement_levels("gradient", T, Mf, None, sp, 2, [200, 300])
    nz, ny, nx = T.shape
    coarse = (nx // 2, ny // 2, nz // 2)
    Tc = _scale_device(gpu_lib, T[None], sp, coarse, False)[0]
    Mc = _scale_device(gpu_lib, Mf[None], sp, coarse, True)[0]
    # the coarse level's spacing keeps the physical extent
    spc = tuple(s * n / m for s, n, m in zip(sp, (nx, ny, nz), coarse))
    fc, _, _ = gpu_lib.solve_host("gradient", Tc, Mc != 0, spc, it_max=200)
    f0 = _scale_device(gpu_lib, fc, spc, (nx, ny, nz), False)
    ff, info1, _ = gpu_lib.solve_host("gradient", T, M, sp, field=f0, it_max=300)
    np.testing.assert_array_equal(field, np.moveaxis(ff, 0, -1))
    assert info["iterations"] == info1["iterations"]


def test_multilevel_displacement(gpu_lib, oracle32):
    """levels=2 solves the coarse level, then the full level from the
    upsampled field; the result meets the tolerance where the single-level
    solve does and keeps the input's shape."""
    from disptools_amd import displacements
    from disptools_amd.workloads import sphere_case
    T, M, sp = sphere_case(40)
    one = displacements.displacement(T, mask=M, it_max=500, gpu_id=0)
    two = displacements.displacement(T, mask=M, levels=2, it_max=[300, 500], gpu_id=0)
    assert two.shape == one.shape == (40, 40, 40, 3)
    assert np.isfinite(two).all()
    for f in (one, two):
        J = oracle32.jacobian(np.ascontiguousarray(np.moveaxis(f, -1, 0)), sp)
        _, mx, _ = oracle32.compute_error(T, J, M, 0.2)
        print("masked max error", mx)
        assert mx < 2e-3


def test_multilevel_product_path_runs_without_torch():
    """displacement(levels=2) with numpy in and out, in a process that never
    imports torch: the library allocates the pyramid and swaps the layouts
    itself."""
    import subprocess
    import sys
    code = (
        "import sys, numpy as np\n"
        "from disptools_amd import displacements\n"
        "from disptools_amd.workloads import sphere_case\n"
        "T, M, sp = sphere_case(24, radius=6.0)\n"
        "f = displacements.displacement(T, mask=M, levels=2, it_max=[100, 100], gpu_id=0)\n"
        "assert f.shape == (24, 24, 24, 3) and np.isfinite(f).all()\n"
        "assert 'torch' not in sys.modules, 'torch was imported'\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().endswith("ok")
