"""HIP jacobian against the oracle (bit-exact: the kernel evaluates the
reference's stencils and determinant in FP64 with the same operation order)."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.npz")

SHAPES = [
    ((3, 3, 3), (1.0, 1.0, 1.0)),
    ((4, 5, 3), (1.0, 1.0, 1.5)),
    ((7, 6, 5), (2.0, 3.0, 4.0)),
    ((11, 10, 9), (1.0, 1.0, 1.5)),
    ((16, 16, 16), (0.8, 0.8, 1.0)),
    ((9, 17, 33), (1.0, 2.0, 0.5)),
    ((40, 40, 40), (1.0, 1.0, 1.0)),
    ((70, 66, 130), (1.0, 1.0, 1.5)),
    ((37, 70, 131), (1.0, 0.7, 1.3)),
    ((128, 128, 128), (1.0, 1.0, 1.5)),
]


@pytest.mark.parametrize("shape,spacing", SHAPES)
def test_jacobian_bit_exact_vs_oracle(gpu_lib, oracle32, shape, spacing):
    print("jacobian", shape, spacing, flush=True)
    rng = np.random.default_rng(sum(shape))
    f = rng.normal(0, 0.3, (3,) + shape).astype(np.float32)
    J = gpu_lib.jacobian_host(f, spacing)
    ref = oracle32.jacobian(f, spacing)
    assert np.array_equal(J.view(np.uint32), ref.view(np.uint32))


def test_jacobian_golden(gpu_lib):
    g = np.load(GOLDEN)
    for name in ("jac_567", "jac_91011", "jac_16"):
        J = gpu_lib.jacobian_host(g[name + "/field"], tuple(g[name + "/spacing"]))
        assert np.array_equal(J, g[name + "/J"]), name


def test_jacobian_quadratic_closed_form(gpu_lib):
    shape, spacing = (9, 10, 11), (1.0, 1.5, 0.5)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = [x * spacing[0], y * spacing[1], z * spacing[2]]
    coef = np.random.default_rng(2).normal(0, 0.05, (3, 3))
    f = np.stack([sum(coef[c, b] * X[b] ** 2 for b in range(3)) for c in range(3)]).astype(np.float32)
    grad = np.zeros(shape + (3, 3))
    for c in range(3):
        for b in range(3):
            grad[..., c, b] = coef[c, b] * 2 * X[b]
    J = gpu_lib.jacobian_host(f, spacing)
    # f is rounded to float32 before the kernel sees it, so the quadratic is exact
    # only to ~1e-7 relative in the samples; a wrong stencil is off by >1e-2
    assert np.allclose(J, np.linalg.det(np.eye(3) + grad), rtol=0, atol=2e-5)


def test_extension_jacobian_matches(gpu_lib, oracle32):
    from disptools_amd import load_extension
    ext = load_extension()
    f = np.random.default_rng(7).normal(0, 0.2, (3, 12, 13, 14)).astype(np.float32)
    out = np.zeros((12, 13, 14), np.float32)
    ext.jacobian((1.0, 1.0, 1.5), f, out)
    assert np.array_equal(out, oracle32.jacobian(f, (1.0, 1.0, 1.5)))


def test_python_jacobian_itk_layout(gpu_lib, oracle32):
    from disptools_amd import displacements
    f = np.random.default_rng(8).normal(0, 0.2, (3, 10, 11, 12)).astype(np.float32)
    J = displacements.jacobian(np.moveaxis(f, 0, -1), spacing=(2.0, 1.0, 1.0))
    assert np.array_equal(J, oracle32.jacobian(f, (2.0, 1.0, 1.0)))


def test_jacobian_too_small_is_an_error(gpu_lib):
    with pytest.raises(gpu_lib.HipError):
        gpu_lib.jacobian_host(np.zeros((3, 2, 5, 5), np.float32))


def test_regularise(gpu_lib):
    from disptools_amd import displacements
    rng = np.random.default_rng(1)
    J = rng.random((20, 21, 22)).astype(np.float32)
    out = displacements.regularise(J, 0.3)
    assert np.array_equal(out, np.maximum(J, np.float32(0.3)))


def test_jacobian_too_large_is_an_error(gpu_lib):
    """More than 357913941 voxels (12 bytes per voxel beyond 32-bit byte
    offsets): refused before any buffer is touched."""
    with pytest.raises(gpu_lib.HipError, match="too large"):
        gpu_lib.jacobian_device(1000, 1000, 1000, (1.0, 1.0, 1.0), 0, 0)
