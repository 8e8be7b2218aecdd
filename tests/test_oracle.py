"""Pins the CPU oracle (oracle/dt_oracle.c) before it is trusted as a checker.

* the reference's own known answers (test/c/test_jacobian_macros.cpp:70-148);
* closed forms of the Jacobian (u = 0, affine, quadratic, cubic fields);
* properties of the voxel error (error.h:87-172) that a wrong branch breaks;
* the gradient as the exact derivative of the weighted 2nd-order Jacobian;
* the reference's Jacobian test (test/test_displacements.py:43-55), with the
  ITK filter replaced by its closed form (2nd-order central differences).
"""

import numpy as np
import pytest

from oracle.oracle import ERROR_EXPLODED


# --------------------------------------------------------------------------
# Reference gtest known answers: forward one-sided derivatives on data[i] = i^2
# --------------------------------------------------------------------------

def _gtest_data(dtype):
    size = (5, 6, 7)  # x, y, z
    n = size[0] * size[1] * size[2] * 3
    data = (np.arange(n, dtype=np.float64) ** 2).astype(dtype)
    return data.reshape(3, size[2], size[1], size[0])


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gtest_forward_derivatives(oracle32, comp, axis):
    spacing = (2.0, 3.0, 4.0)
    f = _gtest_data(np.float32)
    p = (2, 1, 3)
    # the gtest passes 1/spacing of the component's axis (test_jacobian_macros.cpp:73-146)
    inv = np.float32(1.0) / np.float32(spacing[comp])
    got = oracle32.stencil(2, f, p, comp, axis, inv)

    def sample(dp):
        q = list(p)
        q[axis] += dp
        return float(f[comp, q[2], q[1], q[0]])

    expected = (1.0 / spacing[comp]) * (-1.5 * sample(0) + 2.0 * sample(1) - 0.5 * sample(2))
    assert got == pytest.approx(expected, abs=1e-6, rel=1e-6)


# --------------------------------------------------------------------------
# Closed forms
# --------------------------------------------------------------------------

def _grid(shape):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)


def test_zero_field_gives_unit_jacobian(oracle32):
    f = np.zeros((3, 7, 6, 5), np.float32)
    assert np.all(oracle32.jacobian(f, (2.0, 3.0, 4.0)) == 1.0)


@pytest.mark.parametrize("shape", [(3, 3, 3), (5, 6, 7), (9, 8, 12)])
def test_affine_field_exact(oracle64, shape):
    rng = np.random.default_rng(1)
    A = rng.normal(0, 0.2, size=(3, 3))
    spacing = (0.7, 1.3, 2.0)
    x, y, z = _grid(shape)
    coords = [x * spacing[0], y * spacing[1], z * spacing[2]]
    f = np.stack([sum(A[c, b] * coords[b] for b in range(3)) for c in range(3)])
    J = oracle64.jacobian(f, spacing)
    assert np.allclose(J, np.linalg.det(np.eye(3) + A), rtol=0, atol=1e-12)


def _poly_field(shape, spacing, degree, rng):
    x, y, z = _grid(shape)
    X = [x * spacing[0], y * spacing[1], z * spacing[2]]
    coef = rng.normal(0, 0.05, size=(3, 3))
    f = np.stack([sum(coef[c, b] * X[b] ** degree for b in range(3)) for c in range(3)])
    grad = np.zeros(shape + (3, 3))
    for c in range(3):
        for b in range(3):
            grad[..., c, b] = coef[c, b] * degree * X[b] ** (degree - 1)
    return f, grad, X, coef


def test_quadratic_field_exact_everywhere(oracle64):
    # every stencil of the region map (one-sided on the faces included) is exact
    # for quadratics, so J = det(I + grad u) at every voxel
    rng = np.random.default_rng(2)
    shape, spacing = (7, 8, 9), (1.0, 1.5, 0.5)
    f, grad, _, _ = _poly_field(shape, spacing, 2, rng)
    J = oracle64.jacobian(f, spacing)
    assert np.allclose(J, np.linalg.det(np.eye(3) + grad), rtol=0, atol=1e-11)


def test_cubic_field_region_map(oracle64):
    # 4th-order central is exact for cubics; 2nd-order central is not, with
    # error coef * h^2 along each axis: the penultimate layer must show it
    rng = np.random.default_rng(3)
    shape, spacing = (9, 10, 11), (1.0, 1.0, 1.0)
    f, grad, X, coef = _poly_field(shape, spacing, 3, rng)
    J = oracle64.jacobian(f, spacing)
    nz, ny, nx = shape
    exact = np.linalg.det(np.eye(3) + grad)
    inner = (slice(2, nz - 2), slice(2, ny - 2), slice(2, nx - 2))
    assert np.allclose(J[inner], exact[inner], rtol=0, atol=1e-11)
    c2 = grad.copy()
    for c in range(3):
        for b in range(3):
            c2[..., c, b] += coef[c, b] * spacing[b] ** 2
    approx2 = np.linalg.det(np.eye(3) + c2)
    pen = np.zeros(shape, bool)
    pen[1:-1, 1:-1, 1:-1] = True
    pen[2:-2, 2:-2, 2:-2] = False
    assert np.allclose(J[pen], approx2[pen], rtol=0, atol=1e-10)
    assert not np.allclose(J[pen], exact[pen], rtol=0, atol=1e-6)


def test_reference_jacobian_test_restated(oracle32):
    # test/test_displacements.py:43-55: sin(x^2+y^2+z^2) on linspace(0, 0.5, 10),
    # spacing 1, compared on [1:n-1] with rtol 1e-3 against ITK, whose filter
    # uses 2nd-order central differences: restated with that closed form.
    n = 10
    t = np.linspace(0, 0.5, n)
    xx, yy, zz = np.meshgrid(t, t, t)
    u = np.sin(xx ** 2 + yy ** 2 + zz ** 2).astype(np.float32)
    field_zyx3 = np.stack([u, u, u], axis=-1)
    f = np.ascontiguousarray(np.moveaxis(field_zyx3, -1, 0))
    J = oracle32.jacobian(f, (1.0, 1.0, 1.0)).astype(np.float64)
    g = np.zeros((n, n, n, 3, 3))
    fd = f.astype(np.float64)
    for c in range(3):
        g[1:-1, 1:-1, 1:-1, c, 0] = (fd[c, 1:-1, 1:-1, 2:] - fd[c, 1:-1, 1:-1, :-2]) / 2
        g[1:-1, 1:-1, 1:-1, c, 1] = (fd[c, 1:-1, 2:, 1:-1] - fd[c, 1:-1, :-2, 1:-1]) / 2
        g[1:-1, 1:-1, 1:-1, c, 2] = (fd[c, 2:, 1:-1, 1:-1] - fd[c, :-2, 1:-1, 1:-1]) / 2
    itk_like = np.linalg.det(np.eye(3) + g)[1:-1, 1:-1, 1:-1]
    assert np.allclose(J[1:-1, 1:-1, 1:-1], itk_like, 1e-3)


# --------------------------------------------------------------------------
# Voxel error
# --------------------------------------------------------------------------

def test_error_properties(oracle32):
    rng = np.random.default_rng(4)
    shape = (6, 7, 8)
    T = (1 + rng.normal(0, 0.3, shape)).astype(np.float32)
    Jf = (1 + rng.normal(0, 0.3, shape)).astype(np.float32)
    mask = rng.random(shape) < 0.4
    tol = 0.2
    E, mx, e = oracle32.compute_error(T, Jf, mask, tol)
    d = (Jf - T)
    # inside the mask the error is the plain difference
    assert np.array_equal(e[mask], d[mask])
    # outside: zero below the tolerance, shrunk towards zero by tol above it, same sign
    small = (~mask) & (np.abs(d) < tol)
    big = (~mask) & (np.abs(d) >= tol)
    assert np.all(e[small] == 0)
    assert np.all(np.sign(e[big]) == np.sign(d[big]))
    assert np.allclose(np.abs(e[big]), np.abs(d[big]) - tol, atol=1e-6)
    # energy over all voxels, max over the mask only
    assert E == pytest.approx(float(np.sum(e.astype(np.float64) ** 2)), rel=1e-5)
    assert mx == pytest.approx(float(np.max(np.abs(e[mask]))))
    outside_max = np.max(np.abs(e[~mask]))
    assert outside_max > mx or np.all(np.abs(e[~mask]) <= mx)


def test_error_tolerance_zero_is_difference(oracle32):
    rng = np.random.default_rng(5)
    shape = (4, 5, 6)
    T = rng.random(shape).astype(np.float32)
    Jf = rng.random(shape).astype(np.float32)
    mask = np.zeros(shape, bool)
    _, mx, e = oracle32.compute_error(T, Jf, mask, 0.0)
    assert np.array_equal(e, Jf - T)
    assert mx == 0.0


# --------------------------------------------------------------------------
# Gradient = 2 d/du sum_q w(q) J2(q) with frozen weights w = e + zeta (J - delta)[J < delta]
# --------------------------------------------------------------------------

def _j2_at(f, q, inv):
    z, y, x = q
    D = np.zeros((3, 3))
    for c in range(3):
        D[c, 0] = (f[c, z, y, x + 1] - f[c, z, y, x - 1]) * inv[0] * 0.5
        D[c, 1] = (f[c, z, y + 1, x] - f[c, z, y - 1, x]) * inv[1] * 0.5
        D[c, 2] = (f[c, z + 1, y, x] - f[c, z - 1, y, x]) * inv[2] * 0.5
    return np.linalg.det(np.eye(3) + D)


@pytest.mark.parametrize("delta,zeta", [(1e-3, 10.0), (1.2, 3.0)])
def test_gradient_is_adjoint_of_second_order_jacobian(oracle64, delta, zeta):
    rng = np.random.default_rng(6)
    shape = (8, 9, 10)
    spacing = (1.0, 1.5, 0.8)
    inv = [1.0 / s for s in spacing]
    f = rng.normal(0, 0.2, (3,) + shape)
    T = 1 + rng.normal(0, 0.1, shape)
    mask = rng.random(shape) < 0.5
    J = oracle64.jacobian(f, spacing)
    _, _, e = oracle64.compute_error(T, J, mask, 0.05)
    g, G = oracle64.gradient(f, J, e, delta, zeta, spacing)
    w = e + np.where(J < delta, zeta * (J - delta), 0.0)
    if delta > 1:
        assert np.any(J < delta)   # the penalty branch is exercised
    for p in [(3, 4, 4), (4, 4, 5), (5, 6, 7), (2, 2, 2), (4, 3, 6)]:
        z, y, x = p
        nb = [(z, y, x - 1), (z, y, x + 1), (z, y - 1, x), (z, y + 1, x), (z - 1, y, x), (z + 1, y, x)]
        for c in range(3):
            # F is linear in u_c(p): a central difference is exact
            h = 1e-3
            fp = f.copy(); fp[c, z, y, x] += h
            fm = f.copy(); fm[c, z, y, x] -= h
            Fp = sum(w[q] * _j2_at(fp, q, inv) for q in nb)
            Fm = sum(w[q] * _j2_at(fm, q, inv) for q in nb)
            fd = (Fp - Fm) / (2 * h)
            assert g[c, z, y, x] == pytest.approx(2 * fd, rel=1e-7, abs=1e-10)
    # boundary untouched, and G is the squared norm
    assert np.all(g[:, 0] == 0) and np.all(g[:, :, :, -1] == 0)
    assert G == pytest.approx(float(np.sum(g ** 2)), rel=1e-12)


# --------------------------------------------------------------------------
# Solver special cases fixed by the control flow
# --------------------------------------------------------------------------

def test_all_ones_target_explodes_at_first_iteration(oracle32):
    # E == 0 is not a normal number: "error exploded" at it = 1, initial guess returned
    T = np.ones((8, 8, 8), np.float32)
    M = np.ones((8, 8, 8), bool)
    for alg in ("gradient", "greedy"):
        f, info = oracle32.solve(alg, T, M, it_max=100)
        assert info["reason"] == ERROR_EXPLODED
        assert info["iterations"] == 1
        assert np.all(f == 0)


def test_it_max_zero_returns_initial_guess(oracle32):
    from conftest import sphere
    T, M = sphere(10, 3)
    guess = np.random.default_rng(0).normal(0, 0.01, (3, 10, 10, 10)).astype(np.float32)
    for alg in ("gradient", "greedy"):
        f, info = oracle32.solve(alg, T, M, field=guess, it_max=0)
        assert np.array_equal(f, guess)
