"""CPU restatement of the image pyramid step, drawing.scale_image.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``; the product package never
imports it.

The reference builds every pyramid level with ``drawing.scale_image``
(disptools/drawing.py:23-69): ``sitk.SmoothingRecursiveGaussian(image, 2.0)``
followed by ``ResampleImageFilter`` onto a grid of the new size with the same
origin and physical extent (spacing ``s * n / n'``), linear interpolation for
images and fields, nearest neighbour for masks (displacements.py:458-461,
476).  SimpleITK (``SimpleITK>=1.1.0`` in the reference's setup.py:107; no
version pinned) is not installed here, so this module restates the published
algorithms it documents:

* ``SmoothingRecursiveGaussian``: along each axis the fourth-order recursive
  approximation of the Gaussian of R. Deriche, "Recursively implementing the
  Gaussian and its derivatives", INRIA research report 1893 (1993), with the
  constants of that report (a0 = 1.680, a1 = 3.735, b0 = 1.783, b1 = 1.723,
  w0 = 0.6318, w1 = 1.997, c0 = -0.6803, c1 = -0.2598): the sampled impulse
  response h(n) = (a0 cos(w0 n/s) + a1 sin(w0 n/s)) exp(-b0 n/s)
  + (c0 cos(w1 n/s) + c1 sin(w1 n/s)) exp(-b1 n/s), s = sigma / spacing in
  voxels, split into a causal filter (n >= 0) and an anti-causal one (n < 0),
  each a 4-pole recursion, normalised to unit DC gain; the signal continues
  as its end samples beyond the borders (each recursion starts from its
  steady state for the end sample);
* ``ResampleImageFilter`` with the identity transform: output voxel i sits at
  continuous input index c = i * n / n'; it is inside iff -0.5 <= c < n - 0.5
  (0 outside); linear interpolation with the neighbours clamped into the
  image, or nearest = floor(c + 0.5).

Agreement with SimpleITK itself is therefore not pinned; what is pinned
(``tests/test_pyramid_oracle.py``) is the mathematics: the impulse response
against the sampled Gaussian within the approximation's error, unit DC gain,
symmetry, second moment sigma^2, preservation of linear ramps by both the
smoothing (interior) and the linear resampling.
"""

import numpy as np

# Deriche (1993), fourth-order Gaussian.
DERICHE_GAUSSIAN = dict(a0=1.680, a1=3.735, b0=1.783, b1=1.723, w0=0.6318, w1=1.997, c0=-0.6803, c1=-0.2598)


def deriche_coefficients(sigma):
    """Coefficients of the recursive Gaussian of standard deviation `sigma`
    (in samples): causal numerator n[0..3], anti-causal numerator m[1..4],
    common denominator 1 + d[1..4] z^-1..z^-4, and the DC normalisation."""
    k = DERICHE_GAUSSIAN
    s = float(sigma)

    def section(a, b, w, beta):
        # a cos(w n / s) e^{-beta n / s} + b sin(w n / s) e^{-beta n / s}, n >= 0:
        # (a + (b sin W - a cos W) e^{-B} z^-1) / (1 - 2 cos W e^{-B} z^-1 + e^{-2B} z^-2)
        W, B = w / s, beta / s
        num = np.array([a, (b * np.sin(W) - a * np.cos(W)) * np.exp(-B)])
        den = np.array([1.0, -2.0 * np.cos(W) * np.exp(-B), np.exp(-2.0 * B)])
        return num, den

    p0, q0 = section(k["a0"], k["a1"], k["w0"], k["b0"])
    p1, q1 = section(k["c0"], k["c1"], k["w1"], k["b1"])
    den = np.convolve(q0, q1)                                   # 1, d1 .. d4
    num = np.convolve(p0, q1)[:4] + np.convolve(p1, q0)[:4]     # n0 .. n3
    n = num
    d = den[1:]
    # anti-causal: sum_{m>=1} h[m] z^m = N(1/z)/D(1/z) - n0, numerator m_k = n_k - n0 d_k
    m = np.array([n[1] - n[0] * d[0], n[2] - n[0] * d[1], n[3] - n[0] * d[2], -n[0] * d[3]])
    dsum = 1.0 + d.sum()
    scale = dsum / (n.sum() + m.sum())
    return n, m, d, scale


def recursive_gaussian_1d(x, sigma, axis=-1):
    """Deriche's recursive Gaussian along `axis` (FP64), the signal
    continued by its end samples."""
    x = np.moveaxis(np.asarray(x, dtype=np.float64), axis, -1)
    if sigma <= 0.0:
        return np.moveaxis(x.copy(), -1, axis)
    n, m, d, scale = deriche_coefficients(sigma)
    L = x.shape[-1]
    dsum = 1.0 + d.sum()
    yp = np.empty_like(x)
    # causal pass, from the steady state of x[0]
    xm = [x[..., 0]] * 3
    ym = [x[..., 0] * (n.sum() / dsum)] * 4
    for i in range(L):
        xi = x[..., i]
        y = n[0] * xi + n[1] * xm[0] + n[2] * xm[1] + n[3] * xm[2] \
            - d[0] * ym[0] - d[1] * ym[1] - d[2] * ym[2] - d[3] * ym[3]
        yp[..., i] = y
        xm = [xi, xm[0], xm[1]]
        ym = [y, ym[0], ym[1], ym[2]]
    out = np.empty_like(x)
    # anti-causal pass, from the steady state of x[L-1]
    xp = [x[..., L - 1]] * 4
    ya = [x[..., L - 1] * (m.sum() / dsum)] * 4
    for i in range(L - 1, -1, -1):
        y = m[0] * xp[0] + m[1] * xp[1] + m[2] * xp[2] + m[3] * xp[3] \
            - d[0] * ya[0] - d[1] * ya[1] - d[2] * ya[2] - d[3] * ya[3]
        out[..., i] = (yp[..., i] + y) * scale
        xp = [x[..., i], xp[0], xp[1], xp[2]]
        ya = [y, ya[0], ya[1], ya[2]]
    return np.moveaxis(out, -1, axis)


def smooth(vol, spacing, sigma=2.0):
    """SmoothingRecursiveGaussian of a [ncomp, nz, ny, nx] volume: sigma in
    physical units, sigma / spacing samples along each axis (x, y, z)."""
    out = np.asarray(vol, dtype=np.float64)
    for a, ax in ((0, 3), (1, 2), (2, 1)):
        out = recursive_gaussian_1d(out, sigma / spacing[a], axis=ax)
    return out


def resample(vol, size_xyz, nearest=False):
    """ResampleImageFilter onto size_xyz with the same origin and extent."""
    v = np.asarray(vol, dtype=np.float64)
    nc, nz, ny, nx = v.shape
    mx, my, mz = size_xyz
    cx = np.arange(mx) * (nx / mx)
    cy = np.arange(my) * (ny / my)
    cz = np.arange(mz) * (nz / mz)

    def idx(c, n):
        if nearest:
            i0 = np.clip(np.floor(c + 0.5).astype(int), 0, n - 1)
            return i0, i0, np.zeros_like(c)
        f = np.floor(c)
        return np.clip(f.astype(int), 0, n - 1), np.clip(f.astype(int) + 1, 0, n - 1), c - f

    x0, x1, ax_ = idx(cx, nx)
    y0, y1, ay_ = idx(cy, ny)
    z0, z1, az_ = idx(cz, nz)
    res = np.zeros((nc, mz, my, mx))
    axb, ayb, azb = ax_[None, None, :], ay_[None, :, None], az_[:, None, None]
    for c in range(nc):
        s = v[c]

        def g(zi, yi, xi):
            return s[np.ix_(zi, yi, xi)]

        c00 = g(z0, y0, x0) + axb * (g(z0, y0, x1) - g(z0, y0, x0))
        c10 = g(z0, y1, x0) + axb * (g(z0, y1, x1) - g(z0, y1, x0))
        c01 = g(z1, y0, x0) + axb * (g(z1, y0, x1) - g(z1, y0, x0))
        c11 = g(z1, y1, x0) + axb * (g(z1, y1, x1) - g(z1, y1, x0))
        c0 = c00 + ayb * (c10 - c00)
        c1 = c01 + ayb * (c11 - c01)
        res[c] = c0 + azb * (c1 - c0)
    inside = ((cz >= -0.5) & (cz < nz - 0.5))[:, None, None] & ((cy >= -0.5) & (cy < ny - 0.5))[None, :, None] \
        & ((cx >= -0.5) & (cx < nx - 0.5))[None, None, :]
    return np.where(inside[None], res, 0.0)


def scale_image(vol, spacing, size_xyz, nearest=False, sigma=2.0):
    """drawing.scale_image of a [ncomp, nz, ny, nx] volume: the smoothed
    volume resampled onto size_xyz; returns it and its spacing."""
    nc, nz, ny, nx = np.shape(vol)
    out = resample(smooth(vol, spacing, sigma), size_xyz, nearest)
    new_spacing = tuple(s * n / m for s, n, m in zip(spacing, (nx, ny, nz), size_xyz))
    return out, new_spacing


def pyramid_sizes(size_xyz, levels):
    """Grid sizes (x, y, z) of the levels, finest first: each level halves the
    previous one (displacements.py:457-459)."""
    sizes = [tuple(int(v) for v in size_xyz)]
    for _ in range(1, levels):
        sizes.append(tuple(v // 2 for v in sizes[-1]))
    return sizes
