"""Synthetic inputs of the benchmark configurations (seeded, numpy PCG64).

Volumes are numpy arrays indexed [z, y, x]; spacing tuples are (dx, dy, dz).

* ``sphere_case``      -- 40^3 sphere: target 1.1 inside a ball of radius 10
                          centred at n//2 and 1.0 outside, mask = the ball, as
                          built by the reference's ``drawing.create_sphere``
                          (drawing.py:123-242, Euclidean rho <= radius).
* ``smooth_target``    -- exp(g), g i.i.d. N(0, sigma_g^2) smoothed by a Gaussian
                          of ``smooth`` voxels (edge mode "nearest"), clipped.
* ``ellipsoid_mask``   -- union of k ~ U{3..8} ellipsoids, semi-axes U[8, 30]
                          voxels, centres U[0.2 n, 0.8 n], Haar-random rotations.
* ``volume_128``       -- 128^3, spacing (1, 1, 1.5), target as above with
                          sigma_g = 0.05, smoothing 4, clip [0.8, 1.25].
* ``batch_uneven``     -- 64 volumes, each axis U{48..160}, spacing 1.
* ``precision_256``    -- 256^3, spacing (0.8, 0.8, 1.0), sigma_g = 0.08,
                          smoothing 6, clip [0.7, 1.4], two spheres of radii 40
                          and 60 centred at (x, y, z) = (96, 128, 128), (160, 128, 128).
* ``serving_stream``   -- Poisson arrivals with log-normal cube sides.
* ``volume_case_torch`` -- the ``volume_case`` distribution drawn on a torch
                          device (noise from torch's generator, so the values
                          differ from ``volume_case`` for the same seed): used
                          where hundreds of volumes are needed quickly.
"""

import math

import numpy as np
from scipy.ndimage import gaussian_filter

SEED_128 = 12345


def sphere_case(n=40, radius=10.0, inside=1.1, outside=1.0):
    c = n // 2
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    ball = r <= radius
    target = np.where(ball, inside, outside).astype(np.float32)
    return target, ball.copy(), (1.0, 1.0, 1.0)


def smooth_target(shape, rng, sigma_g=0.05, smooth=4.0, clip=(0.8, 1.25)):
    g = rng.normal(0.0, sigma_g, size=shape)
    g = gaussian_filter(g, smooth, mode="nearest", truncate=4.0)
    return np.clip(np.exp(g), clip[0], clip[1]).astype(np.float32)


def _haar_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoid_mask(shape, rng, k_range=(3, 8), axes_range=(8.0, 30.0)):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([x, y, z], axis=-1).astype(np.float64)
    mask = np.zeros(shape, bool)
    k = int(rng.integers(k_range[0], k_range[1] + 1))
    n = np.array([nx, ny, nz], np.float64)
    for _ in range(k):
        semi = rng.uniform(axes_range[0], axes_range[1], size=3)
        centre = rng.uniform(0.2 * n, 0.8 * n)
        rot = _haar_rotation(rng)
        local = (pts - centre) @ rot        # R^T (p - c) for row vectors
        mask |= np.sum((local / semi) ** 2, axis=-1) <= 1.0
    return mask


def volume_case(shape, seed, sigma_g=0.05, smooth=4.0, clip=(0.8, 1.25)):
    rng = np.random.default_rng(seed)
    target = smooth_target(shape, rng, sigma_g, smooth, clip)
    mask = ellipsoid_mask(shape, rng)
    return target, mask


def volume_128(seed=SEED_128):
    target, mask = volume_case((128, 128, 128), seed)
    return target, mask, (1.0, 1.0, 1.5)


def batch_uneven(count=64, seed=2024, lo=48, hi=160, workers=1):
    """Volumes of the batch configuration; `workers` threads build them (each
    volume has its own seed, so the result does not depend on `workers`)."""
    rng = np.random.default_rng(seed)
    shapes = [tuple(int(v) for v in rng.integers(lo, hi + 1, size=3)) for _ in range(count)]

    def one(i):
        target, mask = volume_case(shapes[i], seed * 1000 + i)
        return target, mask, (1.0, 1.0, 1.0)

    if workers <= 1:
        return [one(i) for i in range(count)]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(one, range(count)))


def precision_256():
    rng = np.random.default_rng(256)
    shape = (256, 256, 256)
    target = smooth_target(shape, rng, sigma_g=0.08, smooth=6.0, clip=(0.7, 1.4))
    z, y, x = np.meshgrid(np.arange(256), np.arange(256), np.arange(256), indexing="ij")
    mask = ((x - 96) ** 2 + (y - 128) ** 2 + (z - 128) ** 2 <= 40 ** 2) | \
           ((x - 160) ** 2 + (y - 128) ** 2 + (z - 128) ** 2 <= 60 ** 2)
    return target, mask, (0.8, 0.8, 1.0)


def serving_stream(count, qps, seed=7, median=96, sigma_ln=0.35, lo=32, hi=256, greedy_frac=0.2):
    """Requests (arrival time, side, algorithm, seed) of a Poisson stream."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / qps)
        side = int(np.clip(round(median * math.exp(rng.normal(0.0, sigma_ln))), lo, hi))
        alg = "greedy" if rng.random() < greedy_frac else "gradient"
        out.append(dict(arrival=t, side=side, algorithm=alg, seed=int(seed * 100000 + i)))
    return out


def _gauss_kernel(sigma, truncate=4.0):
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return (k / k.sum()).astype(np.float32), r


def volume_case_torch(shape, seed, device, sigma_g=0.05, smooth=4.0, clip=(0.8, 1.25)):
    """Target and mask of ``volume_case``'s distribution, built on `device`:
    separable Gaussian smoothing with edge replication ("nearest"), ellipsoid
    parameters from numpy's PCG64 (seed), membership evaluated on the device."""
    import torch
    import torch.nn.functional as F
    nz, ny, nx = shape
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    g = torch.randn(shape, generator=gen, device=device, dtype=torch.float32) * sigma_g
    k, r = _gauss_kernel(smooth)
    v = g
    for axis in range(3):
        # one axis at a time: taps summed in a fixed order over shifted views of
        # the edge-replicated volume (no convolution library: its algorithm
        # search per new shape costs more than the solves being served)
        pad = [0, 0, 0, 0, 0, 0]
        pad[2 * (2 - axis)] = r
        pad[2 * (2 - axis) + 1] = r
        vp = F.pad(v[None, None], pad, mode="replicate")[0, 0]
        n = v.shape[axis]
        acc = torch.zeros_like(v)
        for j, kj in enumerate(k):
            acc.add_(vp.narrow(axis, j, n), alpha=float(kj))
        v = acc
    target = torch.clamp(torch.exp(v), clip[0], clip[1]).contiguous()

    rng = np.random.default_rng(seed)
    z = torch.arange(nz, device=device, dtype=torch.float32).view(nz, 1, 1)
    y = torch.arange(ny, device=device, dtype=torch.float32).view(1, ny, 1)
    x = torch.arange(nx, device=device, dtype=torch.float32).view(1, 1, nx)
    mask = torch.zeros(shape, dtype=torch.bool, device=device)
    kcount = int(rng.integers(3, 9))
    n = np.array([nx, ny, nz], np.float64)
    for _ in range(kcount):
        semi = rng.uniform(8.0, 30.0, size=3)
        centre = rng.uniform(0.2 * n, 0.8 * n)
        rot = _haar_rotation(rng)
        px, py, pz = x - float(centre[0]), y - float(centre[1]), z - float(centre[2])
        acc = 0.0
        for j in range(3):
            lj = px * float(rot[0, j]) + py * float(rot[1, j]) + pz * float(rot[2, j])
            acc = acc + (lj / float(semi[j])) ** 2
        mask |= acc <= 1.0
    return target, mask.to(torch.uint8).contiguous()
