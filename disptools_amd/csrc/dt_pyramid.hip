// dt_pyramid.hip -- resampling kernels of the multi-resolution driver.
//
// The reference builds each pyramid level with drawing.scale_image
// (disptools/drawing.py:23-69): SimpleITK SmoothingRecursiveGaussian with
// sigma = 2 (physical units), then ResampleImageFilter onto a grid of the new
// size with the same origin and physical extent (spacing s * n / n'), linear
// (images, fields) or nearest-neighbour (masks) interpolation, value 0
// outside the input.  Here:
//   k_smooth_axis   separable Gaussian along one axis, sigma in voxels,
//                   truncated at 4 sigma, edge samples replicated (the
//                   recursive filter's zero-flux start);
//   k_resample      ITK's continuous-index rules: index c = i * s_out / s_in,
//                   inside iff -0.5 <= c < n - 0.5, linear weights with the
//                   neighbours clamped to the image, or nearest = floor(c + 0.5).
// Volumes are [ncomp][nz][ny][nx] float, x fastest.
#include <hip/hip_runtime.h>

#include "dt_kernels.hpp"

namespace dt {

__global__ void __launch_bounds__(256)
k_smooth_axis(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz, int ncomp, int axis,
              const float *__restrict__ w, int radius)
{
    const long long n = (long long)nx * ny * nz;
    const long long total = n * ncomp;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long long stride = axis == 0 ? 1 : (axis == 1 ? nx : (long long)nx * ny);
    for (long long t = blockIdx.x * 256ll + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const long long v = t % n;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        const int p = axis == 0 ? x : (axis == 1 ? y : z);
        const float *base = in + (t - v) + (v - (long long)p * stride);
        float acc = 0.0f;
        for (int k = -radius; k <= radius; ++k) {
            const int q = min(max(p + k, 0), len - 1);
            acc = fmaf(w[k + radius], base[(long long)q * stride], acc);
        }
        out[t] = acc;
    }
}

__global__ void __launch_bounds__(256)
k_resample(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz, int mx, int my, int mz,
           int ncomp, float rx, float ry, float rz, int nearest)
{
    const long long m = (long long)mx * my * mz;
    const long long n = (long long)nx * ny * nz;
    for (long long t = blockIdx.x * 256ll + threadIdx.x; t < m * ncomp; t += (long long)gridDim.x * 256) {
        const int c = (int)(t / m);
        const long long v = t - (long long)c * m;
        const int x = (int)(v % mx), y = (int)((v / mx) % my), z = (int)(v / ((long long)mx * my));
        const float cx = (float)x * rx, cy = (float)y * ry, cz = (float)z * rz;
        const bool inside = cx >= -0.5f && cx < nx - 0.5f && cy >= -0.5f && cy < ny - 0.5f && cz >= -0.5f &&
                            cz < nz - 0.5f;
        const float *src = in + (long long)c * n;
        float r = 0.0f;
        if (inside) {
            if (nearest) {
                const int ix = min(max((int)floorf(cx + 0.5f), 0), nx - 1);
                const int iy = min(max((int)floorf(cy + 0.5f), 0), ny - 1);
                const int iz = min(max((int)floorf(cz + 0.5f), 0), nz - 1);
                r = src[((long long)iz * ny + iy) * nx + ix];
            } else {
                const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
                const float ax = cx - fx, ay = cy - fy, az = cz - fz;
                const int x0 = min(max((int)fx, 0), nx - 1), x1 = min(max((int)fx + 1, 0), nx - 1);
                const int y0 = min(max((int)fy, 0), ny - 1), y1 = min(max((int)fy + 1, 0), ny - 1);
                const int z0 = min(max((int)fz, 0), nz - 1), z1 = min(max((int)fz + 1, 0), nz - 1);
                auto at = [&](int xx, int yy, int zz) { return src[((long long)zz * ny + yy) * nx + xx]; };
                const float c00 = at(x0, y0, z0) + ax * (at(x1, y0, z0) - at(x0, y0, z0));
                const float c10 = at(x0, y1, z0) + ax * (at(x1, y1, z0) - at(x0, y1, z0));
                const float c01 = at(x0, y0, z1) + ax * (at(x1, y0, z1) - at(x0, y0, z1));
                const float c11 = at(x0, y1, z1) + ax * (at(x1, y1, z1) - at(x0, y1, z1));
                const float c0 = c00 + ay * (c10 - c00), c1 = c01 + ay * (c11 - c01);
                r = c0 + az * (c1 - c0);
            }
        }
        out[t] = r;
    }
}

static unsigned grid_for(long long work)
{
    long long b = (work + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

hipError_t launch_smooth_axis(const float *in, float *out, int nx, int ny, int nz, int ncomp, int axis,
                              const float *d_w, int radius, hipStream_t st)
{
    hipLaunchKernelGGL(k_smooth_axis, dim3(grid_for((long long)nx * ny * nz * ncomp)), dim3(256), 0, st, in, out,
                       nx, ny, nz, ncomp, axis, d_w, radius);
    return hipGetLastError();
}

hipError_t launch_resample(const float *in, float *out, int nx, int ny, int nz, int mx, int my, int mz, int ncomp,
                           float rx, float ry, float rz, int nearest, hipStream_t st)
{
    hipLaunchKernelGGL(k_resample, dim3(grid_for((long long)mx * my * mz * ncomp)), dim3(256), 0, st, in, out, nx,
                       ny, nz, mx, my, mz, ncomp, rx, ry, rz, nearest);
    return hipGetLastError();
}

}  // namespace dt
