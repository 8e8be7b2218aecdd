// dt_pyramid.hip -- kernels of the multi-resolution driver.
//
// The reference builds each pyramid level with drawing.scale_image
// (disptools/drawing.py:23-69): SimpleITK SmoothingRecursiveGaussian with
// sigma = 2 (physical units), then ResampleImageFilter onto a grid of the new
// size with the same origin and physical extent (spacing s * n / n'), linear
// (images, fields) or nearest-neighbour (masks) interpolation, value 0
// outside the input.  Here:
//   k_gauss_axis    the recursive Gaussian along one axis (Deriche's
//                   fourth-order filter: a causal and an anti-causal 4-pole
//                   recursion in FP64, unit DC gain, the signal continued by
//                   its end samples; coefficients from gauss_coefficients on
//                   the host), one line per thread;
//   k_resample      ITK's continuous-index rules: index c = i * s_out / s_in,
//                   inside iff -0.5 <= c < n - 0.5, linear weights with the
//                   neighbours clamped to the image, or nearest = floor(c + 0.5);
//   k_nonzero_u8    a float mask to the solver's byte mask (v != 0, as the
//                   reference's astype(bool), displacements.py:136);
//   k_interleave / k_deinterleave   [3][n] <-> [n][3] (the solver's
//                   component-major field and the [z][y][x][3] arrays of the
//                   Python API).
// Volumes are [ncomp][nz][ny][nx] float, x fastest.
#include <hip/hip_runtime.h>

#include "dt_kernels.hpp"

namespace dt {

// One line per thread: line t of axis `axis` (x fastest among the other two
// coordinates, so that the lines of a wave are adjacent in memory for the y
// and z axes).  `out` receives the causal pass, then the sum of both passes
// times the DC normalisation; `in` and `out` are distinct.
__global__ void __launch_bounds__(256)
k_gauss_axis(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz, int ncomp, int axis,
             GaussCoef g)
{
    const long long n = (long long)nx * ny * nz;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long long stride = axis == 0 ? 1 : (axis == 1 ? nx : (long long)nx * ny);
    const long long per_comp = n / len;
    const long long lines = per_comp * ncomp;
    for (long long t = blockIdx.x * 256ll + threadIdx.x; t < lines; t += (long long)gridDim.x * 256) {
        const long long c = t / per_comp, r = t - c * per_comp;
        long long base;
        if (axis == 0) base = r * nx;                                            // r = z * ny + y
        else if (axis == 1) base = (r / nx) * ((long long)nx * ny) + r % nx;     // r = z * nx + x
        else base = r;                                                           // r = y * nx + x
        const float *x = in + c * n + base;
        float *y = out + c * n + base;
        // causal pass from the steady state of the first sample
        double x0 = x[0];
        double xm1 = x0, xm2 = x0, xm3 = x0;
        double ym1 = x0 * g.nsum, ym2 = ym1, ym3 = ym1, ym4 = ym1;
        for (int i = 0; i < len; ++i) {
            const double xi = x[(long long)i * stride];
            const double yi = g.n[0] * xi + g.n[1] * xm1 + g.n[2] * xm2 + g.n[3] * xm3
                            - g.d[0] * ym1 - g.d[1] * ym2 - g.d[2] * ym3 - g.d[3] * ym4;
            y[(long long)i * stride] = (float)yi;
            xm3 = xm2; xm2 = xm1; xm1 = xi;
            ym4 = ym3; ym3 = ym2; ym2 = ym1; ym1 = yi;
        }
        // anti-causal pass from the steady state of the last sample
        const double xl = x[(long long)(len - 1) * stride];
        double xp1 = xl, xp2 = xl, xp3 = xl, xp4 = xl;
        double yp1 = xl * g.msum, yp2 = yp1, yp3 = yp1, yp4 = yp1;
        for (int i = len - 1; i >= 0; --i) {
            const double yi = g.m[0] * xp1 + g.m[1] * xp2 + g.m[2] * xp3 + g.m[3] * xp4
                            - g.d[0] * yp1 - g.d[1] * yp2 - g.d[2] * yp3 - g.d[3] * yp4;
            float *o = y + (long long)i * stride;
            *o = (float)(((double)*o + yi) * g.scale);
            xp4 = xp3; xp3 = xp2; xp2 = xp1; xp1 = x[(long long)i * stride];
            yp4 = yp3; yp3 = yp2; yp2 = yp1; yp1 = yi;
        }
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

__global__ void __launch_bounds__(256)
k_nonzero_u8(const float *__restrict__ in, unsigned char *__restrict__ out, long long n)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        out[i] = in[i] != 0.0f ? 1 : 0;
}

__global__ void __launch_bounds__(256)
k_interleave(const float *__restrict__ in, float *__restrict__ out, long long n)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        out[3 * i] = in[i];
        out[3 * i + 1] = in[n + i];
        out[3 * i + 2] = in[2 * n + i];
    }
}

__global__ void __launch_bounds__(256)
k_deinterleave(const float *__restrict__ in, float *__restrict__ out, long long n)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        out[i] = in[3 * i];
        out[n + i] = in[3 * i + 1];
        out[2 * n + i] = in[3 * i + 2];
    }
}

static unsigned grid_for(long long work)
{
    long long b = (work + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

hipError_t launch_gauss_axis(const float *in, float *out, int nx, int ny, int nz, int ncomp, int axis,
                             const GaussCoef &g, hipStream_t st)
{
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long long lines = (long long)nx * ny * nz / len * ncomp;
    hipLaunchKernelGGL(k_gauss_axis, dim3(grid_for(lines)), dim3(256), 0, st, in, out, nx, ny, nz, ncomp, axis, g);
    return hipGetLastError();
}

hipError_t launch_nonzero_u8(const float *in, unsigned char *out, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_nonzero_u8, dim3(grid_for(n)), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

hipError_t launch_interleave(const float *in, float *out, long long n, bool to_xyz3, hipStream_t st)
{
    if (to_xyz3) hipLaunchKernelGGL(k_interleave, dim3(grid_for(n)), dim3(256), 0, st, in, out, n);
    else hipLaunchKernelGGL(k_deinterleave, dim3(grid_for(n)), dim3(256), 0, st, in, out, n);
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
