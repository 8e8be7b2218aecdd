// dt_levels.cpp -- the multi-resolution driver of displacement() on the
// device: the image pyramid (drawing.scale_image) and the solve from the
// coarsest level to the finest (reference disptools/displacements.py:454-483).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "disptools_hip.h"
#include "dt_host.hpp"

namespace dt {

namespace {

// Deriche (1993), fourth-order approximation of the Gaussian, h(x) for x >= 0
// in units of sigma: (a0 cos(w0 x) + a1 sin(w0 x)) e^{-b0 x}
//                  + (c0 cos(w1 x) + c1 sin(w1 x)) e^{-b1 x}.
constexpr double DA0 = 1.680, DA1 = 3.735, DB0 = 1.783, DB1 = 1.723;
constexpr double DW0 = 0.6318, DW1 = 1.997, DC0 = -0.6803, DC1 = -0.2598;

// z-transform of a cos(w n / s) e^{-beta n / s} + b sin(w n / s) e^{-beta n / s},
// n >= 0: numerator p[0..1], denominator q[0..2] (q[0] = 1).
void section(double a, double b, double w, double beta, double s, double p[2], double q[3])
{
    const double W = w / s, B = beta / s;
    p[0] = a;
    p[1] = (b * std::sin(W) - a * std::cos(W)) * std::exp(-B);
    q[0] = 1.0;
    q[1] = -2.0 * std::cos(W) * std::exp(-B);
    q[2] = std::exp(-2.0 * B);
}

}  // namespace

GaussCoef gauss_coefficients(double sigma)
{
    double p0[2], q0[3], p1[2], q1[3];
    section(DA0, DA1, DW0, DB0, sigma, p0, q0);
    section(DC0, DC1, DW1, DB1, sigma, p1, q1);
    // causal part H+(z) = p0/q0 + p1/q1 = (p0 q1 + p1 q0) / (q0 q1)
    double den[5] = {0, 0, 0, 0, 0}, num[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) den[i + j] += q0[i] * q1[j];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 3; ++j) num[i + j] += p0[i] * q1[j] + p1[i] * q0[j];
    GaussCoef g;
    double nsum = 0.0, msum = 0.0, dsum = 1.0;
    for (int k = 0; k < 4; ++k) {
        g.n[k] = num[k];
        g.d[k] = den[k + 1];
        nsum += g.n[k];
        dsum += g.d[k];
    }
    // anti-causal part: sum_{m >= 1} h[m] z^m = N(1/z) / D(1/z) - n0
    for (int k = 0; k < 4; ++k) {
        g.m[k] = (k < 3 ? g.n[k + 1] : 0.0) - g.n[0] * g.d[k];
        msum += g.m[k];
    }
    g.nsum = nsum / dsum;
    g.msum = msum / dsum;
    g.scale = dsum / (nsum + msum);
    return g;
}

int scale_volume(int nx, int ny, int nz, int ncomp, float dx, float dy, float dz, const float *in, int mx, int my,
                 int mz, float sigma, bool nearest, float *out, float *t0, float *t1, hipStream_t st)
{
    const int len[3] = {nx, ny, nz};
    const float sp[3] = {dx, dy, dz};
    const float *src = in;
    float *bufs[2] = {t0, t1};
    int k = 0;
    for (int a = 0; a < 3; ++a) {
        const double sv = (double)sigma / (double)sp[a];
        // a line of one sample, or no smoothing asked for: the filter is the identity
        if (!(sv > 0.0) || len[a] < 2) continue;
        float *dst = bufs[k++ & 1];
        DT_HIP(launch_gauss_axis(src, dst, nx, ny, nz, ncomp, a, gauss_coefficients(sv), st));
        src = dst;
    }
    DT_HIP(launch_resample(src, out, nx, ny, nz, mx, my, mz, ncomp, (float)((double)nx / (double)mx),
                           (float)((double)ny / (double)my), (float)((double)nz / (double)mz), nearest ? 1 : 0, st));
    return 0;
}

}  // namespace dt

using namespace dt;

namespace {

struct DeviceBuffers {
    std::vector<void *> ptrs;
    ~DeviceBuffers()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *take(size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptrs.push_back(p);
        return (T *)p;
    }
};

struct StreamGuard {
    hipStream_t st = nullptr;
    ~StreamGuard()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
};

}  // namespace

extern "C" int dt_displacement_levels(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                                      const float *J, const float *mask, const float *guess, size_t gx, size_t gy,
                                      size_t gz, int levels, const long long *it_max, const dt_params *params,
                                      float *field, dt_solve_info *info)
{
    disptools_clear_error();
    if (algorithm != DT_GRADIENT && algorithm != DT_GREEDY && algorithm != DT_MATCHING)
        DT_FAIL("unknown algorithm");
    if (!J || !mask || !field || !params || !it_max) DT_FAIL("null buffer");
    if (levels < 1) DT_FAIL("levels must be a positive integer");
    if (nx < 1 || ny < 1 || nz < 1) DT_FAIL("empty volume");
    if (!(dx > 0.0f) || !(dy > 0.0f) || !(dz > 0.0f)) DT_FAIL("spacing must be positive");
    if (guess && (gx < 1 || gy < 1 || gz < 1)) DT_FAIL("empty initial guess");
    // level sizes, finest first: each halves the previous one (displacements.py:457-459)
    std::vector<int> sx(levels), sy(levels), sz(levels);
    std::vector<float> spx(levels), spy(levels), spz(levels);
    sx[0] = (int)nx; sy[0] = (int)ny; sz[0] = (int)nz;
    for (int l = 1; l < levels; ++l) {
        sx[l] = sx[l - 1] / 2; sy[l] = sy[l - 1] / 2; sz[l] = sz[l - 1] / 2;
        if (sx[l] < 1 || sy[l] < 1 || sz[l] < 1) {
            char m[160];
            std::snprintf(m, sizeof m, "pyramid level %d of size (%d, %d, %d) is empty", l, sx[l], sy[l], sz[l]);
            DT_FAIL(m);
        }
    }
    // spacing of a level: same physical extent as the finest (drawing.py:55-57)
    for (int l = 0; l < levels; ++l) {
        spx[l] = (float)((double)dx * (double)nx / (double)sx[l]);
        spy[l] = (float)((double)dy * (double)ny / (double)sy[l]);
        spz[l] = (float)((double)dz * (double)nz / (double)sz[l]);
    }
    const long long n0 = (long long)nx * ny * nz;
    const long long ng = guess ? (long long)gx * gy * gz : 0;
    StreamGuard sg;
    DT_HIP(hipStreamCreateWithFlags(&sg.st, hipStreamNonBlocking));
    hipStream_t st = sg.st;
    DeviceBuffers B;
    // per level: target, float mask, byte mask; two field volumes and two
    // scratch volumes of the finest size (fields and guesses are no larger)
    const long long big = std::max(3 * n0, 3 * ng);
    std::vector<float *> dJ(levels), dMf(levels);
    std::vector<unsigned char *> dM(levels);
    for (int l = 0; l < levels; ++l) {
        const long long nl = (long long)sx[l] * sy[l] * sz[l];
        dJ[l] = B.take<float>(nl);
        dMf[l] = B.take<float>(nl);
        dM[l] = B.take<unsigned char>(nl);
        if (!dJ[l] || !dMf[l] || !dM[l]) DT_FAIL("cannot allocate device memory for the pyramid");
    }
    float *f[2] = {B.take<float>(big), B.take<float>(big)};
    float *t0 = B.take<float>(big), *t1 = B.take<float>(big);
    if (!f[0] || !f[1] || !t0 || !t1) DT_FAIL("cannot allocate device memory for the pyramid");

    DT_HIP(hipMemcpyAsync(dJ[0], J, n0 * sizeof(float), hipMemcpyHostToDevice, st));
    DT_HIP(hipMemcpyAsync(dMf[0], mask, n0 * sizeof(float), hipMemcpyHostToDevice, st));
    // every coarser level is scaled from the finest target and mask (:460-461)
    for (int l = 1; l < levels; ++l) {
        if (scale_volume(sx[0], sy[0], sz[0], 1, dx, dy, dz, dJ[0], sx[l], sy[l], sz[l], 2.0f, false, dJ[l], t0, t1,
                         st) ||
            scale_volume(sx[0], sy[0], sz[0], 1, dx, dy, dz, dMf[0], sx[l], sy[l], sz[l], 2.0f, true, dMf[l], t0, t1,
                         st))
            return -1;
    }
    for (int l = 0; l < levels; ++l)
        DT_HIP(launch_nonzero_u8(dMf[l], dM[l], (long long)sx[l] * sy[l] * sz[l], st));

    // the field entering the coarsest level: the initial guess scaled onto
    // its grid (:476), or zero
    int cur = 0;
    bool have = false;
    int fx = 0, fy = 0, fz = 0;
    float fdx = dx, fdy = dy, fdz = dz;
    if (guess) {
        DT_HIP(hipMemcpyAsync(t0, guess, 3 * ng * sizeof(float), hipMemcpyHostToDevice, st));
        DT_HIP(launch_interleave(t0, f[cur], ng, false, st));   // [z][y][x][3] -> [3][z][y][x]
        fx = (int)gx; fy = (int)gy; fz = (int)gz;
        have = true;
    }
    dt_solve_info last{};
    for (int k = 0; k < levels; ++k) {
        const int l = levels - 1 - k;
        const long long nl = (long long)sx[l] * sy[l] * sz[l];
        float *dst = f[cur ^ 1];
        if (have) {
            if (scale_volume(fx, fy, fz, 3, fdx, fdy, fdz, f[cur], sx[l], sy[l], sz[l], 2.0f, false, dst, t0, t1, st))
                return -1;
        } else {
            DT_HIP(hipMemsetAsync(dst, 0, 3 * nl * sizeof(float), st));
        }
        cur ^= 1;
        dt_params p = *params;
        p.it_max = it_max[k];
        if (algorithm == DT_MATCHING) {
            dt_matching_info mi{};
            if (dt_matching_device(sx[l], sy[l], sz[l], spx[l], spy[l], spz[l], dJ[l], p.epsilon, p.delta, p.zeta,
                                   p.it_max, f[cur], &mi, nullptr, 0, st))
                return -1;
            std::memset(&last, 0, sizeof last);
            last.iterations = mi.iterations;
        } else {
            if (dt_solve_device(algorithm, sx[l], sy[l], sz[l], spx[l], spy[l], spz[l], dJ[l], dM[l], &p, f[cur], &last,
                                nullptr, 0, st))
                return -1;
        }
        fx = sx[l]; fy = sy[l]; fz = sz[l];
        fdx = spx[l]; fdy = spy[l]; fdz = spz[l];
        have = true;
    }
    // [3][z][y][x] -> [z][y][x][3], one download
    DT_HIP(launch_interleave(f[cur], t0, n0, true, st));
    DT_HIP(hipMemcpyAsync(field, t0, 3 * n0 * sizeof(float), hipMemcpyDeviceToHost, st));
    DT_HIP(hipStreamSynchronize(st));
    if (info) *info = last;
    return 0;
}
