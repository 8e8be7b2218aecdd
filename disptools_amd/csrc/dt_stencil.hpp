// dt_stencil.hpp -- finite-difference stencils of the Jacobian determinant.
//
// Arithmetic follows the reference (src/headers/jacobian_macros.h:4-148,
// src/headers/disptools.h:186-199): sample differences are taken in float,
// weights and the determinant in the compute type CT.  With CT = double and
// contraction disabled the result is bit-identical to the reference C build;
// CT = float is the fast solver path.
#pragma once

#include <hip/hip_runtime.h>

namespace dt {

// Stencil kinds along one axis.
enum Kind : int { K_C4 = 0, K_C2 = 1, K_FWD = 2, K_BWD = 3 };

// Stencil that the reference jacobian() leaves at a voxel (src/jacobian.c):
// the outer face uses one-sided differences on the axes where it touches the
// face and 2nd-order central elsewhere (:112-221); the penultimate layer uses
// 2nd-order central on all axes (:223-241); everything else 4th-order central
// (:87-110).
__host__ __device__ __forceinline__ void classify(int x, int y, int z, int nx, int ny, int nz, int k[3])
{
    const bool ox = (x == 0) | (x == nx - 1);
    const bool oy = (y == 0) | (y == ny - 1);
    const bool oz = (z == 0) | (z == nz - 1);
    if (ox | oy | oz) {
        k[0] = x == 0 ? K_FWD : (x == nx - 1 ? K_BWD : K_C2);
        k[1] = y == 0 ? K_FWD : (y == ny - 1 ? K_BWD : K_C2);
        k[2] = z == 0 ? K_FWD : (z == nz - 1 ? K_BWD : K_C2);
    } else {
        const bool pen = (x == 1) | (x == nx - 2) | (y == 1) | (y == ny - 2) | (z == 1) | (z == nz - 2);
        k[0] = k[1] = k[2] = pen ? K_C2 : K_C4;
    }
}

// Weights of f(-2) .. f(+2) of each stencil kind (times 1/spacing).
__device__ __forceinline__ void stencil_weights(int kind, float w[5])
{
    const bool c4 = kind == K_C4, c2 = kind == K_C2, fw = kind == K_FWD;
    w[0] = c4 ? (float)(1.0 / 12.0) : (c2 || fw ? 0.0f : 0.5f);
    w[1] = c4 ? (float)(-2.0 / 3.0) : (c2 ? -0.5f : (fw ? 0.0f : -2.0f));
    w[2] = c4 || c2 ? 0.0f : (fw ? -1.5f : 1.5f);
    w[3] = c4 ? (float)(2.0 / 3.0) : (c2 ? 0.5f : (fw ? 2.0f : 0.0f));
    w[4] = c4 ? (float)(-1.0 / 12.0) : (c2 ? 0.0f : (fw ? -0.5f : 0.0f));
}

// Derivative along one axis from the five samples f(-2) .. f(+2).
template <typename CT>
__host__ __device__ __forceinline__ CT dstencil(int kind, float m2, float m1, float c0, float p1, float p2, float inv)
{
    switch (kind) {
    case K_C4:
        return ((CT)(p1 - m1) * (CT)(2.0 / 3.0) - (CT)(p2 - m2) * (CT)(1.0 / 12.0)) * (CT)inv;
    case K_C2:
        return (CT)((p1 - m1) * inv) * (CT)0.5;
    case K_FWD:
        return ((CT)-1.5 * (CT)c0 + (CT)2.0 * (CT)p1 - (CT)0.5 * (CT)p2) * (CT)inv;
    default:
        return ((CT)1.5 * (CT)c0 - (CT)2.0 * (CT)m1 + (CT)0.5 * (CT)m2) * (CT)inv;
    }
}

template <typename CT>
__host__ __device__ __forceinline__ CT dstencil_c4(float m2, float m1, float p1, float p2, float inv)
{
    return ((CT)(p1 - m1) * (CT)(2.0 / 3.0) - (CT)(p2 - m2) * (CT)(1.0 / 12.0)) * (CT)inv;
}

// dstencil<double> for lanes whose stencil along the axis is 4th- or
// 2nd-order central (c4 picks): both forms rounded as dstencil's, one kept.
__host__ __device__ __forceinline__ double dstencil_c4c2(bool c4, float m2, float m1, float p1, float p2, float inv)
{
    const double a = dstencil_c4<double>(m2, m1, p1, p2, inv);
    const double b = (double)((p1 - m1) * inv) * 0.5;
    return c4 ? a : b;
}

// dstencil<double> without branches, for lanes of one wave that hold
// different stencil kinds.  Every kind is evaluated with the operations of
// dstencil in the same order -- a - b*w as a + b*(-w), which rounds the same
// -- from operands and weights picked by the kind:
//   C4   ((p1 - m1) * 2/3 + (p2 - m2) * -1/12) * inv
//   FWD  ((c0 * -1.5 + p1 * 2) + p2 * -0.5) * inv
//   BWD  ((c0 * 1.5 + m1 * -2) + m2 * 0.5) * inv
//   C2   (float)((p1 - m1) * inv) * 0.5
// The selectors depend on the kind only; a voxel forms them once per axis
// and shares them between the three field components.
struct StencilSel {
    bool c4, fwd, c2;
    double wa, wb, wc;
};

__host__ __device__ __forceinline__ StencilSel stencil_sel(int kind)
{
    StencilSel s;
    s.c4 = kind == K_C4;
    s.fwd = kind == K_FWD;
    s.c2 = kind == K_C2;
    s.wa = s.c4 ? 2.0 / 3.0 : (s.fwd ? -1.5 : 1.5);
    s.wb = s.c4 ? -(1.0 / 12.0) : (s.fwd ? 2.0 : -2.0);
    s.wc = s.fwd ? -0.5 : 0.5;
    return s;
}

__host__ __device__ __forceinline__ double dstencil_sel(const StencilSel &s, float m2, float m1, float c0, float p1,
                                                        float p2, float inv)
{
    const float d1 = p1 - m1, d2 = p2 - m2;
    const double a = (double)(s.c4 ? d1 : c0);
    const double b = (double)(s.c4 ? d2 : (s.fwd ? p1 : m1));
    const double c = (double)(s.fwd ? p2 : m2);
    const double ab = a * s.wa + b * s.wb;
    const double abc = ab + c * s.wc;
    const double r = (s.c4 ? ab : abc) * (double)inv;
    const double q = (double)(d1 * inv) * 0.5;
    return s.c2 ? q : r;
}

// det(D + I), cofactor expansion along the first row (disptools.h:186-199).
template <typename CT>
__host__ __device__ __forceinline__ CT det_plus_identity(const CT d[3][3])
{
    const CT a11 = d[0][0] + (CT)1.0, a12 = d[0][1], a13 = d[0][2];
    const CT a21 = d[1][0], a22 = d[1][1] + (CT)1.0, a23 = d[1][2];
    const CT a31 = d[2][0], a32 = d[2][1], a33 = d[2][2] + (CT)1.0;
    return a11 * (a22 * a33 - a23 * a32)
         - a12 * (a21 * a33 - a23 * a31)
         + a13 * (a21 * a32 - a22 * a31);
}

// Fused multiply-add in the arithmetic type (__builtin_fma is the double one).
__device__ __forceinline__ float fma_ct(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_ct(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a * b - c * d as one fused multiply-add over the rounded product c * d.
template <typename CT>
__device__ __forceinline__ CT fms2(CT a, CT b, CT c, CT d)
{
#pragma clang fp contract(off)
    return fma_ct(a, b, -(c * d));
}

// det(I + D) - 1 = tr D + (sum of the principal 2x2 minors of D) + det D,
// summed from the smallest terms up.  For |D| << 1 this keeps J - 1 to a
// relative accuracy of the compute type instead of an absolute error of half
// an ulp of 1.0.  Every fused multiply-add is written out and nothing else
// is contracted, so every compilation of a pass rounds it the same way.
template <typename CT>
__device__ __forceinline__ CT det_plus_identity_minus_one(const CT d[3][3])
{
#pragma clang fp contract(off)
    const CT m01 = fms2(d[0][0], d[1][1], d[0][1], d[1][0]);
    const CT m02 = fms2(d[0][0], d[2][2], d[0][2], d[2][0]);
    const CT m12 = fms2(d[1][1], d[2][2], d[1][2], d[2][1]);
    const CT c1 = fms2(d[1][0], d[2][2], d[1][2], d[2][0]);
    const CT c2 = fms2(d[1][0], d[2][1], d[1][1], d[2][0]);
    const CT det = fma_ct(d[0][0], m12, fma_ct(-d[0][1], c1, d[0][2] * c2));
    return ((det + (m01 + m02 + m12)) + (d[0][0] + d[1][1] + d[2][2]));
}

// Voxel error from dJ = J - 1 and dT = T - 1 (exact in float for T in [0.5, 2]):
// err = J - T = dJ - dT, then the same branches as voxel_error.
__device__ __forceinline__ float voxel_error_delta(float dJ, float dT, bool in_mask, float tol)
{
    const float err = dJ - dT;
    if (in_mask) return err;
    const float a = fabsf(err);
    return a < tol ? 0.0f : copysignf(a - tol, err);
}

// Voxel error (src/headers/error.h:128-142): full error inside the mask,
// soft-thresholded by `tol` outside it.
__device__ __forceinline__ float voxel_error(float J, float T, bool in_mask, float tol)
{
    const float err = J - T;
    if (in_mask) return err;
    const float a = fabsf(err);
    return a < tol ? 0.0f : copysignf(a - tol, err);
}

}  // namespace dt
