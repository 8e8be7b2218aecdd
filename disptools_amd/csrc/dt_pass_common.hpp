// dt_pass_common.hpp -- buffer addressing and per-launch constants shared by
// the solver pass kernels.
#pragma once

#include "dt_device.hpp"

namespace dt {

constexpr int RSRC_WORD3 = 0x00020000;   // raw buffer, gfx9-family data format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_WORD3);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, float v, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, 0);
}

struct PassParams {
    int nx, ny, nz;
    float inv[3];     // 1 / spacing
    float c4a[3];     // (2/3) * inv
    float c4b[3];     // (1/12) * inv
    float c2[3];      // 0.5 * inv
    float tol, delta, zeta;
};

__device__ __forceinline__ PassParams load_params(const Inst &I)
{
    PassParams P;
    P.nx = uni(I.nx); P.ny = uni(I.ny); P.nz = uni(I.nz);
    P.inv[0] = uni(I.ix); P.inv[1] = uni(I.iy); P.inv[2] = uni(I.iz);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        P.c4a[a] = (float)(2.0 / 3.0) * P.inv[a];
        P.c4b[a] = (float)(1.0 / 12.0) * P.inv[a];
        P.c2[a] = 0.5f * P.inv[a];
    }
    P.tol = uni(I.p.tolerance); P.delta = uni(I.p.delta); P.zeta = uni(I.p.zeta);
    return P;
}

// J - 1 at voxel (gx, gy, zz) from its samples sm[c][axis][m2, m1, p1, p2] and
// centre v0[c], with the stencil the region map assigns there.  Waves with no
// boundary lane take the 4th-order form; the others a weighted 5-point form
// whose weights encode each lane's stencil (no divergent branches).
template <typename CT>
__device__ __forceinline__ float jacobian_delta(const PassParams &P, int gx, int gy, int zz,
                                                const float (&sm)[3][3][4], const float (&v0)[3])
{
    CT d4[3][3];
    const bool fast = gx >= 2 && gx <= P.nx - 3 && gy >= 2 && gy <= P.ny - 3 && zz >= 2 && zz <= P.nz - 3;
    if (__builtin_amdgcn_ballot_w64(!fast) == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a)
                d4[c][a] = (CT)(sm[c][a][2] - sm[c][a][1]) * (CT)P.c4a[a] -
                           (CT)(sm[c][a][3] - sm[c][a][0]) * (CT)P.c4b[a];
    } else {
        int kind[3];
        classify(gx, gy, zz, P.nx, P.ny, P.nz, kind);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float w[5];
            stencil_weights(kind[a], w);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                d4[c][a] = ((CT)w[0] * sm[c][a][0] + (CT)w[1] * sm[c][a][1] + (CT)w[2] * v0[c] +
                            (CT)w[3] * sm[c][a][2] + (CT)w[4] * sm[c][a][3]) * (CT)P.inv[a];
        }
    }
    return (float)det_plus_identity_minus_one<CT>(d4);
}

// Stencil coefficients of a column (gx, gy) on the planes away from the z
// faces (2 <= z <= nz - 3), where the region map depends on x and y alone.
// Every stencil kind is  d = A (p1 - m1) + B (p2 - m2) + G c  (1/spacing
// folded in), because samples outside the volume are zero:
//   4th-order central  A = 2/3,  B = -1/12, G = 0
//   2nd-order central  A = 1/2,  B = 0,     G = 0
//   forward  (m1 = m2 = 0)  A = 2, B = -1/2, G = -3/2   (-3/2 c + 2 p1 - 1/2 p2)
//   backward (p1 = p2 = 0)  A = 2, B = -1/2, G = +3/2   (3/2 c - 2 m1 + 1/2 m2)
struct ColCoef {
    float A[3], B[3], G[3];
};

__device__ __forceinline__ ColCoef column_coefficients(const PassParams &P, int gx, int gy)
{
    int kind[3];
    classify(gx, gy, 2, P.nx, P.ny, 5, kind);   // a plane away from the z faces
    ColCoef K;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int k = kind[a];
        const float inv = P.inv[a];
        K.A[a] = k == K_C4 ? P.c4a[a] : (k == K_C2 ? P.c2[a] : 2.0f * inv);
        K.B[a] = k == K_C4 ? -P.c4b[a] : (k == K_C2 ? 0.0f : -0.5f * inv);
        K.G[a] = k == K_FWD ? -1.5f * inv : (k == K_BWD ? 1.5f * inv : 0.0f);
    }
    return K;
}

// Coefficients of column (gx, gy) on a plane zz at the z faces (zz < 2 or
// zz > nz - 3).  There the region map gives every axis the outer-face rule
// (src/jacobian.c:112-241): one-sided along the axes where the voxel lies on
// the outer face, 2nd-order central along the others (on planes 1 and
// nz - 2 off the x / y faces that is the penultimate layer's 2nd-order
// stencil, the same thing).  The values are column_coefficients' for those
// kinds.
__device__ __forceinline__ ColCoef face_coefficients(const PassParams &P, int gx, int gy, int zz)
{
    const int g[3] = {gx, gy, zz}, n[3] = {P.nx, P.ny, P.nz};
    ColCoef F;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool lo = g[a] == 0, hi = g[a] == n[a] - 1;
        const float inv = P.inv[a];
        F.A[a] = (lo || hi) ? 2.0f * inv : P.c2[a];
        F.B[a] = (lo || hi) ? -0.5f * inv : 0.0f;
        F.G[a] = lo ? -1.5f * inv : (hi ? 1.5f * inv : 0.0f);
    }
    return F;
}

// J - 1 from per-lane stencil coefficients.  A lane with the 4th-order
// stencil rounds as jacobian_delta's interior form (G c = 0 adds an exact
// zero).  ZFACE: the plane is at a z face, where the z stencil may use the
// centre sample; elsewhere it never does (G[2] = 0) and that term is left out.
template <typename CT, bool ZFACE = true>
__device__ __forceinline__ float jacobian_delta_coef(const ColCoef &K, const float (&sm)[3][3][4],
                                                     const float (&v0)[3])
{
    CT d4[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const CT D1 = (CT)(sm[c][a][2] - sm[c][a][1]);
            const CT D2 = (CT)(sm[c][a][3] - sm[c][a][0]);
            const CT g = (a < 2 || ZFACE) ? fma_ct(D2, (CT)K.B[a], (CT)K.G[a] * (CT)v0[c]) : D2 * (CT)K.B[a];
            d4[c][a] = fma_ct(D1, (CT)K.A[a], g);
        }
    return (float)det_plus_identity_minus_one<CT>(d4);
}

// Penalised weight of the adjoint stencil (displacement_field_gradient.c:71-88):
// w = e + zeta (J - delta) where J < delta.
__device__ __forceinline__ float penalised_weight(float e, float J, const PassParams &P)
{
    return J < P.delta ? fma_ct(P.zeta, J - P.delta, e) : e;
}

}  // namespace dt
