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
    float c2[3];      // 0.5 * inv
    float tol, delta, zeta;
};

__device__ __forceinline__ PassParams load_params(const Inst &I)
{
    PassParams P;
    P.nx = uni(I.nx); P.ny = uni(I.ny); P.nz = uni(I.nz);
    P.inv[0] = uni(I.ix); P.inv[1] = uni(I.iy); P.inv[2] = uni(I.iz);
#pragma unroll
    for (int a = 0; a < 3; ++a) P.c2[a] = uni(I.coef[a][K_C2].A);
    P.tol = uni(I.p.tolerance); P.delta = uni(I.p.delta); P.zeta = uni(I.p.zeta);
    return P;
}

// The stencil coefficient table of an instance (Inst::coef), read where a
// step needs it: a lane loads the row of its own stencil kind (one 16-byte
// load per axis), so a wave whose lanes hold different stencils selects
// nothing and branches nowhere.  (The products are formed on the host: a
// float product formed here is a vector instruction, and its wave-uniform
// result would occupy a vector register for the whole pass.)
struct StencilCoef {
    const DT_GLOBAL CoefRow *t;   // [axis][kind]
    __device__ __forceinline__ CoefRow row(int a, int kind) const
    {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v v = *(const DT_GLOBAL f4v *)(t + 4 * a + kind);
        return CoefRow{v.x, v.y, v.z, v.w};
    }
    // the 4th-order central weights, wave-uniform: (2/3) / spacing, (1/12) / spacing
    __device__ __forceinline__ float c4a(int a) const { return uni(t[4 * a + K_C4].A); }
    __device__ __forceinline__ float c4b(int a) const { return -uni(t[4 * a + K_C4].B); }
};

__device__ __forceinline__ StencilCoef load_coef(const Inst &I)
{
    StencilCoef c;
    c.t = gptr(uni(&I.coef[0][0]));
    return c;
}

// J - 1 at voxel (gx, gy, zz) from its samples sm[c][axis][m2, m1, p1, p2] and
// centre v0[c], with the stencil the region map assigns there.  Waves with no
// boundary lane take the 4th-order form; the others a weighted 5-point form
// whose weights encode each lane's stencil (no divergent branches).
template <typename CT>
__device__ __forceinline__ float jacobian_delta(const PassParams &P, const StencilCoef &cf, int gx, int gy, int zz,
                                                const float (&sm)[3][3][4], const float (&v0)[3])
{
    CT d4[3][3];
    const bool fast = gx >= 2 && gx <= P.nx - 3 && gy >= 2 && gy <= P.ny - 3 && zz >= 2 && zz <= P.nz - 3;
    if (__builtin_amdgcn_ballot_w64(!fast) == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a)
                d4[c][a] = (CT)(sm[c][a][2] - sm[c][a][1]) * (CT)cf.c4a(a) -
                           (CT)(sm[c][a][3] - sm[c][a][0]) * (CT)cf.c4b(a);
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

__device__ __forceinline__ ColCoef coef_rows(const StencilCoef &cf, const int (&kind)[3])
{
    ColCoef K;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const CoefRow r = cf.row(a, kind[a]);
        K.A[a] = r.A; K.B[a] = r.B; K.G[a] = r.G;
    }
    return K;
}

// The region map there (classify on such a plane): a column on the outer x
// or y face takes one-sided differences along the axes where it lies on the
// face and 2nd-order central ones along the others; a column in the
// penultimate layer takes 2nd-order central differences; the rest 4th-order.
__device__ __forceinline__ ColCoef column_coefficients(const PassParams &P, const StencilCoef &cf, int gx, int gy)
{
    const bool ox = gx == 0 || gx == P.nx - 1, oy = gy == 0 || gy == P.ny - 1;
    const bool pen = gx == 1 || gx == P.nx - 2 || gy == 1 || gy == P.ny - 2;
    const int central = (ox || oy || pen) ? K_C2 : K_C4;
    const int kind[3] = {gx == 0 ? K_FWD : (gx == P.nx - 1 ? K_BWD : central),
                         gy == 0 ? K_FWD : (gy == P.ny - 1 ? K_BWD : central), central};
    return coef_rows(cf, kind);
}

// Coefficients of column (gx, gy) on a plane zz at the z faces (zz < 2 or
// zz > nz - 3).  There the region map gives every axis the outer-face rule
// (src/jacobian.c:112-241): one-sided along the axes where the voxel lies on
// the outer face, 2nd-order central along the others (on planes 1 and
// nz - 2 off the x / y faces that is the penultimate layer's 2nd-order
// stencil, the same thing).
__device__ __forceinline__ ColCoef face_coefficients(const PassParams &P, const StencilCoef &cf, int gx, int gy,
                                                     int zz)
{
    const int g[3] = {gx, gy, zz}, n[3] = {P.nx, P.ny, P.nz};
    int kind[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) kind[a] = g[a] == 0 ? K_FWD : (g[a] == n[a] - 1 ? K_BWD : K_C2);
    return coef_rows(cf, kind);
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
