// dt_kernels.hpp -- launchers of the gfx950 kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "dt_internal.hpp"

namespace dt {

hipError_t launch_jacobian(const float *f, float *J, int nx, int ny, int nz, float ix, float iy, float iz,
                           hipStream_t st);
hipError_t launch_regularise(float *J, long long n, float eps, hipStream_t st);

// Variants of the gradient pass (dt_grad_stream.hip): column tile TX x TY,
// one thread per column and one per point of the ring around the tile.
// The column tile changes the speed only: the sums are exact, so every
// variant gives the same results.
constexpr int GRAD_VARIANTS = 5;
constexpr int DEFAULT_GRAD_VARIANT = 4;   // 64 x 12
constexpr int ENGINE_GRAD_VARIANT = 3;    // 32 x 24 (engine lanes)
inline void grad_variant_dims(int v, int &tx, int &ty)
{
    switch (v) {
    case 0: tx = 32; ty = 16; break;
    case 1: tx = 32; ty = 20; break;
    case 2: tx = 64; ty = 8; break;
    case 3: tx = 32; ty = 24; break;
    default: tx = 64; ty = 12; break;
    }
}
// precision: 32 = FP32 stencil arithmetic, 64 = FP64 stencil arithmetic.
// parity: the pass's index mod 2 (which control copy it reads).
hipError_t launch_grad_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                            int parity, hipStream_t st);
// resident blocks per CU of a variant (from the compiled resource usage)
int grad_pass_blocks_per_cu(int precision, int variant);
// Coefficients of the recursive Gaussian along one axis (dt_pyramid.hip):
// causal numerator n[0..3], anti-causal numerator m[0..3] (z^1 .. z^4),
// denominator 1 + d[0..3] z^-1 .. z^-4, the steady states per unit input
// (nsum = sum n / (1 + sum d), msum likewise) and the DC normalisation.
struct GaussCoef {
    double n[4], m[4], d[4];
    double nsum, msum, scale;
};
// Deriche's fourth-order recursive Gaussian of standard deviation `sigma`
// samples (sigma > 0).
GaussCoef gauss_coefficients(double sigma);
hipError_t launch_gauss_axis(const float *in, float *out, int nx, int ny, int nz, int ncomp, int axis,
                             const GaussCoef &g, hipStream_t st);
hipError_t launch_nonzero_u8(const float *in, unsigned char *out, long long n, hipStream_t st);
// to_xyz3: [3][n] -> [n][3]; else [n][3] -> [3][n]
hipError_t launch_interleave(const float *in, float *out, long long n, bool to_xyz3, hipStream_t st);
hipError_t launch_resample(const float *in, float *out, int nx, int ny, int nz, int mx, int my, int mz, int ncomp,
                           float rx, float ry, float rz, int nearest, hipStream_t st);
// Greedy pass variants (dt_greedy_stream.hip): GREEDY_FUSED = 32 x 16
// column tiles, GREEDY_FUSED_WIDE = 32 x 32.
constexpr int GREEDY_FUSED = 1;
constexpr int GREEDY_FUSED_WIDE = 2;
inline void greedy_variant_dims(int v, int &tx, int &ty)
{
    tx = 32;
    ty = v == GREEDY_FUSED_WIDE ? 32 : 16;
}
hipError_t launch_greedy_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                              int parity, hipStream_t st);
int greedy_pass_blocks_per_cu(int precision, int variant);
// Copies buf[ctrl->result_buf] of the instance into buf[0] (nothing when the
// result already sits there); n voxels.
hipError_t launch_take_result(const Inst *inst, long long n, hipStream_t st);

}  // namespace dt
