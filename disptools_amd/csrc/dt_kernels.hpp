// dt_kernels.hpp -- launchers of the gfx950 kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "dt_internal.hpp"

namespace dt {

// Column tile of the volume kernels (one thread per column, 4 wave64 per block).
constexpr int TILE_X = 32;
constexpr int TILE_Y = 8;

hipError_t launch_jacobian(const float *f, float *J, int nx, int ny, int nz, float ix, float iy, float iz,
                           hipStream_t st);
hipError_t launch_regularise(float *J, long long n, float eps, hipStream_t st);
// Variants of the fused gradient pass: column tile TX x TY (one thread per
// column) and the waves per SIMD its registers are capped for (3: may spill a
// few registers; 2: no spill, lower occupancy; 4: spills more).  Variants
// 0-7 are the two-barrier form (dt_grad_pass.hip) kept for comparison.
// Variants SPLIT_BASE .. SPLIT_BASE+2 run each pass as two kernels
// (dt_split_pass.hip); variants from STREAM_BASE on are the single-barrier
// streaming form (dt_grad_stream.hip).
constexpr int GRAD_VARIANTS = 16;
constexpr int SPLIT_BASE = 8;
constexpr int STREAM_BASE = 11;
inline void grad_variant_dims(int v, int &tx, int &ty)
{
    switch (v) {
    case STREAM_BASE: tx = 32; ty = 16; break;
    case STREAM_BASE + 1: tx = 32; ty = 20; break;
    case STREAM_BASE + 2: tx = 64; ty = 8; break;
    case STREAM_BASE + 3: tx = 32; ty = 24; break;
    case STREAM_BASE + 4: tx = 64; ty = 12; break;
    case SPLIT_BASE: tx = 32; ty = 16; break;
    case SPLIT_BASE + 1: tx = 32; ty = 8; break;
    case SPLIT_BASE + 2: tx = 64; ty = 8; break;
    case 1: case 5: case 6: tx = 32; ty = 16; break;
    case 2: tx = 64; ty = 4; break;
    case 3: tx = 64; ty = 8; break;
    default: tx = 32; ty = 8; break;
    }
}
// precision: 32 = FP32 stencil arithmetic, 64 = FP64 stencil arithmetic (reference-exact)
hipError_t launch_grad_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                            hipStream_t st);
// resident blocks per CU of a variant (from the compiled resource usage)
int grad_pass_blocks_per_cu(int precision, int variant);
hipError_t launch_split_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                             hipStream_t st);
int split_pass_blocks_per_cu(int precision, int variant);
hipError_t launch_stream_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                              hipStream_t st);
int stream_pass_blocks_per_cu(int precision, int variant);
hipError_t launch_smooth_axis(const float *in, float *out, int nx, int ny, int nz, int ncomp, int axis,
                              const float *d_w, int radius, hipStream_t st);
hipError_t launch_resample(const float *in, float *out, int nx, int ny, int nz, int mx, int my, int mz, int ncomp,
                           float rx, float ry, float rz, int nearest, hipStream_t st);
// Greedy pass variants: 0 = step kernel + evaluation kernel on TILE_X x TILE_Y
// tiles (dt_kernels.hip); GREEDY_FUSED = one streaming kernel on 32 x 16
// tiles (dt_greedy_stream.hip).
// GREEDY_FUSED_WIDE: the fused kernel on 32 x 32 tiles (16 waves per block).
constexpr int GREEDY_FUSED = 1;
constexpr int GREEDY_FUSED_WIDE = 2;
inline bool greedy_fused(int v) { return v == GREEDY_FUSED || v == GREEDY_FUSED_WIDE; }
inline void greedy_variant_dims(int v, int &tx, int &ty)
{
    tx = v == GREEDY_FUSED ? 32 : (v == GREEDY_FUSED_WIDE ? 32 : TILE_X);
    ty = v == GREEDY_FUSED ? 16 : (v == GREEDY_FUSED_WIDE ? 32 : TILE_Y);
}
hipError_t launch_greedy_step(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st);
hipError_t launch_greedy_eval(int precision, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st);
hipError_t launch_greedy_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                              hipStream_t st);
int greedy_pass_blocks_per_cu(int precision, int variant);
// Copies buf[ctrl->result_buf] of the instance into buf[0] (nothing when the
// result already sits there); n voxels.
hipError_t launch_take_result(const Inst *inst, long long n, hipStream_t st);

}  // namespace dt
