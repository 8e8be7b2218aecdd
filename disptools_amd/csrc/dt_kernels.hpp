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
// precision: 32 = FP32 stencil arithmetic, 64 = FP64 stencil arithmetic (reference-exact)
hipError_t launch_grad_pass(int precision, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st);
hipError_t launch_greedy_step(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st);
hipError_t launch_greedy_eval(int precision, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st);

}  // namespace dt
