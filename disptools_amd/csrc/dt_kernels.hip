// dt_kernels.hip -- gfx950 kernels of the displacement-generation path
// besides the solver passes (dt_grad_stream.hip, dt_greedy_stream.hip) and
// the standalone Jacobian (dt_jacobian.hip):
//
//   k_regularise   J = max(J, epsilon) (src/jacobian.c:29-55)
//   k_take_result  the solve's result into the caller's field buffer
#include <math.h>

#include "dt_device.hpp"
#include "dt_kernels.hpp"

namespace dt {

// ---------------------------------------------------------------------------
// Regularisation and result copy
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
k_regularise(float *__restrict__ J, long long n, float eps)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        if (J[i] < eps) J[i] = eps;
}

// The solve's result into the caller's field (buf[0]) when the controller
// left it in another buffer (src/displacement_field_gradient.c:421 returns
// field_[old]).
__global__ void __launch_bounds__(256)
k_take_result(const Inst *__restrict__ inst)
{
    const Inst &I = *inst;
    // the control copy that records the end of the solve (both, once a pass
    // after the end has run; they then agree)
    const Ctrl *c = I.ctrl[1].phase == PH_DONE && I.ctrl[0].phase != PH_DONE ? I.ctrl + 1 : I.ctrl;
    const int rb = __builtin_amdgcn_readfirstlane(c->result_buf);
    if (rb <= 0 || rb >= NBUF) return;
    const float *__restrict__ src = I.buf[rb];
    float *__restrict__ dst = I.buf[0];
    const long long n3 = 3 * I.n;
    const long long step = (long long)gridDim.x * 256;
    const long long g = blockIdx.x * 256ll + threadIdx.x;
    long long done = 0;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
        const long long q = n3 >> 2;
        for (long long i = g; i < q; i += step) ((float4 *)dst)[i] = ((const float4 *)src)[i];
        done = q << 2;
    }
    for (long long i = done + g; i < n3; i += step) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

hipError_t launch_take_result(const Inst *inst, long long n, hipStream_t st)
{
    long long blocks = (3 * n / 4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_take_result, dim3((unsigned)blocks), dim3(256), 0, st, inst);
    return hipGetLastError();
}

hipError_t launch_regularise(float *J, long long n, float eps, hipStream_t st)
{
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_regularise, dim3((unsigned)blocks), dim3(256), 0, st, J, n, eps);
    return hipGetLastError();
}

// Gradient pass variants, one object per (precision, tile) (dt_grad_stream.hip);
// variant k is entry k of this list.
#define DT_STREAM_VARIANTS(X) X(0, 32x16_w3) X(1, 32x20_w3) X(2, 64x8_w3) X(3, 32x24_w3) X(4, 64x12_w3)
#define DT_DECLARE_STREAM(k, name)                                                                       \
    hipError_t stream_launch_f32_##name(const Inst *insts, const Tile *tiles, int nblocks, int parity,  \
                                        hipStream_t st);                                                \
    hipError_t stream_launch_f64_##name(const Inst *insts, const Tile *tiles, int nblocks, int parity,  \
                                        hipStream_t st);                                                \
    int stream_occupancy_f32_##name();                                                                   \
    int stream_occupancy_f64_##name();
DT_STREAM_VARIANTS(DT_DECLARE_STREAM)

hipError_t launch_grad_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                            int parity, hipStream_t st)
{
    const bool d = precision == 64;
#define DT_LAUNCH_CASE(k, name)                                                            \
    case k:                                                                                \
        return d ? stream_launch_f64_##name(insts, tiles, nblocks, parity, st)             \
                 : stream_launch_f32_##name(insts, tiles, nblocks, parity, st);
    switch (variant) {
        DT_STREAM_VARIANTS(DT_LAUNCH_CASE)
    default: return hipErrorInvalidValue;
    }
#undef DT_LAUNCH_CASE
}

int grad_pass_blocks_per_cu(int precision, int variant)
{
    const bool d = precision == 64;
#define DT_OCC_CASE(k, name) \
    case k: return d ? stream_occupancy_f64_##name() : stream_occupancy_f32_##name();
    switch (variant) {
        DT_STREAM_VARIANTS(DT_OCC_CASE)
    default: return 1;
    }
#undef DT_OCC_CASE
}

}  // namespace dt
