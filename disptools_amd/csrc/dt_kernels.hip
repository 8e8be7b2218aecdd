// dt_kernels.hip -- gfx950 kernels of the displacement-generation path.
//
//   k_jacobian     det(I + grad u) over a volume (reference jacobian, src/jacobian.c:70-242)
//   k_grad_pass    one fused pass of the gradient solver: trial field -> Jacobian ->
//                  voxel error -> adjoint gradient -> speculative next candidate,
//                  with FP64 energy sums and the line-search controller in the
//                  last block (src/displacement_field_gradient.c:26-431)
//   k_greedy_step  greedy update of the interior (src/displacement_field_greedy.c:23-67)
//   k_greedy_eval  Jacobian + error of a greedy trial + controller (:85-247)
//   k_regularise   J = max(J, epsilon) (src/jacobian.c:29-55)
//
// All volume kernels work on tiles of TX x TY columns and TZ planes; a launch
// walks a tile table that may hold tiles of several instances (batched solves).
#include <math.h>

#include "dt_device.hpp"
#include "dt_kernels.hpp"

namespace dt {

// ---------------------------------------------------------------------------
// Greedy kernels
// ---------------------------------------------------------------------------

// u_new = u_old + (eta d_c / 2)(e(p + e_c) - e(p - e_c)) on the interior
// (displacement_field_greedy.c:36-63); the outer face of u_new is not written.
template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
k_greedy_step(const Inst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    const Ctrl *C = I.ctrl;
    if (!C->active || !C->step_on) return;
    const TileCoord tc = decode_tile<TX, TY>(I, tl.tile);
    const int gx = tc.x0 + (int)threadIdx.x % TX, gy = tc.y0 + (int)threadIdx.x / TX;
    const int nx = I.nx, ny = I.ny, nz = I.nz;
    if (gx < 1 || gx > nx - 2 || gy < 1 || gy > ny - 2) return;
    const float *__restrict__ uo = I.buf[C->base];
    float *__restrict__ un = I.buf[C->g_new];
    const float *__restrict__ e = I.err[C->e_cur];
    const float eta = C->eta;
    const float ddx = eta * I.dx * .5f, ddy = eta * I.dy * .5f, ddz = eta * I.dz * .5f;
    const long long n = I.n, sy = nx, sz = (long long)nx * ny;
    for (int z = max(tc.z0, 1); z < min(tc.z1, nz - 1); ++z) {
        const long long i = ((long long)z * ny + gy) * nx + gx;
        const float sx_ = ddx * (e[i + 1] - e[i - 1]);
        const float sy_ = ddy * (e[i + sy] - e[i - sy]);
        const float sz_ = ddz * (e[i + sz] - e[i - sz]);
        un[i] = uo[i] + sx_;
        un[n + i] = uo[n + i] + sy_;
        un[2 * n + i] = uo[2 * n + i] + sz_;
    }
}

// Jacobian at one voxel read straight from global memory (region map of src/jacobian.c).
template <typename CT>
__device__ __forceinline__ float jacobian_global(const float *__restrict__ f, long long n, int x, int y, int z,
                                                 int nx, int ny, int nz, float ix, float iy, float iz)
{
    int kind[3];
    classify(x, y, z, nx, ny, nz, kind);
    const long long i = ((long long)z * ny + y) * nx + x;
    const long long st[3] = {1, nx, (long long)nx * ny};
    const float inv[3] = {ix, iy, iz};
    CT d[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *p = f + c * n + i;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const long long s = st[a];
            const int k = kind[a];
            const float v0 = p[0];
            const float m1 = (k == K_C4 || k == K_C2 || k == K_BWD) ? p[-s] : 0.0f;
            const float p1 = (k == K_C4 || k == K_C2 || k == K_FWD) ? p[s] : 0.0f;
            const float m2 = (k == K_C4 || k == K_BWD) ? p[-2 * s] : 0.0f;
            const float p2 = (k == K_C4 || k == K_FWD) ? p[2 * s] : 0.0f;
            d[c][a] = dstencil<CT>(k, m2, m1, v0, p1, p2, inv[a]);
        }
    }
    return (float)det_plus_identity<CT>(d);
}

// Jacobian + error of the greedy trial buf[g_new] (or of the initial guess),
// energy sums, controller.
template <typename CT, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
k_greedy_eval(const Inst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    constexpr int NT = TX * TY;
    __shared__ double red[3 * (NT / 64)];
    __shared__ int flag;
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    Ctrl *C = I.ctrl;
    if (!C->active) return;
    const TileCoord tc = decode_tile<TX, TY>(I, tl.tile);
    const int gx = tc.x0 + (int)threadIdx.x % TX, gy = tc.y0 + (int)threadIdx.x / TX;
    const int nx = I.nx, ny = I.ny, nz = I.nz;
    const float *__restrict__ f = I.buf[C->phase == PH_INIT ? C->base : C->g_new];
    float *__restrict__ e = I.err[C->e_new];
    double Esum = 0.0, dummy = 0.0;
    float mx = 0.0f;
    if (gx < nx && gy < ny) {
        for (int z = tc.z0; z < tc.z1; ++z) {
            const float J = jacobian_global<CT>(f, I.n, gx, gy, z, nx, ny, nz, I.ix, I.iy, I.iz);
            const long long i = ((long long)z * ny + gy) * nx + gx;
            const bool in_mask = I.mask[i] != 0;
            const float ev = voxel_error(J, I.target[i], in_mask, I.p.tolerance);
            e[i] = ev;
            Esum += (double)ev * (double)ev;
            if (in_mask) mx = fmaxf(mx, fabsf(ev));
        }
    }
    block_reduce3<NT>(Esum, mx, dummy, red);
    if (arrive<NT>(I, tl.tile, Esum, mx, 0.0, &flag)) {
        double E, G;
        float m;
        fold_partials<NT>(I, E, m, G, red);
        if (threadIdx.x == 0) {
            greedy_controller(I, C, E, m, 3);
            publish_status(I, C);
            __hip_atomic_store(I.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Standalone Jacobian and regularisation
// ---------------------------------------------------------------------------

template <typename CT>
__global__ void __launch_bounds__(256)
k_jacobian(const float *__restrict__ f, float *__restrict__ J, int nx, int ny, int nz,
           float ix, float iy, float iz)
{
    const long long n = (long long)nx * ny * nz;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= nx || y >= ny) return;
    J[((long long)z * ny + y) * nx + x] = jacobian_global<CT>(f, n, x, y, z, nx, ny, nz, ix, iy, iz);
}

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
    const int rb = __builtin_amdgcn_readfirstlane(I.ctrl->result_buf);
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

hipError_t launch_jacobian(const float *f, float *J, int nx, int ny, int nz, float ix, float iy, float iz,
                           hipStream_t st)
{
    dim3 grid((nx + 63) / 64, (ny + 3) / 4, nz);
    hipLaunchKernelGGL(k_jacobian<double>, grid, dim3(256), 0, st, f, J, nx, ny, nz, ix, iy, iz);
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

hipError_t launch_greedy_step(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL((k_greedy_step<TILE_X, TILE_Y>), dim3(nblocks), dim3(TILE_X * TILE_Y), 0, st,
                       insts, tiles);
    return hipGetLastError();
}

hipError_t launch_greedy_eval(int precision, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    if (precision == 64)
        hipLaunchKernelGGL((k_greedy_eval<double, TILE_X, TILE_Y>), dim3(nblocks), dim3(TILE_X * TILE_Y), 0,
                           st, insts, tiles);
    else
        hipLaunchKernelGGL((k_greedy_eval<float, TILE_X, TILE_Y>), dim3(nblocks), dim3(TILE_X * TILE_Y), 0,
                           st, insts, tiles);
    return hipGetLastError();
}

}  // namespace dt
