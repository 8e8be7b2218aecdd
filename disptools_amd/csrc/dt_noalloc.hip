// dt_noalloc.hip -- the solve entry point that allocates nothing.
//
// dt_solve_device_noalloc is dt_solve_device for callers that cannot allocate
// device or pinned memory (code under graph capture, tight serving loops,
// harnesses that forbid it): the solver's state (extra field buffers, error
// maps, instance record, control copies, accumulators, tile table) lives in
// `scratch`, a file-scope device array of DT_NOALLOC_SCRATCH_BYTES that each
// GPU the library runs on holds until the process ends.  Set-up is a kernel
// (no host staging), the host follows the solve by reading the control
// copies after each chunk of passes (no host-coherent word), and no events
// are created.  Calls on one GPU share its array: the library serialises its
// own calls, and calls from elsewhere that overlap on one GPU would race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "disptools_hip.h"
#include "dt_host.hpp"
#include "dt_internal.hpp"
#include "dt_kernels.hpp"

namespace dt {

static __device__ __align__(256) unsigned char scratch[DT_NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` per device, from hipGetSymbolAddress on the
// first call with that device current (the array's name in host code is a
// host-side stand-in, not a device address).
constexpr int NOALLOC_MAX_DEVICES = 64;
static unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES];

static int scratch_address(unsigned char **out)
{
    int dev = 0;
    DT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) DT_FAIL("device index beyond the no-allocation table");
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        DT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) DT_FAIL("no device address for the scratch array");
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Instance record, control copies and accumulators at the start of scratch;
// the tile table follows.
struct alignas(256) NoAllocHead {
    Inst inst;
    Ctrl ctrl[2];
    Acc acc[ACC_SLOTS * ACC_COPIES];
};

// Writes the instance record, both control copies, zero accumulators and the
// tile table in the XCD order of append_tiles_xcd (block b on XCD b % 8 walks
// a contiguous run of tiles), and clears `clear` (the greedy solver's second
// field buffer) -- all on the device.
__global__ void __launch_bounds__(256)
k_noalloc_setup(NoAllocHead *head, Tile *tiles, Inst h, Ctrl c0, int ntiles, float *clear, long long nclear)
{
    const long long g = blockIdx.x * 256ll + threadIdx.x, step = (long long)gridDim.x * 256;
    if (g == 0) {
        head->inst = h;
        head->ctrl[0] = c0;
        head->ctrl[1] = c0;
    }
    unsigned long long *acc = (unsigned long long *)head->acc;
    for (long long i = g; i < (long long)(sizeof(head->acc) / 8); i += step) acc[i] = 0ull;
    // Tile t = x * per + k runs on XCD x (k-th of its run); the table lists
    // the valid (k, x) pairs in (k, x) order: its position is the number of
    // valid pairs before it.
    constexpr int XCDS = 8;
    const int per = (ntiles + XCDS - 1) / XCDS;
    const int cols = h.ntx * h.nty;
    for (long long t = g; t < (long long)ntiles; t += step) {
        const int x = (int)(t / per), k = (int)(t % per);
        long long pos = 0;
#pragma unroll
        for (int xx = 0; xx < XCDS; ++xx) {
            const int cnt = min(max(ntiles - xx * per, 0), per);   // tiles of XCD xx
            pos += min(cnt, k) + ((xx < x && cnt > k) ? 1 : 0);
        }
        const int z0 = (int)(t / cols) * h.tz;
        tiles[pos] = Tile{0, (int)(t % cols), z0, min(z0 + h.tz, h.nz)};
    }
    for (long long i = g; i < nclear; i += step) clear[i] = 0.0f;
}

}  // namespace dt

using namespace dt;

extern "C" int dt_solve_device_noalloc(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy,
                                       float dz, const float *d_target, const unsigned char *d_mask,
                                       const dt_params *params, float *d_field, dt_solve_info *info,
                                       void *stream)
{
    std::lock_guard<std::mutex> g(solve_lock());
    disptools_clear_error();
    if (validate(algorithm, nx, ny, nz, dx, dy, dz, params)) return -1;
    if (!d_target || !d_mask || !d_field) DT_FAIL("null buffer");
    hipStream_t st = (hipStream_t)stream;
    const int precision = precision_from_env();
    const int variant = algorithm == DT_GRADIENT ? variant_from_env() : GREEDY_FUSED;
    const Plan P = make_plan(algorithm, (int)nx, (int)ny, (int)nz, precision, variant, 0);
    const size_t head = align_up(sizeof(NoAllocHead), 256);
    const size_t tiles_bytes = align_up((size_t)P.ntiles * sizeof(Tile), 256);
    const size_t state = scratch_bytes(P, 0, false);
    const size_t need = head + tiles_bytes + state;
    if (need > (size_t)DT_NOALLOC_SCRATCH_BYTES) {
        char m[256];
        std::snprintf(m, sizeof m,
                      "volume too large for the no-allocation entry point: needs %zu bytes of scratch, at most %zu",
                      need, (size_t)DT_NOALLOC_SCRATCH_BYTES);
        DT_FAIL(m);
    }
    unsigned char *base = nullptr;
    if (scratch_address(&base)) return -1;
    NoAllocHead *d_head = (NoAllocHead *)base;
    Tile *d_tiles = (Tile *)(base + head);
    Carver cv(base + head + tiles_bytes);
    Inst h = make_inst(P, dx, dy, dz, *params, d_target, d_mask, d_field, cv, nullptr, 0, d_head->ctrl,
                       d_head->acc);
    h.status = nullptr;
    h.final_ctrl = nullptr;
    const long long n = P.n;
    float *clear = algorithm == DT_GREEDY ? h.buf[1] : nullptr;
    const long long nclear = algorithm == DT_GREEDY ? 3 * n : 0;
    long long work = std::max<long long>((long long)P.ntiles, nclear);
    long long blocks = std::min<long long>(std::max<long long>((work + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_noalloc_setup, dim3((unsigned)blocks), dim3(256), 0, st, d_head, d_tiles, h,
                       initial_ctrl(algorithm, *params), P.ntiles, clear, nclear);
    DT_HIP(hipGetLastError());

    // chunks of passes; after each, the control copies tell whether the solve
    // has ended (the controller step of a pass's trial runs at the start of
    // the next pass, so every chunk is followed by more until the end shows)
    const long long pass_cap = max_passes(*params);
    constexpr int CHUNK = 4;
    long long launched = 0;
    Ctrl c[2];
    for (;;) {
        for (int i = 0; i < CHUNK; ++i, ++launched) {
            const int parity = (int)(launched & 1);
            if (algorithm == DT_GRADIENT) {
                DT_HIP(launch_grad_pass(precision, P.variant, &d_head->inst, d_tiles, P.ntiles, parity, st));
            } else {
                DT_HIP(launch_greedy_pass(precision, P.variant, &d_head->inst, d_tiles, P.ntiles, parity, st));
            }
        }
        DT_HIP(hipMemcpyAsync(c, d_head->ctrl, sizeof c, hipMemcpyDeviceToHost, st));
        DT_HIP(hipStreamSynchronize(st));
        if (latest_ctrl(c).phase == PH_DONE) break;
        if (launched > pass_cap) DT_FAIL("solver did not terminate (check alpha > 1 and beta < 1)");
    }
    DT_HIP(launch_take_result(&d_head->inst, n, st));
    DT_HIP(hipStreamSynchronize(st));
    const Ctrl &fin = latest_ctrl(c);
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->iterations = fin.it;
        info->evaluations = fin.evaluations;
        info->rejections = fin.rejections;
        info->launches = launched;
        info->reason = fin.reason;
        info->final_error = fin.E_result;
        info->final_max_error = fin.m_result;
        info->trace_len = fin.trace_len;
        info->field_io = fin.field_io;
    }
    return 0;
}
