// dt_host.hpp -- host-side pieces shared by the single-volume solve and the
// batch engine: error reporting, instance planning (tile shape, z-chunk depth,
// device buffers), the initial control block and the XCD-aware tile table.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "disptools_hip.h"
#include "dt_internal.hpp"
#include "dt_kernels.hpp"

namespace dt {

// Sets the process-global error state (the contract of src/disptools.c:8-43).
void set_error(int line, const char *file, const char *func, const char *msg);

// Serialises the single-volume entry points (they share the device arena).
std::mutex &solve_lock();
// Device scratch of at least `bytes` on the current device (grow-only, reused
// by later calls; valid until the next call under solve_lock()).
int arena_device_buffer(size_t bytes, void **ptr);
// Host-coherent completion word of the current device, cleared (one per
// device, used under solve_lock()): `host` is what the host polls, `dev` what
// the pass kernels store to; `final_*`: where the device publishes the final
// control block of a gradient or greedy solve.
int status_word(unsigned long long **host, unsigned long long **dev, Ctrl **final_host = nullptr,
                Ctrl **final_dev = nullptr);
// Waits until the completion word reports `target` active passes or the end
// of the solve (`done`).  0: reached; 1: the stream drained without the word
// reaching the target (the caller reads the control block); -1: device error.
int follow_status(const unsigned long long *w, long long target, hipStream_t st, bool &done);
// Pinned host memory for completion words and published control blocks.
hipError_t host_words_alloc(void **p, size_t bytes);

// One iteration of a host spin-wait.
inline void cpu_relax()
{
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#else
    std::this_thread::yield();
#endif
}

#define DT_FAIL(msg)                                         \
    do {                                                     \
        ::dt::set_error(__LINE__, __FILE__, __func__, (msg)); \
        return -1;                                           \
    } while (0)

#define DT_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            char m_[1024];                                                            \
            std::snprintf(m_, sizeof m_, "%s: %s", #call, hipGetErrorString(e_));     \
            ::dt::set_error(__LINE__, __FILE__, __func__, m_);                        \
            return -1;                                                                \
        }                                                                             \
    } while (0)

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves 256-byte aligned sub-buffers out of a base pointer.
struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *b) : base((char *)b) {}
    template <typename T>
    T *take(size_t count)
    {
        off = align_up(off, 256);
        T *p = (T *)(base + off);
        off += count * sizeof(T);
        return p;
    }
};

// Tiling of one instance.
struct Plan {
    int algorithm = DT_GRADIENT;
    int precision = 32;      // stencil arithmetic: 32 or 64 bits
    int variant = 0;         // pass variant of the algorithm (column tile shape)
    int nx = 0, ny = 0, nz = 0;
    int tx = 64, ty = 12, tz = 4;
    int ntx = 0, nty = 0, ntz = 0, ntiles = 0;
    long long n = 0;
};

// Named range on the calling thread for rocprofv3 --marker-trace (roctx).
struct MarkerRange {
    explicit MarkerRange(const char *fmt, ...);
    ~MarkerRange();
    MarkerRange(const MarkerRange &) = delete;
    MarkerRange &operator=(const MarkerRange &) = delete;
};

// Per-iteration log lines of the reference entry points (DISPTOOLS_VERBOSE).
bool verbose_from_env();
void print_iteration(long long it, double E, double m, double eta);

// Compute units of the current device.
int device_cus();
// z-chunk depth that minimises (blocks per CU, or rounds of resident blocks)
// x (TZ + overhead).
int choose_tz(int nx, int ny, int nz, int tx, int ty, int cus, int resident, int overhead, bool shared_cu);
// DISPTOOLS_GRAD_VARIANT / DISPTOOLS_GREEDY_VARIANT / DISPTOOLS_STENCIL_PRECISION overrides.
int variant_from_env();
int greedy_variant_from_env();
int precision_from_env();
// DISPTOOLS_STATUS_WORD=0: the kernels publish no completion word; the host
// follows a solve through its fallback (stream drained -> control block read
// from device memory).  For testing that path.
bool status_words_from_env();
// Tiling of an instance; `variant` is the pass variant of the algorithm
// (gradient: 0 .. GRAD_VARIANTS-1, greedy: GREEDY_FUSED or GREEDY_FUSED_WIDE);
// tz <= 0 picks the depth for a launch of this instance alone (choose_tz).
// Depths are multiples of 4 (the groups of planes of the exact sums).
Plan make_plan(int algorithm, int nx, int ny, int nz, int precision, int variant, int tz);

// Device bytes of an instance's solver state besides target, mask and the
// caller's field (extra field buffers, error maps, trace; the two control
// copies and the accumulators too when `own_ctrl`).
size_t scratch_bytes(const Plan &P, long long trace_cap, bool own_ctrl);

// Fills the instance record: geometry, parameters, the caller's buffers and
// the solver state carved from `cv` (same order and sizes as scratch_bytes).
// ctrl (two copies) and acc (ACC_SLOTS * ACC_COPIES) are carved too when null.
Inst make_inst(const Plan &P, float dx, float dy, float dz, const dt_params &p, const float *target,
               const uint8_t *mask, float *field, Carver &cv, double *trace, long long trace_cap, Ctrl *ctrl,
               Acc *acc);

// Control block of a solve that starts from the initial guess in buf[0]
// (both copies start from it; the accumulators start at zero).
Ctrl initial_ctrl(int algorithm, const dt_params &p);

// The newer of an instance's two control copies (the one that records the end
// of the solve, or the one with more evaluated passes).
const Ctrl &latest_ctrl(const Ctrl (&c)[2]);

// The tiles of one instance: `cols` column tiles times the z-chunks of depth
// `tz` over `nz` planes (matching: cols blocks, nz = tz = 1).
struct TileRun {
    int inst, cols, nz, tz;
};
inline TileRun plan_run(int inst, const Plan &P) { return TileRun{inst, P.ntx * P.nty, P.nz, P.tz}; }

// Appends the tiles of `runs` to `out` (per run: columns fastest, then
// z-chunks), in the order that gives block b (dispatched to XCD b % 8) a
// contiguous run of the concatenated tile list, so the halos a block re-reads
// sit in its own XCD's L2.
void append_tiles_xcd(std::vector<Tile> &out, const std::vector<TileRun> &runs);

// Checks the arguments of a solve (sets the error state and returns -1).
int validate(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, const dt_params *p);

// One scale step of the image pyramid on the device (drawing.scale_image):
// the [ncomp][nz][ny][nx] volume `in` (spacing dx, dy, dz) smoothed by the
// recursive Gaussian of `sigma` physical units per axis, then resampled onto
// mx x my x mz (linear, or nearest when `nearest`) into `out`; t0 and t1 are
// scratch volumes of the input's size.  Enqueued on `st`.
int scale_volume(int nx, int ny, int nz, int ncomp, float dx, float dy, float dz, const float *in, int mx, int my,
                 int mz, float sigma, bool nearest, float *out, float *t0, float *t1, hipStream_t st);

// Largest number of passes a solve may take before the host declares it
// stuck (a line search that never terminates, e.g. alpha <= 1 or beta >= 1).
long long max_passes(const dt_params &p);

}  // namespace dt
