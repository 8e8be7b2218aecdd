// dt_runtime.cpp -- host runtime and C ABI of the HIP displacement library.
//
// A solve keeps all of its state on the device.  The host enqueues chunks of
// pass kernels (two chunks in flight) and follows the completion word the
// pass controller stores in host-coherent memory, so the line search and the
// stopping tests never wait for the host and nothing but pass kernels enters
// the stream; passes launched after the solve has stopped return at once.  Reference behaviour: src/displacement_field_gradient.c:191-431,
// src/displacement_field_greedy.c:85-247; the host-driven CUDA loop this
// replaces is cuda/displacement_field_gradient.cu:164-468.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdarg>
#include <cstddef>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "disptools_hip.h"
#include "dt_host.hpp"
#include "dt_internal.hpp"
#include "dt_kernels.hpp"
#include "dt_stencil.hpp"

// ---------------------------------------------------------------------------
// Error state (same contract as src/disptools.c:8-43)
// ---------------------------------------------------------------------------

static disptools_error_state g_error = {false, -1, "", "", "", ""};

void dt::set_error(int line, const char *file, const char *func, const char *msg)
{
    g_error.error = true;
    g_error.line = line;
    std::snprintf(g_error.file, sizeof g_error.file, "%s", file);
    std::snprintf(g_error.function, sizeof g_error.function, "%s", func);
    std::snprintf(g_error.message, sizeof g_error.message, "%s", msg);
    g_error.trace[0] = '\0';
}

extern "C" disptools_error_state *disptools_get_error(void) { return &g_error; }
extern "C" bool disptools_has_error(void) { return g_error.error; }
extern "C" void disptools_clear_error(void)
{
    g_error.error = false;
    g_error.line = -1;
    g_error.file[0] = g_error.function[0] = g_error.message[0] = g_error.trace[0] = '\0';
}
extern "C" int get_float_type_size(void) { return 32; }

extern "C" bool set_device(const int id)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return false;
    if (id < 0 || id >= count) return false;
    return hipSetDevice(id) == hipSuccess;
}

namespace dt {

// ---------------------------------------------------------------------------
// Device arena: one grow-only allocation per device, reused across calls
// ---------------------------------------------------------------------------

struct Arena {
    void *ptr = nullptr;
    size_t bytes = 0;
    char *coherent = nullptr;       // host-coherent: completion word, final control block
    char *coherent_dev = nullptr;   // its device address
    char *stage = nullptr;                      // pinned staging of a solve's head
    size_t stage_bytes = 0;
    hipEvent_t kev[2] = {nullptr, nullptr};     // timing events of a solve's passes
};

// Instance record, control copies and accumulators of a single-volume solve;
// its tile table follows in the same allocation, so one copy stages all of
// them.
struct alignas(256) SolveHead {
    Inst inst;
    Ctrl ctrl[2];
    Acc acc[ACC_SLOTS * ACC_COPIES];
};

static int arena_stage(Arena &a, size_t bytes)
{
    if (a.stage_bytes >= bytes) return 0;
    if (a.stage) DT_HIP(hipHostFree(a.stage));
    a.stage = nullptr;
    a.stage_bytes = 0;
    const size_t want = align_up(bytes, 4096);
    DT_HIP(hipHostMalloc((void **)&a.stage, want, hipHostMallocDefault));
    a.stage_bytes = want;
    return 0;
}

static std::mutex g_lock;
static std::vector<Arena> g_arenas;

hipError_t host_words_alloc(void **p, size_t bytes)
{
    // host-coherent where the platform offers it; otherwise plain pinned
    // memory, where the device's stores may only land at the end of a kernel
    // (the host then follows a solve through its drained-stream fallback)
    if (hipHostMalloc(p, bytes, hipHostMallocCoherent) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

static int arena_get(size_t bytes, Arena **out)
{
    int dev = 0;
    DT_HIP(hipGetDevice(&dev));
    if ((int)g_arenas.size() <= dev) g_arenas.resize(dev + 1);
    Arena &a = g_arenas[dev];
    if (!a.kev[0]) {
        DT_HIP(hipEventCreate(&a.kev[0]));
        DT_HIP(hipEventCreate(&a.kev[1]));
    }
    if (!a.coherent) {
        DT_HIP(host_words_alloc((void **)&a.coherent, 256 + sizeof(Ctrl)));
        DT_HIP(hipHostGetDevicePointer((void **)&a.coherent_dev, a.coherent, 0));
    }
    if (a.bytes < bytes) {
        if (a.ptr) DT_HIP(hipFree(a.ptr));
        a.ptr = nullptr;
        a.bytes = 0;
        const size_t want = align_up(bytes, (size_t)1 << 21);
        if (hipMalloc(&a.ptr, want) != hipSuccess) {
            (void)hipGetLastError();
            a.ptr = nullptr;
            DT_FAIL("cannot allocate device memory for the solver");
        }
        a.bytes = want;
    }
    *out = &a;
    return 0;
}

std::mutex &solve_lock() { return g_lock; }

int arena_device_buffer(size_t bytes, void **ptr)
{
    Arena *A = nullptr;
    if (arena_get(bytes, &A)) return -1;
    *ptr = A->ptr;
    return 0;
}

int status_word(unsigned long long **host, unsigned long long **dev, Ctrl **final_host, Ctrl **final_dev)
{
    Arena *A = nullptr;
    if (arena_get(0, &A)) return -1;
    *(volatile unsigned long long *)A->coherent = 0;
    *host = (unsigned long long *)A->coherent;
    *dev = (unsigned long long *)A->coherent_dev;
    if (final_host) *final_host = (Ctrl *)(A->coherent + 256);
    if (final_dev) *final_dev = (Ctrl *)(A->coherent_dev + 256);
    return 0;
}

int follow_status(const unsigned long long *w, long long target, hipStream_t st, bool &done)
{
    const volatile unsigned long long *vw = w;
    auto reached = [&](unsigned long long v) {
        if (v & 1ull) done = true;
        return done || (long long)(v >> 1) >= target;
    };
    auto last = std::chrono::steady_clock::now();
    for (unsigned spin = 1;; ++spin) {
        if (reached(*vw)) return 0;
        // every few milliseconds: has the device failed, or has the stream
        // drained without the word reaching the target?
        if ((spin & 1023u) == 0) {
            const auto now = std::chrono::steady_clock::now();
            if (now - last > std::chrono::milliseconds(2)) {
                last = now;
                const hipError_t q = hipStreamQuery(st);
                if (q == hipSuccess) return reached(*vw) ? 0 : 1;
                if (q != hipErrorNotReady) {
                    (void)hipGetLastError();
                    return -1;
                }
            }
        }
        // spin while a chunk is short; long waits give the core away
        if (spin < 65536) cpu_relax();
        else std::this_thread::yield();
    }
}

// ---------------------------------------------------------------------------
// Instance planning (shared with the batch engine)
// ---------------------------------------------------------------------------

size_t scratch_bytes(const Plan &P, long long trace_cap, bool own_ctrl)
{
    size_t b = 0;
    auto add = [&](size_t s) { b = align_up(b, 256) + s; };
    const long long n = P.n;
    const int extra_fields = P.algorithm == DT_GRADIENT ? NBUF - 1 : 1;
    for (int i = 0; i < extra_fields; ++i) add(3 * (size_t)n * sizeof(float));
    if (P.algorithm == DT_GREEDY) { add((size_t)n * sizeof(float)); add((size_t)n * sizeof(float)); }
    if (trace_cap > 0) add((size_t)trace_cap * 4 * sizeof(double));
    if (own_ctrl) { add(ACC_SLOTS * ACC_COPIES * sizeof(Acc)); add(2 * sizeof(Ctrl)); }
    return align_up(b, 256);
}

const Ctrl &latest_ctrl(const Ctrl (&c)[2])
{
    if (c[1].phase == PH_DONE && c[0].phase != PH_DONE) return c[1];
    if (c[0].phase == PH_DONE) return c[0];
    return c[1].evaluations > c[0].evaluations ? c[1] : c[0];
}

Inst make_inst(const Plan &P, float dx, float dy, float dz, const dt_params &p, const float *target,
               const uint8_t *mask, float *field, Carver &cv, double *trace, long long trace_cap, Ctrl *ctrl,
               Acc *acc)
{
    Inst h{};
    h.nx = P.nx; h.ny = P.ny; h.nz = P.nz;
    h.ntx = P.ntx; h.nty = P.nty; h.ntz = P.ntz; h.tz = P.tz;
    h.variant = P.variant;
    h.n = P.n;
    h.ix = (float)(1.0 / (double)dx);
    h.iy = (float)(1.0 / (double)dy);
    h.iz = (float)(1.0 / (double)dz);
    h.dx = dx; h.dy = dy; h.dz = dz;
    const float inv[3] = {h.ix, h.iy, h.iz};
    for (int a = 0; a < 3; ++a) {
        const float c4a = (float)(2.0 / 3.0) * inv[a], c4b = (float)(1.0 / 12.0) * inv[a];
        const float c2 = 0.5f * inv[a], c20 = 2.0f * inv[a], c15 = 1.5f * inv[a];
        h.coef[a][K_C4] = CoefRow{c4a, -c4b, 0.0f, 0.0f};
        h.coef[a][K_C2] = CoefRow{c2, 0.0f, 0.0f, 0.0f};
        h.coef[a][K_FWD] = CoefRow{c20, -c2, -c15, 0.0f};
        h.coef[a][K_BWD] = CoefRow{c20, -c2, c15, 0.0f};
    }
    h.p = p;
    h.target = target;
    h.mask = mask;
    h.buf[0] = field;
    const long long n = P.n;
    const int extra_fields = P.algorithm == DT_GRADIENT ? NBUF - 1 : 1;
    for (int i = 1; i <= extra_fields; ++i) h.buf[i] = cv.take<float>(3 * (size_t)n);
    for (int i = extra_fields + 1; i < NBUF; ++i) h.buf[i] = nullptr;
    if (P.algorithm == DT_GREEDY) {
        h.err[0] = cv.take<float>((size_t)n);
        h.err[1] = cv.take<float>((size_t)n);
    }
    if (trace_cap > 0) {
        double *t = cv.take<double>((size_t)trace_cap * 4);
        h.trace = trace ? t : nullptr;
        h.trace_cap = trace ? trace_cap : 0;
    } else {
        h.trace = nullptr;
        h.trace_cap = 0;
    }
    h.acc = acc ? acc : cv.take<Acc>(ACC_SLOTS * ACC_COPIES);
    h.ctrl = ctrl ? ctrl : cv.take<Ctrl>(2);
    return h;
}

Ctrl initial_ctrl(int algorithm, const dt_params &p)
{
    Ctrl c{};
    c.active = 1;
    c.phase = PH_INIT;
    c.mode = TRIAL_DIRECT;
    c.src = 0;
    c.cand = 0;
    c.base = 0;
    c.eta = p.eta;
    c.out_trial = -1;
    if (algorithm == DT_GRADIENT) {
        c.out_next = 1;
        c.eta_next = p.eta * p.alpha;   // first search step (displacement_field_gradient.c:290-291)
    } else {
        c.out_next = -1;
        c.step_on = 0;
        c.g_new = 1;
        c.e_cur = 1;
        c.e_new = 0;
    }
    c.result_buf = 0;
    // the first pass adds into slot 0 and clears slot 1 (already clear)
    c.acc_cur = 0;
    c.acc_zero = 1;
    c.acc_prev = 2;
    c.evaluated = 0;
    c.copy_now = 0;
    c.result_copied = 0;
    return c;
}

// Per-device values that do not change while the process runs (queried once:
// hipGetDeviceProperties and the occupancy calculator cost tens of
// microseconds, which a small solve would pay on every call).
static std::mutex g_query_lock;
static std::map<std::tuple<int, int, int, int>, int> g_queries;   // (device, kind, a, b) -> value

template <typename F>
static int cached_query(int kind, int a, int b, F query)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return query();
    const auto key = std::make_tuple(dev, kind, a, b);
    {
        std::lock_guard<std::mutex> g(g_query_lock);
        auto it = g_queries.find(key);
        if (it != g_queries.end()) return it->second;
    }
    const int v = query();
    std::lock_guard<std::mutex> g(g_query_lock);
    g_queries[key] = v;
    return v;
}

int device_cus()
{
    return cached_query(0, 0, 0, [] {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
                cus = v;
        }
        return cus;
    });
}

// A block streams TZ planes plus a pipeline of `overhead` plane steps.
// Streaming passes (`shared_cu`): blocks that share a CU share its issue slots
// and bandwidth, so a launch takes about ceil(blocks / CUs) block times and
// deep chunks pay off by cutting the halo planes (measured: the fused greedy
// pass at 128^3 runs 8 % faster with 16-plane chunks, one block per CU, than
// with 8-plane chunks, two per CU).  Other kernels: rounds of resident blocks.
int choose_tz(int nx, int ny, int nz, int tx, int ty, int cus, int resident, int overhead, bool shared_cu)
{
    const long long cols = (long long)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
    int best = std::max(1, std::min(nz, 16));
    double best_cost = 1e300;
    // multiples of 4: z-chunks start on the groups of planes of the exact sums
    for (int tz = 4; tz <= std::max(4, nz + 3); tz += 4) {
        const long long blocks = cols * ((nz + tz - 1) / tz);
        const long long per = shared_cu ? (blocks + cus - 1) / cus : (blocks + resident - 1) / resident;
        const double cost = (double)per * (std::min(tz, nz) + overhead);
        if (cost < best_cost - 1e-9) { best_cost = cost; best = std::min(tz, nz); }
    }
    return best;
}

// Default pass variant (DEFAULT_GRAD_VARIANT): 64 x 12 column tiles with the
// tile ring in waves of its own (measured fastest at 128^3 among the compiled
// variants).  Single solves start from it and switch tile by
// model_grad_variant; engine lanes use ENGINE_GRAD_VARIANT;
// DISPTOOLS_GRAD_VARIANT overrides both.
int variant_from_env()
{
    const char *s = std::getenv("DISPTOOLS_GRAD_VARIANT");
    if (!s) return DEFAULT_GRAD_VARIANT;
    const int v = std::atoi(s);
    return (v >= 0 && v < GRAD_VARIANTS) ? v : DEFAULT_GRAD_VARIANT;
}

// Greedy pass variant from DISPTOOLS_GREEDY_VARIANT: 1 = 32 x 16 column
// tiles, 2 = 32 x 32; -1 when unset.
int greedy_variant_from_env()
{
    const char *s = std::getenv("DISPTOOLS_GREEDY_VARIANT");
    if (s && std::strcmp(s, "1") == 0) return GREEDY_FUSED;
    if (s && std::strcmp(s, "2") == 0) return GREEDY_FUSED_WIDE;
    return -1;
}

// Fused greedy tile of a single solve: 32 x 32 tiles (16 waves, one block
// per CU) on wide planes, 32 x 16 on small ones (measured per pass: 64^3
// 19.3 vs 22.6 us, 96^3 25.8 vs 25.6, 128^3 42.1 vs 36.9, 200^3 137 vs 116);
// the FP64-stencil build keeps 32 x 16 (the wide tile spills there).
static int greedy_solve_variant(int nx, int ny, int precision)
{
    const int v = greedy_variant_from_env();
    if (v >= 0) return v;
    return (precision == 32 && (long long)nx * ny >= 8192) ? GREEDY_FUSED_WIDE : GREEDY_FUSED;
}

// Variants the autotuner times: every column tile of the gradient pass.
static const int TUNE_CANDIDATES[] = {0, 1, 2, 3, 4};

struct TuneKey {
    int dev, nx, ny, nz;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(dev, nx, ny, nz) < std::tie(o.dev, o.nx, o.ny, o.nz);
    }
};
static std::map<TuneKey, int> g_tuned;

int precision_from_env()
{
    const char *s = std::getenv("DISPTOOLS_STENCIL_PRECISION");
    if (s && std::strcmp(s, "64") == 0) return 64;
    return 32;
}

bool status_words_from_env()
{
    const char *s = std::getenv("DISPTOOLS_STATUS_WORD");
    return !(s && std::strcmp(s, "0") == 0);
}

static int plan_tz(const Plan &P)
{
    const int cus = device_cus();
    const bool fused_greedy = P.algorithm == DT_GREEDY;
    const bool stream = true;
    const int per_cu = cached_query(1 + P.algorithm, P.precision, P.variant, [&] {
        return P.algorithm == DT_GRADIENT ? grad_pass_blocks_per_cu(P.precision, P.variant)
                                          : greedy_pass_blocks_per_cu(P.precision, P.variant);
    });
    return choose_tz(P.nx, P.ny, P.nz, P.tx, P.ty, cus, cus * per_cu, fused_greedy ? 4 : 3, stream);
}

Plan make_plan(int algorithm, int nx, int ny, int nz, int precision, int variant, int tz)
{
    Plan P;
    P.algorithm = algorithm;
    P.precision = precision;
    P.variant = algorithm == DT_GRADIENT ? variant : (variant == GREEDY_FUSED_WIDE ? GREEDY_FUSED_WIDE : GREEDY_FUSED);
    P.nx = nx; P.ny = ny; P.nz = nz;
    P.n = (long long)nx * ny * nz;
    if (algorithm == DT_GRADIENT) grad_variant_dims(P.variant, P.tx, P.ty);
    else greedy_variant_dims(P.variant, P.tx, P.ty);
    // a given depth is rounded up to a multiple of 4
    P.tz = tz > 0 ? std::min((tz + 3) / 4 * 4, nz) : plan_tz(P);
    P.ntx = (nx + P.tx - 1) / P.tx;
    P.nty = (ny + P.ty - 1) / P.ty;
    P.ntz = (nz + P.tz - 1) / P.tz;
    P.ntiles = P.ntx * P.nty * P.ntz;
    return P;
}

void append_tiles_xcd(std::vector<Tile> &out, const std::vector<TileRun> &runs)
{
    constexpr int XCDS = 8;
    std::vector<Tile> flat;
    for (const TileRun &r : runs)
        for (int z0 = 0; z0 < r.nz; z0 += r.tz)
            for (int c = 0; c < r.cols; ++c) flat.push_back(Tile{r.inst, c, z0, std::min(z0 + r.tz, r.nz)});
    const size_t total = flat.size();
    const size_t per = (total + XCDS - 1) / XCDS;
    for (size_t k = 0; k < per; ++k)
        for (size_t x = 0; x < XCDS; ++x) {
            const size_t t = x * per + k;
            if (t < total) out.push_back(flat[t]);
        }
}

long long max_passes(const dt_params &p)
{
    const long long per_iteration = 64;
    return 4096 + per_iteration * (std::max<long long>(p.it_max, 0) + 2);
}

int validate(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, const dt_params *p)
{
    if (algorithm != DT_GRADIENT && algorithm != DT_GREEDY) DT_FAIL("invalid algorithm");
    if (nx < 3 || ny < 3 || nz < 3)
        DT_FAIL("every image dimension must be at least 3 voxels (the stencils read 3 samples)");
    if (nx > (1u << 30) || ny > (1u << 30) || nz > (1u << 30)) DT_FAIL("image too large");
    // a field buffer is addressed by one 32-bit byte offset (12 bytes per voxel)
    if ((unsigned long long)nx * ny * nz * 12ull >= (1ull << 32))
        DT_FAIL("image too large: at most 357913941 voxels per solve");
    if (!(dx > 0.0f) || !(dy > 0.0f) || !(dz > 0.0f)) DT_FAIL("spacing must be positive");
    if (!p) DT_FAIL("null parameters");
    return 0;
}

// Times the fused-pass variants on a synthetic instance of this shape: a
// ball of J = 1.05 (mask = the ball) in J = 1, zero initial field, the solve
// control running with its stop tests disabled (epsilon = 0, theta = -1,
// iota = 0), so every timed launch evaluates a trial.
static int tune_gradient(int nx, int ny, int nz, float dx, float dy, float dz, int passes, dt_tune_result *out)
{
    if (passes <= 0) passes = 8;
    const int warm = 2;
    const long long n = (long long)nx * ny * nz;
    const int precision = precision_from_env();
    std::vector<float> T((size_t)n);
    std::vector<uint8_t> M((size_t)n);
    const double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
    const double r = std::max(1.0, std::min(nx, std::min(ny, nz)) / 3.0);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t i = ((size_t)z * ny + y) * nx + x;
                const bool in = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) <= r * r;
                T[i] = in ? 1.05f : 1.0f;
                M[i] = in ? 1 : 0;
            }
    dt_params p{};
    p.epsilon = 0.0f; p.tolerance = 0.2f; p.eta = 0.1f; p.eta_max = 0.4f; p.alpha = 1.2f; p.beta = 0.5f;
    p.gamma = 0.1f; p.delta = 1e-3f; p.zeta = 10.0f; p.theta = -1.0f; p.iota = 0.0f; p.strict = 0;
    p.it_max = 1LL << 40;

    size_t need = align_up(12 * (size_t)n, 256) + align_up(4 * (size_t)n, 256) + align_up((size_t)n, 256);
    size_t most = 0;
    for (int v : TUNE_CANDIDATES) {
        const Plan P = make_plan(DT_GRADIENT, nx, ny, nz, precision, v, 0);
        most = std::max(most, scratch_bytes(P, 0, true) + align_up(sizeof(Inst), 256) +
                                  align_up((size_t)P.ntiles * sizeof(Tile), 256));
    }
    Arena *A = nullptr;
    if (arena_get(need + most + 1024, &A)) return -1;
    Carver cv(A->ptr);
    float *d_field = cv.take<float>(3 * (size_t)n);
    float *d_T = cv.take<float>((size_t)n);
    uint8_t *d_M = cv.take<uint8_t>((size_t)n);
    const size_t base_off = cv.off;

    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemcpyAsync(d_T, T.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_M, M.data(), (size_t)n, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error(__LINE__, __FILE__, __func__, "cannot stage the tuning volume");
        rc = -1;
    }
    out->count = 0;
    out->best_variant = DEFAULT_GRAD_VARIANT;
    double best = 1e300;
    for (int v : TUNE_CANDIDATES) {
        if (rc) break;
        const Plan P = make_plan(DT_GRADIENT, nx, ny, nz, precision, v, 0);
        cv.off = base_off;
        const Inst h = make_inst(P, dx, dy, dz, p, d_T, d_M, d_field, cv, nullptr, 0, nullptr, nullptr);
        Inst *d_inst = cv.take<Inst>(1);
        Tile *d_tiles = cv.take<Tile>((size_t)P.ntiles);
        std::vector<Tile> tiles;
        append_tiles_xcd(tiles, {plan_run(0, P)});
        const Ctrl c0[2] = {initial_ctrl(DT_GRADIENT, p), initial_ctrl(DT_GRADIENT, p)};
        float ms = 0.0f;
        if (hipMemsetAsync(d_field, 0, 12 * (size_t)n, st) != hipSuccess ||
            hipMemcpyAsync(d_inst, &h, sizeof(Inst), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            hipMemcpyAsync(h.ctrl, c0, sizeof c0, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemsetAsync(h.acc, 0, ACC_SLOTS * ACC_COPIES * sizeof(Acc), st) != hipSuccess) {
            set_error(__LINE__, __FILE__, __func__, "cannot stage a tuning instance");
            rc = -1;
            break;
        }
        int launched = 0;   // the pass index's parity selects the control copy
        for (int i = 0; i < warm && !rc; ++i, ++launched)
            if (launch_grad_pass(precision, v, d_inst, d_tiles, P.ntiles, launched & 1, st) != hipSuccess) rc = -1;
        if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = -1;
        for (int i = 0; i < passes && !rc; ++i, ++launched)
            if (launch_grad_pass(precision, v, d_inst, d_tiles, P.ntiles, launched & 1, st) != hipSuccess) rc = -1;
        if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = -1;
        Ctrl fc[2];
        if (!rc && hipMemcpy(fc, h.ctrl, sizeof fc, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
        const Ctrl fin = latest_ctrl(fc);
        if (rc) {
            set_error(__LINE__, __FILE__, __func__, "tuning pass failed on the device");
            break;
        }
        const double us = 1e3 * ms / passes;
        if (out->count < DT_TUNE_MAX) {
            out->variant[out->count] = v;
            out->us_per_pass[out->count] = us;
            out->count++;
        }
        // a solve that stopped during the timing ran idle passes: not comparable
        if (fin.phase != PH_DONE && us < best) {
            best = us;
            out->best_variant = v;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (rc) return -1;
    int dev = 0;
    DT_HIP(hipGetDevice(&dev));
    g_tuned[TuneKey{dev, nx, ny, nz}] = out->best_variant;
    return 0;
}

// Pass variant of a solve.  Gradient: DISPTOOLS_GRAD_VARIANT, else the tuned
// variant of the shape when DISPTOOLS_AUTOTUNE=1, else the default; greedy:
// the fused pass unless DISPTOOLS_GREEDY_VARIANT=0.
// Column tile of a single FP32 gradient solve from a cost model of the
// streaming pass: t = 8.85 + S (0.301 + 1.80 NT / 1000) microseconds, S the
// plane steps of a pass (rounds of blocks x (TZ + 3), as choose_tz plans
// them) and NT the block's threads, fitted to measured pass times of the
// five streaming tiles over 18 volume shapes (32^3 to 256^3 and non-cubic:
// within 1-2 us).  The 64 x 12 tile stays unless another is predicted at
// least 10 % faster; over those shapes the rule is 6.5 % faster in sum than
// 64 x 12 everywhere and never slower (the measured best is 2 % faster
// still).  Deterministic: it depends on the shape and the CU count only.
int model_grad_variant(int nx, int ny, int nz, int precision)
{
    if (precision != 32) return DEFAULT_GRAD_VARIANT;
    const int cus = device_cus();
    auto predict = [&](int v) {
        int tx = 0, ty = 0;
        grad_variant_dims(v, tx, ty);
        const int nring = 2 * tx + 2 * ty;
        const double nt = tx * ty + (nring + 63) / 64 * 64;
        const int tz = choose_tz(nx, ny, nz, tx, ty, cus, cus, 3, true);
        const long long cols = (long long)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
        const long long blocks = cols * ((nz + tz - 1) / tz);
        const double steps = (double)((blocks + cus - 1) / cus) * (std::min(tz, nz) + 3);
        return 8.85 + steps * (0.301 + 1.80 * nt / 1000.0);
    };
    const double t_def = predict(DEFAULT_GRAD_VARIANT);
    int best = DEFAULT_GRAD_VARIANT;
    double t_best = t_def;
    for (int v = 0; v < GRAD_VARIANTS; ++v) {
        const double t = predict(v);
        if (t < t_best) { t_best = t; best = v; }
    }
    return t_def >= 1.10 * t_best ? best : DEFAULT_GRAD_VARIANT;
}

static int solve_variant(int algorithm, int nx, int ny, int nz, float dx, float dy, float dz)
{
    if (algorithm != DT_GRADIENT) return greedy_solve_variant(nx, ny, precision_from_env());
    if (std::getenv("DISPTOOLS_GRAD_VARIANT")) return variant_from_env();
    const char *t = std::getenv("DISPTOOLS_AUTOTUNE");
    if (!t || std::strcmp(t, "1") != 0) return model_grad_variant(nx, ny, nz, precision_from_env());
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return DEFAULT_GRAD_VARIANT;
    auto it = g_tuned.find(TuneKey{dev, nx, ny, nz});
    if (it != g_tuned.end()) return it->second;
    dt_tune_result r{};
    if (tune_gradient(nx, ny, nz, dx, dy, dz, 0, &r)) {
        disptools_clear_error();
        return DEFAULT_GRAD_VARIANT;
    }
    return r.best_variant;
}

// Core solve on device-resident buffers.
static int solve_device(int algorithm, int nx, int ny, int nz, float dx, float dy, float dz,
                        const float *d_target, const uint8_t *d_mask, const dt_params *p, float *d_field,
                        dt_solve_info *info, double *trace, long long trace_cap, hipStream_t st)
{
    const auto t_start = std::chrono::steady_clock::now();
    const MarkerRange range("dt_solve %s %dx%dx%d", algorithm == DT_GRADIENT ? "gradient" : "greedy", nx, ny, nz);
    const int precision = precision_from_env();
    const Plan P = make_plan(algorithm, nx, ny, nz, precision, solve_variant(algorithm, nx, ny, nz, dx, dy, dz), 0);
    const long long n = P.n;
    const int variant = P.variant;
    const int ntiles = P.ntiles;
    if (trace_cap < 0) trace_cap = 0;
    if (!trace) trace_cap = 0;

    Arena *A = nullptr;
    const size_t head = sizeof(SolveHead) + (size_t)ntiles * sizeof(Tile);
    const size_t own = scratch_bytes(P, trace_cap, false);
    if (arena_get(align_up(head, 256) + own + 256, &A)) return -1;
    if (arena_stage(*A, head)) return -1;
    Carver cv(A->ptr);
    SolveHead *d_head = cv.take<SolveHead>(1);
    Tile *d_tiles = cv.take<Tile>((size_t)ntiles);
    Inst *d_inst = &d_head->inst;
    if ((char *)d_tiles != (char *)d_head + sizeof(SolveHead)) DT_FAIL("solve head layout");
    Inst h = make_inst(P, dx, dy, dz, *p, d_target, d_mask, d_field, cv, trace, trace_cap, d_head->ctrl,
                       d_head->acc);
    unsigned long long *status = nullptr;
    Ctrl *fin_word = nullptr;   // the control block, published by the device once done
    if (status_word(&status, &h.status, &fin_word, &h.final_ctrl)) return -1;
    if (!status_words_from_env()) h.status = nullptr, h.final_ctrl = nullptr;

    // head and tile table in one copy from pinned staging (the previous solve
    // on this device has finished with the staging: solves end synchronised)
    SolveHead *sh = (SolveHead *)A->stage;
    sh->inst = h;
    sh->ctrl[0] = sh->ctrl[1] = initial_ctrl(algorithm, *p);
    std::memset(sh->acc, 0, sizeof sh->acc);
    std::vector<Tile> tiles;
    tiles.reserve(ntiles);
    append_tiles_xcd(tiles, {plan_run(0, P)});
    std::memcpy(A->stage + sizeof(SolveHead), tiles.data(), tiles.size() * sizeof(Tile));
    DT_HIP(hipMemcpyAsync(d_head, A->stage, head, hipMemcpyHostToDevice, st));
    if (algorithm == DT_GREEDY) {
        // the second field buffer starts at zero: its outer face is never written
        // by the greedy step (src/disptools.c:61, displacement_field_greedy.c:43-65)
        DT_HIP(hipMemsetAsync(h.buf[1], 0, 3 * (size_t)n * sizeof(float), st));
    }

    // timing events around the solve's back-to-back pass launches: the device
    // time of the passes (no host turnaround sits between them)
    hipEvent_t *kev = A->kev;

    // passes needed in the worst case before the controller stops the solve
    const long long pass_cap = max_passes(*p);
    long long launched = 0;
    int rc = 0;
    auto enqueue = [&](int k) -> int {
        for (int i = 0; i < k; ++i, ++launched) {
            // a pass reads the control copy of its parity
            const int parity = (int)(launched & 1);
            if (algorithm == DT_GRADIENT) {
                DT_HIP(launch_grad_pass(precision, variant, d_inst, d_tiles, ntiles, parity, st));
            } else {
                DT_HIP(launch_greedy_pass(precision, variant, d_inst, d_tiles, ntiles, parity, st));
            }
        }
        return 0;
    };

    // Two chunks in flight; the host follows the completion word that the
    // start of each pass stores in host memory (the passes before it have run)
    // and enqueues the next chunk once the older one has run, so nothing but
    // pass kernels enters the stream.  Small chunks bound the passes that run
    // idle after the solve has stopped.  The controller step of a pass's trial
    // runs at the start of the next pass, which is always enqueued: the host
    // stops only once the word reports the end.
    const int chunk = 2;
    rc = hipEventRecord(kev[0], st) == hipSuccess ? 0 : -1;
    if (!rc) rc = enqueue(chunk);
    long long target = launched;   // active passes the host waits for
    if (!rc) rc = enqueue(chunk);
    bool published = false;   // the end of the solve seen through the completion word
    while (!rc) {
        bool done = false;
        const int w = follow_status(status, target, st, done);
        if (w < 0) { set_error(__LINE__, __FILE__, __func__, "device error during solve"); rc = -1; break; }
        published = w == 0 && done;
        if (w > 0) {
            // the stream drained but the word did not follow: read the control copies
            Ctrl c[2];
            if (hipMemcpy(c, h.ctrl, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) {
                set_error(__LINE__, __FILE__, __func__, "device error during solve");
                rc = -1;
                break;
            }
            done = latest_ctrl(c).phase == PH_DONE;
        }
        if (done) break;
        if (launched > pass_cap) {
            set_error(__LINE__, __FILE__, __func__, "solver did not terminate (check alpha > 1 and beta < 1)");
            rc = -1;
            break;
        }
        target = launched;
        rc = enqueue(chunk);
    }
    const bool timed = hipEventRecord(kev[1], st) == hipSuccess;
    // the pass that ended the solve has copied the result into the caller's
    // field (copy_result_tile); the final control block comes from the device
    // unless the device published it
    Ctrl *fin_copies = (Ctrl *)(A->stage + offsetof(SolveHead, ctrl));
    hipError_t se = hipSuccess;
    if (!rc && !published)
        se = hipMemcpyAsync(fin_copies, h.ctrl, 2 * sizeof(Ctrl), hipMemcpyDeviceToHost, st);
    const hipError_t ss = hipStreamSynchronize(st);
    if (se == hipSuccess) se = ss;
    float kernel_ms = 0.0f;
    if (se != hipSuccess || !timed || hipEventElapsedTime(&kernel_ms, kev[0], kev[1]) != hipSuccess) kernel_ms = 0.0f;
    if (se != hipSuccess && !rc) { set_error(__LINE__, __FILE__, __func__, hipGetErrorString(se)); rc = -1; }
    if (rc) return -1;

    Ctrl fin;
    if (published) {
        std::memcpy(&fin, fin_word, sizeof(Ctrl));
    } else {
        Ctrl c[2];
        std::memcpy(c, fin_copies, sizeof c);
        fin = latest_ctrl(c);
    }
    if (trace && trace_cap > 0) {
        const long long nl = std::min(fin.trace_len, trace_cap);
        if (nl > 0) DT_HIP(hipMemcpy(trace, h.trace, (size_t)nl * 4 * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (info) {
        info->iterations = fin.it;
        info->evaluations = fin.evaluations;
        info->rejections = fin.rejections;
        info->launches = launched;
        info->reason = fin.reason;
        info->final_error = fin.E_result;
        info->final_max_error = fin.m_result;
        info->trace_len = fin.trace_len;
        info->field_io = fin.field_io;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        info->kernel_seconds = (double)kernel_ms * 1e-3;
    }
    return 0;
}

}  // namespace dt

using namespace dt;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" int dt_solve_device(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                               const float *d_target, const unsigned char *d_mask, const dt_params *params,
                               float *d_field, dt_solve_info *info, double *trace, long long trace_cap,
                               void *stream)
{
    std::lock_guard<std::mutex> g(g_lock);
    disptools_clear_error();
    if (validate(algorithm, nx, ny, nz, dx, dy, dz, params)) return -1;
    if (!d_target || !d_mask || !d_field) DT_FAIL("null buffer");
    return solve_device(algorithm, (int)nx, (int)ny, (int)nz, dx, dy, dz, d_target, d_mask, params, d_field,
                        info, trace, trace_cap, (hipStream_t)stream);
}

extern "C" int dt_solve_host(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                             const float *target, const bool *mask, const dt_params *params, float *field,
                             dt_solve_info *info, double *trace, long long trace_cap)
{
    std::lock_guard<std::mutex> g(g_lock);
    disptools_clear_error();
    if (validate(algorithm, nx, ny, nz, dx, dy, dz, params)) return -1;
    if (!target || !mask || !field) DT_FAIL("null buffer");
    const size_t n = nx * ny * nz;
    void *io = nullptr;
    const size_t bytes = align_up(4 * n * sizeof(float), 256) + align_up(n, 256);
    if (hipMalloc(&io, bytes) != hipSuccess) {
        (void)hipGetLastError();
        DT_FAIL("cannot allocate device memory for the inputs");
    }
    float *d_field = (float *)io;
    float *d_target = d_field + 3 * n;
    uint8_t *d_mask = (uint8_t *)((char *)io + align_up(4 * n * sizeof(float), 256));
    int rc = 0;
    hipStream_t st = nullptr;
    if (hipMemcpy(d_field, field, 3 * n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_target, target, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_mask, mask, n, hipMemcpyHostToDevice) != hipSuccess) {
        set_error(__LINE__, __FILE__, __func__, "host to device copy failed");
        rc = -1;
    }
    if (!rc) rc = solve_device(algorithm, (int)nx, (int)ny, (int)nz, dx, dy, dz, d_target, d_mask, params,
                               d_field, info, trace, trace_cap, st);
    if (!rc && hipMemcpy(field, d_field, 3 * n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(__LINE__, __FILE__, __func__, "device to host copy failed");
        rc = -1;
    }
    (void)hipFree(io);
    return rc;
}

// The reference's release build prints one line per iteration and a
// termination message (DISPTOOLS_VERBOSE, setup.py:50; verbose_printf,
// src/headers/disptools.h:87-91); its entry points here print the same lines
// from the solve's trace after the solve, unless DISPTOOLS_VERBOSE=0.
dt::MarkerRange::MarkerRange(const char *fmt, ...)
{
    char name[128];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(name, sizeof name, fmt, ap);
    va_end(ap);
    roctxRangePushA(name);
}

dt::MarkerRange::~MarkerRange() { roctxRangePop(); }

bool dt::verbose_from_env()
{
    const char *s = std::getenv("DISPTOOLS_VERBOSE");
    return !(s && std::strcmp(s, "0") == 0);
}

void dt::print_iteration(long long it, double E, double m, double eta)
{
    std::printf("Iteration %5lld:  total error %6e  max voxel error %6e  eta %6e\n", it, (double)(float)E,
                (double)(float)m, (double)(float)eta);
}

// Lines of displacement_field_gradient.c:272-418 / displacement_field_greedy.c:184-236.
static void print_solve_log(int algorithm, const dt_solve_info &info, const double *tr, long long lines,
                            long long it_max)
{
    for (long long k = 0; k < lines; ++k) {
        const double *row = tr + 4 * k;
        print_iteration((long long)row[0], row[1], row[2], row[3]);
        // greedy: a rejected trial is retried with a smaller step (same iteration)
        if (algorithm == DT_GREEDY && k + 1 < lines && tr[4 * (k + 1)] == row[0])
            std::printf("Error not decreasing, reducing step size to %.4e\n", (double)(float)tr[4 * (k + 1) + 3]);
    }
    switch (info.reason) {
    case DT_STOP_ERROR_EXPLODED: std::printf("Terminating: error exploded.\n"); break;
    case DT_STOP_ETA_SMALL:
    case DT_STOP_NO_IMPROVEMENT: std::printf("Error not decreasing, terminating.\n"); break;
    case DT_STOP_VOXEL_EXPLODED: std::printf("Terminating: voxel error exploded.\n"); break;
    case DT_STOP_TOLERANCE: std::printf("Terminating: reached desired tolerance.\n"); break;
    default: break;
    }
    // printed when the loop counter stopped at it_max (:418), i.e. a stop at
    // the last iteration; a loop that ran out ends at it_max + 1
    if (info.reason != DT_STOP_MAX_ITERATIONS && info.iterations == it_max)
        std::printf("Terminating: reached maximum number of iterations.\n");
    std::fflush(stdout);
}

static void reference_entry(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                            const float *J, const bool *mask, float epsilon, float tolerance, float eta,
                            float eta_max, float alpha, float beta, float gamma, float delta, float zeta,
                            float theta, float iota, bool strict, size_t it_max, float *field)
{
    dt_params p;
    p.epsilon = epsilon; p.tolerance = tolerance; p.eta = eta; p.eta_max = eta_max;
    p.alpha = alpha; p.beta = beta; p.gamma = gamma; p.delta = delta; p.zeta = zeta;
    p.theta = theta; p.iota = iota; p.strict = strict ? 1 : 0;
    p.it_max = (long long)it_max;
    if (!verbose_from_env()) {
        (void)dt_solve_host(algorithm, nx, ny, nz, dx, dy, dz, J, mask, &p, field, nullptr, nullptr, 0);
        return;
    }
    // one line per evaluated trial at most (greedy logs rejected trials too)
    const long long cap = std::min<long long>(max_passes(p), 1ll << 21);
    std::vector<double> tr(4 * (size_t)cap);
    dt_solve_info info{};
    if (dt_solve_host(algorithm, nx, ny, nz, dx, dy, dz, J, mask, &p, field, &info, tr.data(), cap) == 0)
        print_solve_log(algorithm, info, tr.data(), std::min(info.trace_len, cap), p.it_max);
}

extern "C" void generate_displacement_gradient_cuda(
    const size_t nx, const size_t ny, const size_t nz, const float dx, const float dy, const float dz,
    const float *J, const bool *mask, const float epsilon, const float tolerance, float eta,
    const float eta_max, const float alpha, const float beta, const float gamma, const float delta,
    const float zeta, const float theta, const float iota, const bool strict, const size_t it_max,
    float *field)
{
    reference_entry(DT_GRADIENT, nx, ny, nz, dx, dy, dz, J, mask, epsilon, tolerance, eta, eta_max, alpha,
                    beta, gamma, delta, zeta, theta, iota, strict, it_max, field);
}

extern "C" void generate_displacement_greedy_cuda(
    const size_t nx, const size_t ny, const size_t nz, const float dx, const float dy, const float dz,
    const float *J, const bool *mask, const float epsilon, const float tolerance, float eta,
    const float eta_max, const float alpha, const float beta, const float gamma, const float delta,
    const float zeta, const float theta, const float iota, const bool strict, const size_t it_max,
    float *field)
{
    reference_entry(DT_GREEDY, nx, ny, nz, dx, dy, dz, J, mask, epsilon, tolerance, eta, eta_max, alpha,
                    beta, gamma, delta, zeta, theta, iota, strict, it_max, field);
}

// The reference's CPU-named solvers (src/headers/displacement_field_gradient.h:7-30,
// src/headers/displacement_field_greedy.h:7-30), which its extension calls when
// gpu_id < 0 (python_c_extension/_disptools.c:374-384).  This library has one
// implementation, on HIP devices: these run it on the calling thread's current
// device (device 0 unless set_device chose another), with the semantics and
// log lines of src/displacement_field_gradient.c:191-431 and
// src/displacement_field_greedy.c:85-247.  There is no CPU path behind them.
extern "C" void generate_displacement_gradient(
    const size_t nx, const size_t ny, const size_t nz, const float dx, const float dy, const float dz,
    const float *J, const bool *mask, const float epsilon, const float tolerance, float eta,
    const float eta_max, const float alpha, const float beta, const float gamma, const float delta,
    const float zeta, const float theta, const float iota, const bool strict, const size_t it_max,
    float *field)
{
    reference_entry(DT_GRADIENT, nx, ny, nz, dx, dy, dz, J, mask, epsilon, tolerance, eta, eta_max, alpha,
                    beta, gamma, delta, zeta, theta, iota, strict, it_max, field);
}

extern "C" void generate_displacement_greedy(
    const size_t nx, const size_t ny, const size_t nz, const float dx, const float dy, const float dz,
    const float *J, const bool *mask, const float epsilon, const float tolerance, float eta,
    const float eta_max, const float alpha, const float beta, const float gamma, const float delta,
    const float zeta, const float theta, const float iota, const bool strict, const size_t it_max,
    float *field)
{
    reference_entry(DT_GREEDY, nx, ny, nz, dx, dy, dz, J, mask, epsilon, tolerance, eta, eta_max, alpha,
                    beta, gamma, delta, zeta, theta, iota, strict, it_max, field);
}

extern "C" int dt_jacobian_device(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                                  const float *d_field, float *d_J, void *stream)
{
    if (nx < 3 || ny < 3 || nz < 3)
        DT_FAIL("every image dimension must be at least 3 voxels (the stencils read 3 samples)");
    if (nx * ny * nz > 357913941u) DT_FAIL("image too large: at most 357913941 voxels per jacobian");
    if (!d_field || !d_J) DT_FAIL("null buffer");
    const float ix = (float)(1.0 / (double)dx), iy = (float)(1.0 / (double)dy), iz = (float)(1.0 / (double)dz);
    DT_HIP(launch_jacobian(d_field, d_J, (int)nx, (int)ny, (int)nz, ix, iy, iz, (hipStream_t)stream));
    return 0;
}

extern "C" void jacobian_dynamic(const size_t nx, const size_t ny, const size_t nz, const float dx,
                                 const float dy, const float dz, const float *f, float *J)
{
    disptools_clear_error();
    if (nx < 3 || ny < 3 || nz < 3) {
        set_error(__LINE__, __FILE__, __func__,
                  "every image dimension must be at least 3 voxels (the stencils read 3 samples)");
        return;
    }
    const size_t n = nx * ny * nz;
    void *mem = nullptr;
    if (hipMalloc(&mem, 4 * n * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        set_error(__LINE__, __FILE__, __func__, "cannot allocate device memory");
        return;
    }
    float *d_f = (float *)mem, *d_J = d_f + 3 * n;
    if (hipMemcpy(d_f, f, 3 * n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        dt_jacobian_device(nx, ny, nz, dx, dy, dz, d_f, d_J, nullptr) != 0 ||
        hipMemcpy(J, d_J, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        if (!g_error.error) set_error(__LINE__, __FILE__, __func__, "jacobian failed on the device");
    }
    (void)hipFree(mem);
}

extern "C" void regularise(const size_t nx, const size_t ny, const size_t nz, float *J, float epsilon)
{
    disptools_clear_error();
    const size_t n = nx * ny * nz;
    if (n == 0) return;
    void *mem = nullptr;
    if (hipMalloc(&mem, n * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        set_error(__LINE__, __FILE__, __func__, "cannot allocate device memory");
        return;
    }
    float *d_J = (float *)mem;
    if (hipMemcpy(d_J, J, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        launch_regularise(d_J, (long long)n, epsilon, nullptr) != hipSuccess ||
        hipMemcpy(J, d_J, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(__LINE__, __FILE__, __func__, "regularise failed on the device");
    }
    (void)hipFree(mem);
}

extern "C" int dt_tune_gradient(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, int passes,
                                dt_tune_result *out)
{
    std::lock_guard<std::mutex> g(g_lock);
    disptools_clear_error();
    dt_params p{};
    if (validate(DT_GRADIENT, nx, ny, nz, dx, dy, dz, &p)) return -1;
    if (!out) DT_FAIL("null result");
    return tune_gradient((int)nx, (int)ny, (int)nz, dx, dy, dz, passes, out);
}

extern "C" int dt_scale_volume_device(size_t nx, size_t ny, size_t nz, size_t ncomp, float dx, float dy, float dz,
                                      const float *d_in, size_t mx, size_t my, size_t mz, float sigma,
                                      int nearest, float *d_out, void *stream)
{
    std::lock_guard<std::mutex> g(g_lock);
    disptools_clear_error();
    if (!d_in || !d_out) DT_FAIL("null buffer");
    if (nx < 1 || ny < 1 || nz < 1 || mx < 1 || my < 1 || mz < 1 || ncomp < 1) DT_FAIL("empty volume");
    if (!(dx > 0.0f) || !(dy > 0.0f) || !(dz > 0.0f)) DT_FAIL("spacing must be positive");
    if (!(sigma >= 0.0f)) DT_FAIL("sigma must be non-negative");
    const size_t n = nx * ny * nz * ncomp;
    if (n >= (1ull << 40)) DT_FAIL("volume too large");
    hipStream_t st = (hipStream_t)stream;
    Arena *A = nullptr;
    if (arena_get(2 * align_up(n * sizeof(float), 256) + 512, &A)) return -1;
    Carver cv(A->ptr);
    float *t0 = cv.take<float>(n);
    float *t1 = cv.take<float>(n);
    if (scale_volume((int)nx, (int)ny, (int)nz, (int)ncomp, dx, dy, dz, d_in, (int)mx, (int)my, (int)mz, sigma,
                     nearest != 0, d_out, t0, t1, st))
        return -1;
    // the arena is reused by the next call: finish before returning
    DT_HIP(hipStreamSynchronize(st));
    return 0;
}
