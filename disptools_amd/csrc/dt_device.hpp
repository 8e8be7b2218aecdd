// dt_device.hpp -- device helpers shared by the pass kernels: block reductions,
// the last-block arrival protocol, the solver controllers and tile decoding.
#pragma once

#include <math.h>

#include <algorithm>

#include "dt_internal.hpp"
#include "dt_stencil.hpp"

namespace dt {

// Global-address-space pointers.  Buffer pointers read from the instance table
// are generic; casting them lets the compiler emit global_* instead of flat_*
// accesses, which would also count against the LDS wait counter.
#define DT_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ DT_GLOBAL T *gptr(T *p)
{
    return (DT_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ const DT_GLOBAL T *gptr(const T *p)
{
    return (const DT_GLOBAL T *)p;
}

// Wave-uniform copies of values read from the (writable) control block: the
// compiler cannot prove them uniform, and a buffer descriptor built from a
// possibly divergent pointer costs a waterfall loop per access.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned uni(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uni(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
template <typename T>
__device__ __forceinline__ T *uni(T *p)
{
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Block reduction of (sum a, max b, sum c) to thread 0, fixed order.
template <int NT>
__device__ __forceinline__ void block_reduce3(double &a, float &b, double &c, double *red)
{
    constexpr int NW = NT / 64;
    a = wave_sum(a);
    b = wave_max(b);
    c = wave_sum(c);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        red[3 * w + 0] = a;
        red[3 * w + 1] = (double)b;
        red[3 * w + 2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sc = 0.0;
        float mb = 0.0f;
        for (int i = 0; i < NW; ++i) {
            sa += red[3 * i + 0];
            mb = fmaxf(mb, (float)red[3 * i + 1]);
            sc += red[3 * i + 2];
        }
        a = sa; b = mb; c = sc;
    }
}

__device__ __forceinline__ void store_partial(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_partial(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes this block's partials; returns true in every thread of the block
// that arrived last for its instance.
template <int NT>
__device__ bool arrive(const Inst &I, int tile, double E, float m, double G, int *s_flag)
{
    if (threadIdx.x == 0) {
        double *p = I.partials + 4 * (long long)tile;
        store_partial(p + 0, E);
        store_partial(p + 1, (double)m);
        store_partial(p + 2, G);
        const unsigned prev = __hip_atomic_fetch_add(I.counter, 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        *s_flag = (prev + 1 == (unsigned)(I.ntx * I.nty * I.ntz)) ? 1 : 0;
    }
    __syncthreads();
    return *s_flag != 0;
}

// Deterministic fold of an instance's partials by the last block.
template <int NT>
__device__ void fold_partials(const Inst &I, double &E, float &m, double &G, double *red)
{
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    const int ntiles = I.ntx * I.nty * I.ntz;
    double a = 0.0, c = 0.0;
    float b = 0.0f;
    for (int t = threadIdx.x; t < ntiles; t += NT) {
        const double *p = I.partials + 4 * (long long)t;
        a += load_partial(p + 0);
        b = fmaxf(b, (float)load_partial(p + 1));
        c += load_partial(p + 2);
    }
    block_reduce3<NT>(a, b, c, red);
    E = a; m = b; G = c;
}

// ---------------------------------------------------------------------------
// Controllers (run by one thread after the instance's partials are folded)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void log_line(const Inst &I, Ctrl *C, double it, double E, double m, double eta)
{
    if (I.trace && C->trace_len < I.trace_cap) {
        double *t = I.trace + 4 * C->trace_len;
        t[0] = it; t[1] = E; t[2] = m; t[3] = eta;
    }
    C->trace_len++;
}

// Completion word of a solve in host-coherent memory: (active passes << 1) |
// done, stored after every controller step, so the host follows the solve
// without a copy or an event between pass launches.  Once the solve is done
// the control block itself is published first (when the instance has a host
// copy), ordered before the word by a system-scope release.
__device__ __forceinline__ void publish_status(const Inst &I, const Ctrl *C)
{
    if (!I.status) return;
    const bool done = C->phase == PH_DONE;
    if (done && I.final_ctrl) {
        const unsigned long long *src = (const unsigned long long *)C;
        unsigned long long *dst = (unsigned long long *)I.final_ctrl;
        for (int k = 0; k < (int)(sizeof(Ctrl) / 8); ++k)
            __hip_atomic_store(dst + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __atomic_thread_fence(__ATOMIC_RELEASE);
    }
    __hip_atomic_store(I.status, ((unsigned long long)C->evaluations << 1) | (done ? 1ull : 0ull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void finish(Ctrl *C, int reason)
{
    C->reason = reason;
    C->phase = PH_DONE;
    C->active = 0;
    C->result_buf = C->base;
    C->E_result = C->E_last;
    C->m_result = C->accepted ? C->m_last : C->m0;
}

__device__ __forceinline__ void free_pair(int a, int b, int &f0, int &f1)
{
    int k = 0;
    for (int i = 0; i < NBUF; ++i) {
        if (i == a || i == b) continue;
        if (k == 0) f0 = i; else f1 = i;
        ++k;
    }
}

__device__ __forceinline__ float capped_next(float eta, const dt_params &p)
{
    float e = eta * p.alpha;
    return e > p.eta_max ? p.eta_max : e;
}

// Trial base + (eta / eta_c)(cand - base) written to a scratch buffer, with
// its speculative candidate in the other scratch buffer.
__device__ __forceinline__ void start_interp(Ctrl *C, const dt_params &p, bool store)
{
    int f0, f1;
    free_pair(C->base, C->cand_buf, f0, f1);
    C->mode = TRIAL_INTERP;
    C->src = C->base;
    C->cand = C->cand_buf;
    C->s = C->eta / C->eta_c;
    C->out_trial = store ? f0 : -1;
    C->out_next = store ? f1 : -1;
    C->eta_next = capped_next(C->eta, p);
}

// Gradient solver control flow (src/displacement_field_gradient.c:250-421).
// E, m, G are the energy, masked max error and squared gradient norm of the
// trial the pass evaluated.
static __device__ void grad_controller(const Inst &I, Ctrl *C, double E, float m, double G)
{
    const dt_params &p = I.p;
    C->evaluations++;
    // algorithmic traffic of the pass that was just evaluated
    C->field_io += (C->mode == TRIAL_INTERP ? 2 : 1) + (C->out_trial >= 0 ? 1 : 0) + (C->out_next >= 0 ? 1 : 0);
    switch (C->phase) {
    case PH_INIT: {
        // trial = initial guess in buf[src]; its candidate was formed at eta * alpha
        C->base = C->src;
        C->E0 = E; C->m0 = m;
        C->E_last = E;
        C->G_base = G;
        log_line(I, C, 0.0, E, m, C->eta);
        if (p.it_max <= 0) { finish(C, DT_STOP_MAX_ITERATIONS); return; }
        C->cand_buf = C->out_next;
        C->eta_c = C->eta_next;
        C->eta = C->eta_next;              // eta *= alpha (:291)
        C->phase = PH_SEARCH;
        C->mode = TRIAL_DIRECT;            // first search trial is the candidate itself
        C->src = C->cand_buf;
        C->out_trial = -1;
        C->out_next = -1;                  // search trials are never accepted
        return;
    }
    case PH_SEARCH: {
        const bool go_on = (C->eta < p.eta_max) &&
                           (E - C->E_last > -(double)p.gamma * (double)C->eta * C->G_base);
        if (go_on) {
            C->eta *= p.alpha;
            start_interp(C, p, false);
            return;
        }
        C->eta /= p.alpha * p.alpha;       // (:317)
        C->E_last = C->E0;                 // recomputed initial error (:320-326)
        C->m_last = INFINITY;
        C->it = 1;
        C->phase = PH_MAIN;
        C->eta = capped_next(C->eta, p);   // (:342-343)
        start_interp(C, p, true);
        return;
    }
    case PH_MAIN: {
        const bool eta_good = C->eta >= p.iota;
        const bool armijo_fails = E - C->E_last > -(double)p.gamma * (double)C->eta * C->G_base;
        const bool strict_fails = p.strict && (double)m > C->m_last;
        if (eta_good && (armijo_fails || strict_fails)) {
            C->eta *= p.beta;
            C->rejections++;
            start_interp(C, p, true);
            return;
        }
        log_line(I, C, (double)C->it, E, m, C->eta);
        if (!isnormal((float)E)) { finish(C, DT_STOP_ERROR_EXPLODED); return; }
        if (C->eta < p.iota) { finish(C, DT_STOP_ETA_SMALL); return; }
        if (1.0 - E / C->E_last < (double)p.theta) { finish(C, DT_STOP_NO_IMPROVEMENT); return; }
        if (!isnormal(m)) { finish(C, DT_STOP_VOXEL_EXPLODED); return; }
        if (m < p.epsilon) { finish(C, DT_STOP_TOLERANCE); return; }
        // accept the trial
        C->E_last = E;
        C->m_last = m;
        C->G_base = G;
        C->accepted = 1;
        C->base = (C->mode == TRIAL_DIRECT) ? C->src : C->out_trial;
        C->cand_buf = C->out_next;
        C->eta_c = C->eta_next;
        C->it++;
        if (C->it > p.it_max) { finish(C, DT_STOP_MAX_ITERATIONS); return; }
        C->eta = capped_next(C->eta, p);
        int f0, f1;
        free_pair(C->base, C->cand_buf, f0, f1);
        C->mode = TRIAL_DIRECT;
        C->src = C->cand_buf;
        C->out_trial = -1;
        C->out_next = f0;
        C->eta_next = capped_next(C->eta, p);
        return;
    }
    default:
        return;
    }
}

// Greedy solver control flow (src/displacement_field_greedy.c:149-238).
// `step_io` is the field volumes a pass with the step moves (step kernel +
// evaluation kernel: 3; fused pass: 2); the evaluation of the initial guess
// reads one.
static __device__ void greedy_controller(const Inst &I, Ctrl *C, double E, float m, int step_io)
{
    const dt_params &p = I.p;
    C->evaluations++;
    C->field_io += C->phase == PH_INIT ? 1 : step_io;
    const int e_written = C->e_new;
    if (C->phase == PH_INIT) {
        C->E0 = E; C->m0 = m;
        C->E_last = E;
        C->m_last = m;
        C->e_cur = e_written;
        C->e_new = 1 - e_written;
        if (p.it_max <= 0) { finish(C, DT_STOP_MAX_ITERATIONS); return; }
        C->it = 1;
        C->phase = PH_MAIN;
        C->step_on = 1;
        C->g_new = 1 - C->base;
        return;
    }
    // the trial is buf[g_new]; its error map is err[e_written]
    log_line(I, C, (double)C->it, E, m, C->eta);
    C->e_cur = e_written;                 // the next step reads the newest map (:165-181)
    C->e_new = 1 - e_written;
    if (!isnormal((float)E)) { finish(C, DT_STOP_ERROR_EXPLODED); return; }
    if (E >= C->E_last) {
        C->eta *= p.beta;
        C->rejections++;
        if (C->eta < p.iota) { finish(C, DT_STOP_ETA_SMALL); return; }
        return;                            // repeat the iteration from the same base
    }
    if (1.0 - E / C->E_last < (double)p.theta) { finish(C, DT_STOP_NO_IMPROVEMENT); return; }
    if (!isnormal(m)) { finish(C, DT_STOP_VOXEL_EXPLODED); return; }
    if (m < p.epsilon) { finish(C, DT_STOP_TOLERANCE); return; }
    C->E_last = E;
    C->m_last = m;
    C->accepted = 1;
    const int nb = C->g_new;
    C->g_new = C->base;
    C->base = nb;
    C->it++;
    if (C->it > p.it_max) { finish(C, DT_STOP_MAX_ITERATIONS); return; }
}

// ---------------------------------------------------------------------------
// Tile decoding
// ---------------------------------------------------------------------------

struct TileCoord {
    int x0, y0, z0, z1;   // first column, first plane, one past the last plane
};

template <int TX, int TY>
__device__ __forceinline__ TileCoord decode_tile(const Inst &I, int t)
{
    TileCoord c;
    const int tx = t % I.ntx;
    const int r = t / I.ntx;
    const int ty = r % I.nty;
    const int tz = r / I.nty;
    c.x0 = tx * TX;
    c.y0 = ty * TY;
    c.z0 = tz * I.tz;
    c.z1 = min(c.z0 + I.tz, I.nz);
    return c;
}

}  // namespace dt
