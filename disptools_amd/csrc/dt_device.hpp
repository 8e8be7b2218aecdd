// dt_device.hpp -- device helpers shared by the pass kernels: exact pass sums,
// the solver controllers, the controller step at the start of a pass and tile
// decoding.
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

// One LDS cell: the three values of a point (field components, or a relation
// axis' three weighted cofactors), 16-byte aligned, moved by one 96-bit LDS
// access.
typedef float f3 __attribute__((ext_vector_type(3)));
static_assert(sizeof(f3) == 16, "an LDS cell is 16 bytes");

// ---------------------------------------------------------------------------
// Exact sums (the Acc layout: 128-bit fixed point, quantum 2^-64, 32-bit limbs)
// ---------------------------------------------------------------------------

typedef unsigned __int128 u128;

// Adds the term d >= 0 to a fixed-point sum: floor(d * 2^64).  Terms at or
// above 2^32 (or infinite) set the `huge` flag bit and NaN the `nan` bit
// instead: the sum is then reported as +inf or NaN.  The three 32-bit limbs
// come from truncating conversions: floor(d), then the fraction scaled by
// 2^32 twice; every subtraction and scaling is exact (d has 53 significant
// bits, the fraction of an exact integer split is representable), so the
// limbs are the exact binary digits of d down to 2^-64.
__device__ __forceinline__ void fix_add(u128 &acc, double d, unsigned &flags, unsigned huge, unsigned nan)
{
    const bool is_nan = d != d;
    const bool is_huge = d >= 0x1p32;
    const double x = (is_nan || is_huge) ? 0.0 : d;
    const unsigned l2 = __double2uint_rz(x);
    double r = (x - (double)l2) * 0x1p32;
    const unsigned l1 = __double2uint_rz(r);
    r = (r - (double)l1) * 0x1p32;
    const unsigned l0 = __double2uint_rz(r);
    acc += ((u128)l2 << 64) | ((u128)l1 << 32) | (u128)l0;
    flags |= (is_huge ? huge : 0u) | (is_nan ? nan : 0u);
}

// x of the lane a row-local DPP control selects (the whole wave active).
template <int CTRL>
__device__ __forceinline__ unsigned dpp_u32(unsigned x)
{
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x)
{
    return (unsigned long long)dpp_u32<CTRL>((unsigned)x) |
           ((unsigned long long)dpp_u32<CTRL>((unsigned)(x >> 32)) << 32);
}
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long x, int lane)
{
    return (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, lane) |
           ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), lane) << 32);
}

template <int CTRL>
__device__ __forceinline__ u128 dpp_u128(u128 x)
{
    u128 r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) r |= (u128)dpp_u32<CTRL>((unsigned)(x >> (32 * k))) << (32 * k);
    return r;
}
__device__ __forceinline__ u128 readlane_u128(u128 x, int lane)
{
    u128 r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r |= (u128)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> (32 * k)), lane) << (32 * k);
    return r;
}

// Wave reductions (the whole wave active; results wave-uniform).  Four
// row-local DPP steps -- lane xor 1, lane xor 2, mirror within 8, mirror
// within 16 -- leave every lane of a row of 16 holding the row's result; the
// four rows' results are then read into scalar registers and combined.
// (Cross-lane permutes through the LDS cost a round trip per step.)
__device__ __forceinline__ u128 wave_sum_u128(u128 v)
{
    v += dpp_u128<0xB1>(v);
    v += dpp_u128<0x4E>(v);
    v += dpp_u128<0x141>(v);
    v += dpp_u128<0x140>(v);
    return readlane_u128(v, 0) + readlane_u128(v, 16) + readlane_u128(v, 32) + readlane_u128(v, 48);
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    v = max(v, dpp_u32<0xB1>(v));
    v = max(v, dpp_u32<0x4E>(v));
    v = max(v, dpp_u32<0x141>(v));
    v = max(v, dpp_u32<0x140>(v));
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0), b = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
    const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 32), d = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
    return max(max(a, b), max(c, d));
}
__device__ __forceinline__ unsigned wave_or_u32(unsigned v)
{
    v |= dpp_u32<0xB1>(v);
    v |= dpp_u32<0x4E>(v);
    v |= dpp_u32<0x141>(v);
    v |= dpp_u32<0x140>(v);
    return (unsigned)(__builtin_amdgcn_readlane((int)v, 0) | __builtin_amdgcn_readlane((int)v, 16) |
                      __builtin_amdgcn_readlane((int)v, 32) | __builtin_amdgcn_readlane((int)v, 48));
}

// LDS scratch of the end-of-pass reduction: one entry per wave.
struct SumsWave {
    u128 e, g;
    float mx;
    unsigned flags;
};

// The block's copy of accumulator slot `slot`.
__device__ __forceinline__ DT_GLOBAL Acc *block_acc(const Inst &I, int slot)
{
    return gptr(uni(I.acc + (long long)slot * ACC_COPIES + (blockIdx.x % ACC_COPIES)));
}

// One thread adds a block's reduced sums into its accumulator copy (relaxed,
// agent scope: the pass that reads them starts after this kernel has ended).
__device__ __forceinline__ void commit_block_sums(DT_GLOBAL Acc *a, u128 e, u128 g, float mx, unsigned flags)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned long long le = (unsigned long long)(unsigned)(e >> (32 * k));
        const unsigned long long lg = (unsigned long long)(unsigned)(g >> (32 * k));
        if (le) __hip_atomic_fetch_add(&a->e[k], le, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lg) __hip_atomic_fetch_add(&a->g[k], lg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (mx > 0.0f) __hip_atomic_fetch_max(&a->mx, __float_as_uint(mx), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (flags) __hip_atomic_fetch_or(&a->flags, flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Adds a block's sums, held in registers, into its accumulator copy: wave
// reductions, the waves' results in LDS, then wave 0 reduces those and one
// thread commits.  mx >= 0, so its bits order as unsigned integers.
template <int NT>
__device__ void add_pass_sums(const Inst &I, int slot, u128 e, u128 g, float mx, unsigned flags, SumsWave *red)
{
    constexpr int NW = NT / 64;
    static_assert(NW <= 64, "one wave's results per lane");
    DT_GLOBAL Acc *acc = block_acc(I, slot);
    e = wave_sum_u128(e);
    g = wave_sum_u128(g);
    const unsigned m = wave_max_u32(__float_as_uint(mx));
    flags = wave_or_u32(flags);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[w].e = e;
        red[w].g = g;
        red[w].mx = __uint_as_float(m);
        red[w].flags = flags;
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    u128 se = lane < NW ? red[lane].e : (u128)0;
    u128 sg = lane < NW ? red[lane].g : (u128)0;
    unsigned sm = lane < NW ? __float_as_uint(red[lane].mx) : 0u;
    unsigned sf = lane < NW ? red[lane].flags : 0u;
    se = wave_sum_u128(se);
    sg = wave_sum_u128(sg);
    sm = wave_max_u32(sm);
    sf = wave_or_u32(sf);
    if (lane == 0) commit_block_sums(acc, se, sg, __uint_as_float(sm), sf);
}

// Adds a block's sums into its accumulator copy when the exact sums sit in
// LDS, one pair per tile thread (accE/accG[0 .. NOWN)): every wave leaves its
// max and flags in `red`, then after one barrier wave 0 alone sums the LDS
// pairs (NOWN / 64 per lane), reduces across its lanes and commits; the
// other waves are done at the barrier.  The sums are integers, so the order of
// the additions does not change them.
template <int NT, int NOWN>
__device__ void add_pass_sums_lds(const Inst &I, int slot, const u128 *accE, const u128 *accG, float mx,
                                  unsigned flags, SumsWave *red)
{
    constexpr int NW = NT / 64;
    static_assert(NW <= 64, "one wave's results per lane");
    const unsigned m = wave_max_u32(__float_as_uint(mx));
    flags = wave_or_u32(flags);
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        red[threadIdx.x >> 6].mx = __uint_as_float(m);
        red[threadIdx.x >> 6].flags = flags;
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    DT_GLOBAL Acc *acc = block_acc(I, slot);
    u128 e = 0, g = 0;
#pragma unroll 4
    for (int i = lane; i < NOWN; i += 64) {
        e += accE[i];
        g += accG[i];
    }
    unsigned sm = lane < NW ? __float_as_uint(red[lane].mx) : 0u;
    unsigned sf = lane < NW ? red[lane].flags : 0u;
    e = wave_sum_u128(e);
    g = wave_sum_u128(g);
    sm = wave_max_u32(sm);
    sf = wave_or_u32(sf);
    if (lane == 0) commit_block_sums(acc, e, g, __uint_as_float(sm), sf);
}

// 128-bit fixed-point value (quantum 2^-64) as a double; deterministic.
__device__ __forceinline__ double fix_to_double(u128 v)
{
    const unsigned long long hi = (unsigned long long)(v >> 64), lo = (unsigned long long)v;
    return (double)hi + (double)lo * 0x1p-64;
}

// The sums of one accumulator slot, from its ACC_COPIES copies as 64-bit
// words (the layout of Acc: e[4], g[4], mx | flags << 32).
struct PassSums {
    double E, G;
    float m;
};

constexpr int ACC_WORDS = (int)(sizeof(Acc) / 8);
static_assert(ACC_WORDS == 9, "Acc is nine 64-bit words");

// One fold step of the accumulator records held one per lane: every lane adds
// the words of the lane CTRL selects (limb sums add, max and flags combine).
template <int CTRL>
__device__ __forceinline__ void fold_records(unsigned long long (&w)[8], unsigned &mx, unsigned &fl)
{
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] += dpp_u64<CTRL>(w[k]);
    const unsigned m = dpp_u32<CTRL>(mx);
    mx = m > mx ? m : mx;
    fl |= dpp_u32<CTRL>(fl);
}

// The sums of accumulator slot `slot` when lane L < ACC_SLOTS * ACC_COPIES
// holds record L (slot L / ACC_COPIES, copy L % ACC_COPIES): limb words w,
// max bits mx and flags fl (run by a whole wave, `slot` uniform).  The eight
// copies of a slot sit in consecutive lanes 8s .. 8s + 7 and are folded by
// three row-local DPP steps (lane xor 1, lane xor 2, mirror within 8); the
// slot's total is then read from lane 8s.  The words are sums of 32-bit limbs
// (below 2^45 here), so the folds are exact and their order does not matter.
__device__ inline PassSums sums_of_records(unsigned long long (&w)[8], unsigned mx, unsigned fl, int slot)
{
    static_assert(ACC_COPIES == 8, "a slot's copies fill one half-row");
    fold_records<0xB1>(w, mx, fl);    // quad_perm [1, 0, 3, 2]
    fold_records<0x4E>(w, mx, fl);    // quad_perm [2, 3, 0, 1]
    fold_records<0x141>(w, mx, fl);   // row_half_mirror
    const int src = slot * ACC_COPIES;
    u128 e = 0, g = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        e += (u128)readlane_u64(w[l], src) << (32 * l);
        g += (u128)readlane_u64(w[4 + l], src) << (32 * l);
    }
    const unsigned mxb = (unsigned)__builtin_amdgcn_readlane((int)mx, src);
    const unsigned flags = (unsigned)__builtin_amdgcn_readlane((int)fl, src);
    PassSums s;
    s.E = (flags & ACC_E_NAN) ? (double)NAN : ((flags & ACC_E_HUGE) ? (double)INFINITY : fix_to_double(e));
    s.G = (flags & ACC_G_NAN) ? (double)NAN : ((flags & ACC_G_HUGE) ? (double)INFINITY : fix_to_double(g));
    s.m = __uint_as_float(mxb);
    return s;
}

// ---------------------------------------------------------------------------
// Controllers (run at the start of the next pass, on a copy of the state)
// ---------------------------------------------------------------------------

// What the controllers read of an instance besides its control block: the
// solver parameters, the trace and the host-coherent status words.  The
// controller step gathers it with the control block in one round trip (a
// field read where a branch first needs it costs a round trip of its own).
struct CtlEnv {
    dt_params p;
    double *trace;                  // Inst::trace .. Inst::final_ctrl, in order
    long long trace_cap;
    unsigned long long *status;
    Ctrl *final_ctrl;
};

// A log line (it, E, max error, eta); only the instance's first tile writes
// the trace (`leader`), every block counts it.
__device__ __forceinline__ void log_line(const CtlEnv &V, Ctrl *C, bool leader, double it, double E, double m,
                                         double eta)
{
    if (leader && V.trace && C->trace_len < V.trace_cap) {
        double *t = V.trace + 4 * C->trace_len;
        t[0] = it; t[1] = E; t[2] = m; t[3] = eta;
    }
    C->trace_len++;
}

// Completion word of a solve in host-coherent memory: (evaluated passes << 1)
// | done, stored by the instance's first tile after every controller step, so
// the host follows the solve without a copy or an event between pass
// launches.  Once the solve is done the control block itself is published
// first (when the instance has a host copy), ordered before the word by a
// system-scope release.
__device__ __forceinline__ void publish_status(const CtlEnv &V, const Ctrl *C)
{
    if (!V.status) return;
    const bool done = C->phase == PH_DONE;
    if (done && V.final_ctrl) {
        const unsigned long long *src = (const unsigned long long *)C;
        unsigned long long *dst = (unsigned long long *)V.final_ctrl;
        for (int k = 0; k < (int)(sizeof(Ctrl) / 8); ++k)
            __hip_atomic_store(dst + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __atomic_thread_fence(__ATOMIC_RELEASE);
    }
    __hip_atomic_store(V.status, ((unsigned long long)C->evaluations << 1) | (done ? 1ull : 0ull), __ATOMIC_RELAXED,
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
static __device__ __forceinline__ void grad_controller(const CtlEnv &V, Ctrl *C, bool leader, double E, float m,
                                                      double G)
{
    const dt_params &p = V.p;
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
        log_line(V, C, leader, 0.0, E, m, C->eta);
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
        log_line(V, C, leader, (double)C->it, E, m, C->eta);
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
// `step_io` is the field volumes a pass with the step moves (2: the base read,
// the trial written); the evaluation of the initial guess reads one.
static __device__ __forceinline__ void greedy_controller(const CtlEnv &V, Ctrl *C, bool leader, double E, float m,
                                                        int step_io)
{
    const dt_params &p = V.p;
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
    log_line(V, C, leader, (double)C->it, E, m, C->eta);
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
// The start of a pass
// ---------------------------------------------------------------------------

// The controller step (pass_begin_load, then pass_begin_run), run by wave 0
// of every block before the block's work: reads the control
// block of this pass's parity and, when it records an evaluated trial,
// advances the controller (ALG: DT_GRADIENT or DT_GREEDY) by the sums of that
// pass, then leaves the state in `c` (the block's LDS copy).  Every block
// computes the same state from the same inputs; the instance's first tile
// (`leader`) stores it into the other copy for the next pass, clears the
// accumulator slot the next pass adds into and publishes the completion
// word.  On return `c` is the state this pass works from (active == 0: the
// solve has ended, the pass does nothing); the caller synchronises the block
// before reading it.
//
// The lanes load the control words and every slot's accumulator records at
// once (which slot holds the previous pass's sums is in the control block
// itself), so the whole step costs one memory round trip.  The control
// block is gathered into wave-uniform registers (one word per lane, read
// across lanes) and the controller runs there: a controller on an LDS copy
// waits for one LDS round trip per field it touches.
// What wave 0 loads for the controller step: its lane's control or CtlEnv
// word and accumulator record, and the instance's control copies and
// accumulators (kept in registers for the leader's stores: the compiler
// would otherwise reload them from the instance record there, a round trip
// of its own).
struct BeginLoads {
    unsigned long long cw, aw[8], a8;
    Ctrl *ctrl;
    Acc *acc;
};

// The loads of the controller step (pass_begin_run consumes them); a kernel
// may issue other loads in between, whose round trip then overlaps this one.
__device__ __forceinline__ BeginLoads pass_begin_load(const Inst &I, int parity)
{
    constexpr int W = (int)(sizeof(Ctrl) / 8);
    constexpr int PW = (int)(sizeof(dt_params) / 8), VW = (int)(sizeof(CtlEnv) / 8);
    constexpr int NREC = ACC_SLOTS * ACC_COPIES;
    static_assert(W + VW <= 64 && NREC <= 64, "one word and one accumulator record per lane");
    static_assert(sizeof(dt_params) % 8 == 0 && offsetof(CtlEnv, trace) == sizeof(dt_params), "CtlEnv layout");
    static_assert(offsetof(Inst, final_ctrl) - offsetof(Inst, trace) == 24, "Inst::trace .. final_ctrl");
    const int lane = threadIdx.x & 63;
    BeginLoads L;
    L.ctrl = uni(I.ctrl);
    L.acc = uni(I.acc);
    const unsigned long long *src = (const unsigned long long *)(L.ctrl + parity);
    const unsigned long long *rec = (const unsigned long long *)(L.acc + (lane < NREC ? lane : 0));
    // lane k < W: word k of the control block; W <= k < W + VW: word k - W
    // of the CtlEnv (the parameters, then Inst::trace .. Inst::final_ctrl)
    // (every lane loads from a valid address, so the loads need no branches)
    const unsigned long long *wp = lane < W ? src + lane
                                 : lane < W + PW ? (const unsigned long long *)&I.p + (lane - W)
                                 : lane < W + VW ? (const unsigned long long *)&I.trace + (lane - W - PW)
                                                 : src;
    L.cw = *wp;
#pragma unroll
    for (int k = 0; k < 8; ++k) L.aw[k] = rec[k];
    L.a8 = rec[8];
    return L;
}

template <int ALG>
__device__ void pass_begin_run(const Inst &I, int parity, bool leader, Ctrl *c, const BeginLoads &L)
{
    constexpr int W = (int)(sizeof(Ctrl) / 8);
    constexpr int VW = (int)(sizeof(CtlEnv) / 8);
    constexpr int NREC = ACC_SLOTS * ACC_COPIES;
    const int lane = threadIdx.x & 63;
    Ctrl *ctrl = L.ctrl;
    Acc *acc = L.acc;
    const unsigned long long cw = L.cw;
    unsigned long long aw[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) aw[k] = L.aw[k];
    const unsigned long long a8 = L.a8;
    const bool has_rec = lane < NREC;
#pragma unroll
    for (int k = 0; k < 8; ++k) aw[k] = has_rec ? aw[k] : 0ull;
    const unsigned amx = has_rec ? (unsigned)a8 : 0u;
    const unsigned afl = has_rec ? (unsigned)(a8 >> 32) : 0u;
    unsigned long long words[W];
#pragma unroll
    for (int k = 0; k < W; ++k) words[k] = readlane_u64(cw, k);
    Ctrl s;
    __builtin_memcpy(&s, words, sizeof(Ctrl));
    unsigned long long vwords[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) vwords[k] = readlane_u64(cw, W + k);
    CtlEnv V;
    __builtin_memcpy(&V, vwords, sizeof(CtlEnv));
    if (s.active && s.evaluated) {
        const PassSums ps = sums_of_records(aw, amx, afl, s.acc_cur);
        // (only lane 0 of the leader writes log lines)
        if (ALG == DT_GRADIENT) grad_controller(V, &s, leader && lane == 0, ps.E, ps.m, ps.G);
        else greedy_controller(V, &s, leader && lane == 0, ps.E, ps.m, 2);
        const int prev = s.acc_prev;
        s.acc_prev = s.acc_cur;
        s.acc_cur = s.acc_zero;
        s.acc_zero = prev;
    }
    s.evaluated = s.active;
    // the pass in which the solve ends copies its result into buf[0]
    // (once: the copy of the next pass records it as done)
    s.copy_now = (!s.active && s.phase == PH_DONE && !s.result_copied) ? 1 : 0;
    if (s.copy_now) s.result_copied = 1;
    if (lane == 0) *c = s;
    unsigned long long *dst = (unsigned long long *)c;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (leader) {
        // an ended solve stays ended at both parities
        unsigned long long *next = (unsigned long long *)(ctrl + (parity ^ 1));
        if (lane < W) next[lane] = dst[lane];
        if (s.active) {
            unsigned long long *z = (unsigned long long *)(acc + (long long)s.acc_zero * ACC_COPIES);
            for (int i = lane; i < ACC_COPIES * ACC_WORDS; i += 64) z[i] = 0ull;
        }
        if (lane == 0) publish_status(V, &s);
    }
}

template <int ALG>
__device__ __forceinline__ void pass_begin(const Inst &I, int parity, bool leader, Ctrl *c)
{
    const BeginLoads L = pass_begin_load(I, parity);
    pass_begin_run<ALG>(I, parity, leader, c, L);
}

// ---------------------------------------------------------------------------
// Tile decoding
// ---------------------------------------------------------------------------

struct TileCoord {
    int x0, y0, z0, z1;   // first column, first plane, one past the last plane
};

template <int TX, int TY>
__device__ __forceinline__ TileCoord decode_tile(const Inst &I, const Tile &t)
{
    TileCoord c;
    c.x0 = t.tile % I.ntx * TX;
    c.y0 = t.tile / I.ntx * TY;
    c.z0 = t.z0;
    c.z1 = t.z1;
    return c;
}

// The instance's first tile: the one that stores the next control block.
__device__ __forceinline__ bool leader_tile(const Tile &t) { return t.tile == 0 && t.z0 == 0; }

// Run by every block of the pass in which a solve ends (copy_now): the
// block's tile of the result buffer the controller names is copied into the
// caller's field, buf[0] (the reference returns field_[old],
// src/displacement_field_gradient.c:421).  The tiles of a pass partition the
// volume, so the whole result moves without a launch of its own.  Nothing to
// do when the result is already in buf[0].
template <int TX, int TY, int NT>
__device__ void copy_result_tile(const Inst &I, const Ctrl *C, const TileCoord &tc)
{
    const int rb = __builtin_amdgcn_readfirstlane(C->result_buf);
    if (rb <= 0 || rb >= NBUF) return;
    const float *__restrict__ src = I.buf[rb];
    float *__restrict__ dst = I.buf[0];
    const long long plane = (long long)I.nx * I.ny;
    // a thread copies its column eight planes at a time: the 24 loads of a
    // batch are all in flight before its stores
    constexpr int B = 8;
    for (int t = threadIdx.x; t < TX * TY; t += NT) {
        const int x = tc.x0 + t % TX, y = tc.y0 + t / TX;
        if (x >= I.nx || y >= I.ny) continue;
        const long long col = (long long)y * I.nx + x;
        for (int z0 = tc.z0; z0 < tc.z1; z0 += B) {
            float v[3][B];
#pragma unroll
            for (int k = 0; k < B; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (z0 + k < tc.z1) v[c][k] = src[c * I.n + (long long)(z0 + k) * plane + col];
#pragma unroll
            for (int k = 0; k < B; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (z0 + k < tc.z1) dst[c * I.n + (long long)(z0 + k) * plane + col] = v[c][k];
        }
    }
}

}  // namespace dt
