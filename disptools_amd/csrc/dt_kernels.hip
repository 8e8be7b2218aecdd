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

#include "dt_internal.hpp"
#include "dt_stencil.hpp"
#include "dt_kernels.hpp"

namespace dt {

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
__device__ void grad_controller(const Inst &I, Ctrl *C, double E, float m, double G)
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
__device__ void greedy_controller(const Inst &I, Ctrl *C, double E, float m)
{
    const dt_params &p = I.p;
    C->evaluations++;
    // step: read u, write u; evaluation: read u (error maps are counted by the caller)
    C->field_io += C->phase == PH_INIT ? 1 : 3;
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

// ---------------------------------------------------------------------------
// Fused gradient pass
// ---------------------------------------------------------------------------
//
// A block owns a TX x TY column tile and streams its planes z0..z1-1.  Step zz
// (a) stores the prefetched trial plane zz+2 into a ring of five LDS planes
// with a 3-voxel halo, (b) evaluates, on plane zz over the tile plus a 1-voxel
// ring, the Jacobian J, the voxel error e, the penalised weight
// w = e + zeta (J - delta)[J < delta] and the weighted cofactors
// W_ab = w * M_ab, where M_ab are the 2x2 minors of I + D (D = 2nd-order
// central derivatives) that the adjoint stencil needs, and (c) forms the
// gradient on plane zz-1 as the difference of W across each axis:
//
//   g_x = +ix dW_xx - iy dW_xy + iz dW_xz      dW_ab = W_ab(p - e_b) - W_ab(p + e_b)
//   g_y = -ix dW_yx + iy dW_yy - iz dW_yz
//   g_z = +ix dW_zx - iy dW_zy + iz dW_zz
//
// which is the gradient of displacement_field_gradient.c:26-145 reorganised
// so that each minor is computed once per voxel instead of once per neighbour.
// The pass then writes the speculative candidate u - eta_next g.

template <typename CT, int TX, int TY>
struct GradPassSmem {
    static constexpr int HX = TX + 6, HY = TY + 6, PL = HX * HY;
    static constexpr int WX = TX + 2, WY = TY + 2, WPL = WX * WY;
    float xs[5][3][PL];      // trial planes, halo 3
    float ws[2][6][WPL];     // W_a,x (0..2) and W_a,y (3..5) of two planes
    double red[3 * (TX * TY / 64)];
    int flag;
};

__device__ __forceinline__ int ring(int z) { return (z + 10) % 5; }

template <typename CT, int TX, int TY, bool FAST>
__device__ __forceinline__ void stage_point(
    const Inst &I, GradPassSmem<CT, TX, TY> &S, int zz, int lx, int ly, int gx, int gy,
    int what,  // 0 = tile point (all W), 1 = x-halo (W_a,x), 2 = y-halo (W_a,y)
    int wbuf, float wz[3], bool count, double &Esum, float &mx)
{
    using SM = GradPassSmem<CT, TX, TY>;
    constexpr int HX = SM::HX;
    const int c0i = (ly + 3) * HX + (lx + 3);
    const int s0 = ring(zz);
    const int sm2 = ring(zz - 2), sm1 = ring(zz - 1), sp1 = ring(zz + 1), sp2 = ring(zz + 2);
    const float inv[3] = {I.ix, I.iy, I.iz};

    int kind[3] = {K_C4, K_C4, K_C4};
    if (!FAST) classify(gx, gy, zz, I.nx, I.ny, I.nz, kind);

    CT d4[3][3];
    CT d2[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *P = S.xs[s0][c];
        const float v0 = P[c0i];
        // x
        {
            const float m2 = P[c0i - 2], m1 = P[c0i - 1], p1 = P[c0i + 1], p2 = P[c0i + 2];
            d4[c][0] = FAST ? dstencil_c4<CT>(m2, m1, p1, p2, inv[0])
                            : dstencil<CT>(kind[0], m2, m1, v0, p1, p2, inv[0]);
            d2[c][0] = (CT)((p1 - m1) * inv[0]) * (CT)0.5;
        }
        // y
        {
            const float m2 = P[c0i - 2 * HX], m1 = P[c0i - HX], p1 = P[c0i + HX], p2 = P[c0i + 2 * HX];
            d4[c][1] = FAST ? dstencil_c4<CT>(m2, m1, p1, p2, inv[1])
                            : dstencil<CT>(kind[1], m2, m1, v0, p1, p2, inv[1]);
            d2[c][1] = (CT)((p1 - m1) * inv[1]) * (CT)0.5;
        }
        // z
        {
            const float m2 = S.xs[sm2][c][c0i], m1 = S.xs[sm1][c][c0i];
            const float p1 = S.xs[sp1][c][c0i], p2 = S.xs[sp2][c][c0i];
            d4[c][2] = FAST ? dstencil_c4<CT>(m2, m1, p1, p2, inv[2])
                            : dstencil<CT>(kind[2], m2, m1, v0, p1, p2, inv[2]);
            d2[c][2] = (CT)((p1 - m1) * inv[2]) * (CT)0.5;
        }
    }
    d2[0][0] += (CT)1.0;
    d2[1][1] += (CT)1.0;
    d2[2][2] += (CT)1.0;

    const float J = (float)det_plus_identity<CT>(d4);
    const long long vi = ((long long)zz * I.ny + gy) * I.nx + gx;
    const float T = I.target[vi];
    const bool in_mask = I.mask[vi] != 0;
    const float e = voxel_error(J, T, in_mask, I.p.tolerance);
    if (count) {
        Esum += (double)e * (double)e;
        if (in_mask) mx = fmaxf(mx, fabsf(e));
    }
    float w = e;
    if (J < I.p.delta) w += I.p.zeta * (J - I.p.delta);
    const CT cw = (CT)w;

    const int wi = (ly + 1) * SM::WX + (lx + 1);
    if (what != 2) {
        // relation axis x: minors from derivatives along y and z
        S.ws[wbuf][0][wi] = (float)(cw * (d2[1][1] * d2[2][2] - d2[1][2] * d2[2][1]));
        S.ws[wbuf][1][wi] = (float)(cw * (d2[0][1] * d2[2][2] - d2[0][2] * d2[2][1]));
        S.ws[wbuf][2][wi] = (float)(cw * (d2[0][1] * d2[1][2] - d2[0][2] * d2[1][1]));
    }
    if (what != 1) {
        // relation axis y: minors from derivatives along x and z
        S.ws[wbuf][3][wi] = (float)(cw * (d2[1][0] * d2[2][2] - d2[1][2] * d2[2][0]));
        S.ws[wbuf][4][wi] = (float)(cw * (d2[0][0] * d2[2][2] - d2[0][2] * d2[2][0]));
        S.ws[wbuf][5][wi] = (float)(cw * (d2[0][0] * d2[1][2] - d2[0][2] * d2[1][0]));
    }
    if (what == 0) {
        // relation axis z: minors from derivatives along x and y (kept in registers)
        wz[0] = (float)(cw * (d2[1][0] * d2[2][1] - d2[1][1] * d2[2][0]));
        wz[1] = (float)(cw * (d2[0][0] * d2[2][1] - d2[0][1] * d2[2][0]));
        wz[2] = (float)(cw * (d2[0][0] * d2[1][1] - d2[0][1] * d2[1][0]));
    }
}

template <typename CT, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
k_grad_pass(const Inst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    constexpr int NT = TX * TY;
    using SM = GradPassSmem<CT, TX, TY>;
    constexpr int HX = SM::HX, PL = SM::PL;
    __shared__ SM S;

    const Tile tl = tiles[blockIdx.x];
    const Inst &I = insts[tl.inst];
    Ctrl *C = I.ctrl;
    if (!C->active) return;

    const int mode = C->mode;
    const float *__restrict__ src = I.buf[C->src];
    const float *__restrict__ cnd = I.buf[mode == TRIAL_INTERP ? C->cand : C->src];
    const float s = C->s;
    const int out_trial = C->out_trial;
    const int out_next = C->out_next;
    const float eta_next = C->eta_next;

    const TileCoord tc = decode_tile<TX, TY>(I, tl.tile);
    const int nx = I.nx, ny = I.ny, nz = I.nz;
    const long long n = I.n;
    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int gx = tc.x0 + lx, gy = tc.y0 + ly;
    const bool col_in = gx < nx && gy < ny;

    // whole tile (+ 1-voxel ring, all planes) uses the 4th-order interior stencil?
    const bool fast_xy = tc.x0 - 1 >= 2 && tc.x0 + TX <= nx - 3 && tc.y0 - 1 >= 2 && tc.y0 + TY <= ny - 3;

    // trial plane loader: ring slot <- plane gz (zero outside the volume)
    constexpr int PER = (3 * PL + NT - 1) / NT;
    float pre[PER];
    auto fetch = [&](int gz) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * NT;
            float v = 0.0f;
            if (i < 3 * PL) {
                const int c = i / PL, r = i - c * PL;
                const int yy = r / HX, xx = r - yy * HX;
                const int x = tc.x0 - 3 + xx, y = tc.y0 - 3 + yy;
                if (gz >= 0 && gz < nz && x >= 0 && x < nx && y >= 0 && y < ny) {
                    const long long idx = c * n + ((long long)gz * ny + y) * nx + x;
                    if (mode == TRIAL_DIRECT) {
                        v = src[idx];
                    } else {
                        const float b = src[idx];
                        v = b + s * (cnd[idx] - b);
                    }
                }
            }
            pre[k] = v;
        }
    };
    auto commit = [&](int gz) {
        float *dst = &S.xs[ring(gz)][0][0];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * NT;
            if (i < 3 * PL) dst[i] = pre[k];
        }
    };

    const int z0 = tc.z0, z1 = tc.z1;
    // prologue: planes z0-3 .. z0 in LDS, plane z0+1 in registers
    for (int gz = z0 - 3; gz <= z0; ++gz) {
        fetch(gz);
        commit(gz);
    }
    fetch(z0 + 1);

    double Esum = 0.0, Gsum = 0.0;
    float mx = 0.0f;
    float wz_m[3] = {0, 0, 0}, wz_c[3] = {0, 0, 0}, wz_p[3] = {0, 0, 0};
    const int nhalo = 2 * TY + 2 * TX;

    for (int zz = z0 - 1; zz <= z1; ++zz) {
        commit(zz + 2);
        __syncthreads();
        fetch(zz + 3);

        // ---- stage A: plane zz ----
#pragma unroll
        for (int k = 0; k < 3; ++k) { wz_m[k] = wz_c[k]; wz_c[k] = wz_p[k]; wz_p[k] = 0.0f; }
        const int wbuf = zz & 1;
        if (zz >= 0 && zz < nz) {
            const bool fast = fast_xy && zz >= 2 && zz <= nz - 3;
            const bool count = zz >= z0 && zz < z1;
            if (col_in) {
                if (fast) stage_point<CT, TX, TY, true>(I, S, zz, lx, ly, gx, gy, 0, wbuf, wz_p, count, Esum, mx);
                else stage_point<CT, TX, TY, false>(I, S, zz, lx, ly, gx, gy, 0, wbuf, wz_p, count, Esum, mx);
            }
            for (int h = tid; h < nhalo; h += NT) {
                int hx, hy, what;
                if (h < TY) { hx = -1; hy = h; what = 1; }
                else if (h < 2 * TY) { hx = TX; hy = h - TY; what = 1; }
                else if (h < 2 * TY + TX) { hx = h - 2 * TY; hy = -1; what = 2; }
                else { hx = h - 2 * TY - TX; hy = TY; what = 2; }
                const int qx = tc.x0 + hx, qy = tc.y0 + hy;
                if (qx >= 0 && qx < nx && qy >= 0 && qy < ny) {
                    float dummy[3];
                    double e0 = 0.0;
                    float m0 = 0.0f;
                    if (fast) stage_point<CT, TX, TY, true>(I, S, zz, hx, hy, qx, qy, what, wbuf, dummy, false, e0, m0);
                    else stage_point<CT, TX, TY, false>(I, S, zz, hx, hy, qx, qy, what, wbuf, dummy, false, e0, m0);
                }
            }
        }
        __syncthreads();

        // ---- stage B: plane zb = zz - 1 ----
        const int zb = zz - 1;
        if (zb >= z0 && zb < z1 && col_in) {
            const float *P = S.xs[ring(zb)][0];
            const int c0i = (ly + 3) * HX + (lx + 3);
            const float ux = P[c0i], uy = P[PL + c0i], uz = P[2 * PL + c0i];
            float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
            if (gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2 && zb >= 1 && zb <= nz - 2) {
                const int wb = zb & 1;
                const int wi = (ly + 1) * SM::WX + (lx + 1);
                const float *W = &S.ws[wb][0][0];
                constexpr int WPL = SM::WPL, WX = SM::WX;
                const CT ix = I.ix, iy = I.iy, iz = I.iz;
                const CT dxx = (CT)W[0 * WPL + wi - 1] - (CT)W[0 * WPL + wi + 1];
                const CT dyx = (CT)W[1 * WPL + wi - 1] - (CT)W[1 * WPL + wi + 1];
                const CT dzx = (CT)W[2 * WPL + wi - 1] - (CT)W[2 * WPL + wi + 1];
                const CT dxy = (CT)W[3 * WPL + wi - WX] - (CT)W[3 * WPL + wi + WX];
                const CT dyy = (CT)W[4 * WPL + wi - WX] - (CT)W[4 * WPL + wi + WX];
                const CT dzy = (CT)W[5 * WPL + wi - WX] - (CT)W[5 * WPL + wi + WX];
                const CT dxz = (CT)wz_m[0] - (CT)wz_p[0];
                const CT dyz = (CT)wz_m[1] - (CT)wz_p[1];
                const CT dzz = (CT)wz_m[2] - (CT)wz_p[2];
                g0 = (float)(ix * dxx - iy * dxy + iz * dxz);
                g1 = (float)(-ix * dyx + iy * dyy - iz * dyz);
                g2 = (float)(ix * dzx - iy * dzy + iz * dzz);
                Gsum += (double)g0 * g0 + (double)g1 * g1 + (double)g2 * g2;
            }
            const long long vi = ((long long)zb * ny + gy) * nx + gx;
            if (out_next >= 0) {
                float *o = I.buf[out_next];
                o[vi] = ux - eta_next * g0;
                o[n + vi] = uy - eta_next * g1;
                o[2 * n + vi] = uz - eta_next * g2;
            }
            if (out_trial >= 0) {
                float *o = I.buf[out_trial];
                o[vi] = ux;
                o[n + vi] = uy;
                o[2 * n + vi] = uz;
            }
        }
    }

    block_reduce3<NT>(Esum, mx, Gsum, S.red);
    if (arrive<NT>(I, tl.tile, Esum, mx, Gsum, &S.flag)) {
        double E, G;
        float m;
        fold_partials<NT>(I, E, m, G, S.red);
        if (tid == 0) {
            grad_controller(I, C, E, m, G);
            __hip_atomic_store(I.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

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
            greedy_controller(I, C, E, m);
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

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

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

hipError_t launch_grad_pass(int precision, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    if (precision == 64)
        hipLaunchKernelGGL((k_grad_pass<double, TILE_X, TILE_Y>), dim3(nblocks), dim3(TILE_X * TILE_Y), 0, st,
                           insts, tiles);
    else
        hipLaunchKernelGGL((k_grad_pass<float, TILE_X, TILE_Y>), dim3(nblocks), dim3(TILE_X * TILE_Y), 0, st,
                           insts, tiles);
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
