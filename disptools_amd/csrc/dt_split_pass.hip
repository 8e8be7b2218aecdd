// dt_split_pass.hip -- the gradient-solver pass as two kernels per trial.
//
//   k_eval    trial field -> Jacobian J (region map of src/jacobian.c:70-242),
//             voxel error e (src/headers/error.h:128-142), energy and masked max
//             partials, and the penalised weight w = e + zeta (J - delta)[J < delta]
//             (displacement_field_gradient.c:71-88), stored as one float per voxel.
//   k_update  w and the trial field -> weighted cofactors W_ab = w M_ab of the
//             2nd-order derivatives (gradient.c:13-21), the gradient as the
//             difference of W across each axis (gradient.c:26-145 reorganised so
//             each minor is computed once per voxel), the speculative candidate
//             u - eta_next g (gradient.c:150-175), and, in the last block, the
//             line-search controller (gradient.c:250-421).
//
// Compared with the single fused pass this moves one float per voxel through
// memory (w) but removes the recomputation of the full Jacobian on the ring of
// every tile: the ring of k_update only needs central differences.  Both
// kernels keep fewer values live per thread, so more waves fit on a CU.
//
// Layout of a block: TX x TY columns, one thread each, streaming planes
// z0..z1-1 of its tile; the same tile table and partial-sum slots serve both
// kernels (slot 4t+0: sum e^2, 4t+1: max |e| in the mask, 4t+2: sum |g|^2).
#include "dt_device.hpp"
#include "dt_kernels.hpp"
#include "dt_pass_common.hpp"

namespace dt {

// ---------------------------------------------------------------------------
// Shared tile bookkeeping
// ---------------------------------------------------------------------------

// Per-thread byte offsets of the load slots of one halo-H plane of the field.
template <int TX, int TY, int H, int PER>
__device__ __forceinline__ void plane_slots(int x0, int y0, int nx, int ny, unsigned n, unsigned (&vo)[PER],
                                            unsigned &ok)
{
    constexpr int HX = TX + 2 * H, PL = HX * (TY + 2 * H), NT = TX * TY;
    ok = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + k * NT;
        const int c = i / PL, r = i - c * PL;
        const int yy = r / HX, xx = r - yy * HX;
        const int x = x0 - H + xx, y = y0 - H + yy;
        const bool v = i < 3 * PL && x >= 0 && x < nx && y >= 0 && y < ny;
        vo[k] = v ? 4u * ((unsigned)c * n + (unsigned)(y * nx + x)) : 0u;
        ok |= (v ? 1u : 0u) << k;
    }
}

// Trial-field plane loader: raw loads into registers, combined (INTERP) and
// masked when the plane is committed to LDS.
template <int TX, int TY, int H, bool INTERP>
struct PlaneLoader {
    static constexpr int HX = TX + 2 * H, PL = HX * (TY + 2 * H), NT = TX * TY;
    static constexpr int PER = (3 * PL + NT - 1) / NT;
    __amdgpu_buffer_rsrc_t r_src, r_cnd;
    unsigned vo[PER];
    unsigned ok;
    unsigned plane4;
    int nz;
    float s;

    __device__ __forceinline__ void fetch(int gz, float (&a)[PER], float (&b)[PER]) const
    {
        const int gzc = min(max(gz, 0), nz - 1);
        const unsigned so = plane4 * (unsigned)gzc;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            a[k] = bload(r_src, vo[k], so);
            if (INTERP) b[k] = bload(r_cnd, vo[k], so);
        }
    }
    __device__ __forceinline__ void commit(int gz, float *dst, const float (&a)[PER], const float (&b)[PER]) const
    {
        const bool zin = gz >= 0 && gz < nz;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + k * NT;
            float v = a[k];
            if (INTERP) v = v + s * (b[k] - v);
            v = (zin && ((ok >> k) & 1u)) ? v : 0.0f;
            if (i < 3 * PL) dst[i] = v;
        }
    }
};

struct PassBuffers {
    __amdgpu_buffer_rsrc_t r_src, r_cnd, r_next, r_trial, r_w, r_tgt, r_msk;
    bool has_next, has_trial;
    float s, eta_next;
    unsigned n, plane;
};

template <bool INTERP>
__device__ __forceinline__ PassBuffers pass_buffers(const Inst &I, const Ctrl *C)
{
    PassBuffers B;
    B.n = uni((unsigned)I.n);
    B.plane = (unsigned)uni(I.nx) * (unsigned)uni(I.ny);
    const unsigned fbytes = 12u * B.n;
    const int b_src = uni(C->src), b_cnd = uni(INTERP ? C->cand : C->src);
    const int b_next = uni(C->out_next), b_trial = uni(C->out_trial);
    B.has_next = b_next >= 0;
    B.has_trial = INTERP && b_trial >= 0;
    float *p_src = uni(I.buf[b_src]);
    B.r_src = make_rsrc(p_src, fbytes);
    B.r_cnd = make_rsrc(uni(I.buf[b_cnd]), fbytes);
    B.r_next = make_rsrc(B.has_next ? uni(I.buf[b_next]) : p_src, fbytes);
    B.r_trial = make_rsrc(B.has_trial ? uni(I.buf[b_trial]) : p_src, fbytes);
    B.r_w = make_rsrc(uni(I.wmap), 4u * B.n);
    B.r_tgt = make_rsrc(uni(I.target), 4u * B.n);
    B.r_msk = make_rsrc(uni(I.mask), B.n);
    B.s = uni(C->s);
    B.eta_next = uni(C->eta_next);
    return B;
}

__device__ __forceinline__ int mod6(int z) { return (z + 12) % 6; }
__device__ __forceinline__ int mod4(int z) { return (z + 12) & 3; }

// ---------------------------------------------------------------------------
// k_eval: J, e, w
// ---------------------------------------------------------------------------

template <int TX, int TY>
struct EvalSmem {
    static constexpr int HX = TX + 4, PL = HX * (TY + 4);
    float xs[6][3][PL];   // ring of six trial planes, halo 2
    double red[3 * (TX * TY / 64)];
};

template <typename CT, int TX, int TY, bool INTERP>
__device__ __forceinline__ void eval_body(EvalSmem<TX, TY> &S, const Inst &I, const Ctrl *C, const TileCoord tc,
                                          double &Esum, float &mx)
{
    using SM = EvalSmem<TX, TY>;
    constexpr int HX = SM::HX, PL = SM::PL;
    using LD = PlaneLoader<TX, TY, 2, INTERP>;
    constexpr int PER = LD::PER;
    const PassParams P = load_params(I);
    const PassBuffers B = pass_buffers<INTERP>(I, C);
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int z0 = tc.z0, z1 = tc.z1;

    LD L;
    L.r_src = B.r_src; L.r_cnd = B.r_cnd; L.plane4 = 4u * B.plane; L.nz = nz; L.s = B.s;
    plane_slots<TX, TY, 2, PER>(tc.x0, tc.y0, nx, ny, B.n, L.vo, L.ok);

    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int gx = tc.x0 + lx, gy = tc.y0 + ly;
    const bool col_in = gx < nx && gy < ny;
    const unsigned col_off = col_in ? (unsigned)(gy * nx + gx) : 0u;
    const int c0i = (ly + 2) * HX + (lx + 2);
    // w is only consumed by the update kernel of a pass that forms a candidate
    const bool store_w = B.has_next;

    auto own = [&](int gz, float (&v)[3]) {
        const float *Q = S.xs[mod6(gz)][0];
        v[0] = Q[c0i]; v[1] = Q[PL + c0i]; v[2] = Q[2 * PL + c0i];
    };
    float tn, tc_;
    unsigned mn, mc;
    auto fetch_tm = [&](int gz, float &t, unsigned &m) {
        const int gzc = min(max(gz, 0), nz - 1);
        t = bload(B.r_tgt, 4u * col_off, 4u * B.plane * (unsigned)gzc);
        m = (unsigned)__builtin_amdgcn_raw_buffer_load_b8(B.r_msk, (int)col_off, (int)(B.plane * (unsigned)gzc), 0);
    };

    // prologue: planes z0-2 .. z0+1 in LDS, z0+2 in registers
    float pa[PER], qa[PER];
    {
        float a0[PER], a1[PER], a2[PER], a3[PER], b0[PER], b1[PER], b2[PER], b3[PER];
        L.fetch(z0 - 2, a0, b0);
        L.fetch(z0 - 1, a1, b1);
        L.fetch(z0, a2, b2);
        L.fetch(z0 + 1, a3, b3);
        L.fetch(z0 + 2, pa, qa);
        fetch_tm(z0, tc_, mc);
        L.commit(z0 - 2, &S.xs[mod6(z0 - 2)][0][0], a0, b0);
        L.commit(z0 - 1, &S.xs[mod6(z0 - 1)][0][0], a1, b1);
        L.commit(z0, &S.xs[mod6(z0)][0][0], a2, b2);
        L.commit(z0 + 1, &S.xs[mod6(z0 + 1)][0][0], a3, b3);
    }
    __syncthreads();
    // own column: planes zz-3 .. zz+1 at the start of step zz
    float col[5][3];
    col[0][0] = col[0][1] = col[0][2] = 0.0f;
    own(z0 - 2, col[1]);
    own(z0 - 1, col[2]);
    own(z0, col[3]);
    own(z0 + 1, col[4]);

    for (int zz = z0; zz < z1; ++zz) {
        // a six-plane ring lets one barrier per plane suffice: the slot written
        // here (zz + 2) is not read by a thread still on plane zz - 1
        L.commit(zz + 2, &S.xs[mod6(zz + 2)][0][0], pa, qa);
        __syncthreads();
        L.fetch(zz + 3, pa, qa);
        const float T = tc_;
        const unsigned mk = mc;
        fetch_tm(zz + 1, tn, mn);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[0][c] = col[1][c]; col[1][c] = col[2][c]; col[2][c] = col[3][c]; col[3][c] = col[4][c];
        }
        own(zz + 2, col[4]);
        if (col_in) {
            const float *Q = S.xs[mod6(zz)][0];
            float sm[3][3][4];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float *R = Q + c * PL + c0i;
                sm[c][0][0] = R[-2]; sm[c][0][1] = R[-1]; sm[c][0][2] = R[1]; sm[c][0][3] = R[2];
                sm[c][1][0] = R[-2 * HX]; sm[c][1][1] = R[-HX]; sm[c][1][2] = R[HX]; sm[c][1][3] = R[2 * HX];
                sm[c][2][0] = col[0][c]; sm[c][2][1] = col[1][c]; sm[c][2][2] = col[3][c]; sm[c][2][3] = col[4][c];
            }
            const float dJ = jacobian_delta<CT>(P, gx, gy, zz, sm, col[2]);
            const bool in_mask = mk != 0;
            const float e = voxel_error_delta(dJ, T - 1.0f, in_mask, P.tol);
            Esum += (double)e * (double)e;
            if (in_mask) mx = fmaxf(mx, fabsf(e));
            if (store_w) bstore(B.r_w, penalised_weight(e, 1.0f + dJ, P), 4u * col_off, 4u * B.plane * (unsigned)zz);
        }
        tc_ = tn;
        mc = mn;
    }
}

template <typename CT, int TX, int TY, int MINW>
__global__ void __launch_bounds__(TX * TY, MINW)
k_eval(const Inst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    constexpr int NT = TX * TY;
    __shared__ EvalSmem<TX, TY> S;
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    const Ctrl *C = I.ctrl;
    if (!uni(C->active)) return;
    TileCoord tc = decode_tile<TX, TY>(I, tl.tile);
    tc.x0 = uni(tc.x0); tc.y0 = uni(tc.y0); tc.z0 = uni(tc.z0); tc.z1 = uni(tc.z1);

    double Esum = 0.0, unused = 0.0;
    float mx = 0.0f;
    if (uni(C->mode) == TRIAL_INTERP)
        eval_body<CT, TX, TY, true>(S, I, C, tc, Esum, mx);
    else
        eval_body<CT, TX, TY, false>(S, I, C, tc, Esum, mx);
    block_reduce3<NT>(Esum, mx, unused, S.red);
    if (threadIdx.x == 0) {
        double *p = I.partials + 4 * (long long)tl.tile;
        p[0] = Esum;
        p[1] = (double)mx;
    }
}

// ---------------------------------------------------------------------------
// k_update: W, g, candidate, controller
// ---------------------------------------------------------------------------

template <int TX, int TY>
struct UpdSmem {
    static constexpr int HX = TX + 4, PL = HX * (TY + 4);
    static constexpr int WX = TX + 2, WPL = WX * (TY + 2);
    float xs[4][3][PL];   // ring of four trial planes, halo 2
    float ws[6][WPL];     // W_{a,x} (0..2) and W_{a,y} (3..5) of one plane, halo 1
    double red[3 * (TX * TY / 64)];
    int flag;
};

// Weighted minors at a point from its central differences: rel 0 = the three
// minors of relation axis x (d along y and z), 1 = of axis y (d along x and z),
// 2 = of axis z (d along x and y).  d[c][a] includes the identity.
template <typename CT>
__device__ __forceinline__ void weighted_minors(int rel, const CT (&d)[3][3], CT w, float (&out)[3])
{
    if (rel == 0) {
        out[0] = (float)(w * (d[1][1] * d[2][2] - d[1][2] * d[2][1]));
        out[1] = (float)(w * (d[0][1] * d[2][2] - d[0][2] * d[2][1]));
        out[2] = (float)(w * (d[0][1] * d[1][2] - d[0][2] * d[1][1]));
    } else if (rel == 1) {
        out[0] = (float)(w * (d[1][0] * d[2][2] - d[1][2] * d[2][0]));
        out[1] = (float)(w * (d[0][0] * d[2][2] - d[0][2] * d[2][0]));
        out[2] = (float)(w * (d[0][0] * d[1][2] - d[0][2] * d[1][0]));
    } else {
        out[0] = (float)(w * (d[1][0] * d[2][1] - d[1][1] * d[2][0]));
        out[1] = (float)(w * (d[0][0] * d[2][1] - d[0][1] * d[2][0]));
        out[2] = (float)(w * (d[0][0] * d[1][1] - d[0][1] * d[1][0]));
    }
}

template <typename CT, int TX, int TY, bool INTERP>
__device__ __forceinline__ void update_body(UpdSmem<TX, TY> &S, const Inst &I, const Ctrl *C, const TileCoord tc,
                                            double &Gsum)
{
    using SM = UpdSmem<TX, TY>;
    constexpr int HX = SM::HX, PL = SM::PL, WX = SM::WX;
    constexpr int NHALO = 2 * TX + 2 * TY;
    using LD = PlaneLoader<TX, TY, 2, INTERP>;
    constexpr int PER = LD::PER;
    const PassParams P = load_params(I);
    const PassBuffers B = pass_buffers<INTERP>(I, C);
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int z0 = tc.z0, z1 = tc.z1;
    const unsigned n = B.n, plane = B.plane;

    LD L;
    L.r_src = B.r_src; L.r_cnd = B.r_cnd; L.plane4 = 4u * plane; L.nz = nz; L.s = B.s;
    plane_slots<TX, TY, 2, PER>(tc.x0, tc.y0, nx, ny, n, L.vo, L.ok);

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int gx = tc.x0 + lx, gy = tc.y0 + ly;
    const bool col_in = gx < nx && gy < ny;
    const bool col_interior = gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2;
    const unsigned col_off = col_in ? (unsigned)(gy * nx + gx) : 0u;
    const int c0i = (ly + 2) * HX + (lx + 2);
    const int wi = (ly + 1) * WX + (lx + 1);

    // ring point of this thread (at most one): x-ring needs W_{a,x}, y-ring W_{a,y}
    int hx = 0, hy = 0, hrel = 0;
    if (tid < NHALO) {
        const int h = tid;
        if (h < TY) { hx = -1; hy = h; hrel = 0; }
        else if (h < 2 * TY) { hx = TX; hy = h - TY; hrel = 0; }
        else if (h < 2 * TY + TX) { hx = h - 2 * TY; hy = -1; hrel = 1; }
        else { hx = h - 2 * TY - TX; hy = TY; hrel = 1; }
    }
    const int qx = tc.x0 + hx, qy = tc.y0 + hy;
    // a ring minor is used only by an interior neighbour, whose other
    // coordinates equal the ring point's; skip ring points whose central
    // differences would leave the volume
    const bool halo_in = tid < NHALO && qx >= 0 && qx < nx && qy >= 0 && qy < ny &&
                         (hrel == 0 ? (qy >= 1 && qy <= ny - 2) : (qx >= 1 && qx <= nx - 2));
    const unsigned halo_off = halo_in ? (unsigned)(qy * nx + qx) : 0u;
    const int h0i = (hy + 2) * HX + (hx + 2);
    const int hwi = (hy + 1) * WX + (hx + 1);

    auto own = [&](int gz, float (&v)[3]) {
        const float *Q = S.xs[mod4(gz)][0];
        v[0] = Q[c0i]; v[1] = Q[PL + c0i]; v[2] = Q[2 * PL + c0i];
    };
    auto fetch_w = [&](int gz, float &wt, float &wh) {
        const int gzc = min(max(gz, 0), nz - 1);
        const unsigned so = 4u * plane * (unsigned)gzc;
        wt = bload(B.r_w, 4u * col_off, so);
        wh = bload(B.r_w, 4u * halo_off, so);
    };

    // prologue: planes z0-2, z0-1 in LDS, z0 in registers
    float pa[PER], qa[PER];
    float wt_c, wh_c, wt_n, wh_n;
    {
        float a0[PER], a1[PER], b0[PER], b1[PER];
        L.fetch(z0 - 2, a0, b0);
        L.fetch(z0 - 1, a1, b1);
        L.fetch(z0, pa, qa);
        fetch_w(z0 - 1, wt_c, wh_c);
        L.commit(z0 - 2, &S.xs[mod4(z0 - 2)][0][0], a0, b0);
        L.commit(z0 - 1, &S.xs[mod4(z0 - 1)][0][0], a1, b1);
    }
    __syncthreads();
    // own column: planes zz-2 .. zz at the start of step zz
    float col[3][3];
    col[0][0] = col[0][1] = col[0][2] = 0.0f;
    own(z0 - 2, col[1]);
    own(z0 - 1, col[2]);

    float wz_m[3] = {0, 0, 0}, wz_c[3] = {0, 0, 0}, wz_p[3] = {0, 0, 0};
    const CT cix = (CT)P.inv[0], ciy = (CT)P.inv[1], ciz = (CT)P.inv[2];
    const float h0 = P.c2[0], h1 = P.c2[1], h2 = P.c2[2];

    for (int zz = z0 - 1; zz <= z1; ++zz) {
        const int zb = zz - 1;
        const bool do_b = zb >= z0 && zb < z1 && col_in;
        const bool grad_b = do_b && col_interior && zb >= 1 && zb <= nz - 2;

        // gradient on plane zb: x/y part from the W plane of zb
        CT g0 = 0, g1 = 0, g2 = 0;
        if (grad_b) {
            const CT dxx = (CT)S.ws[0][wi - 1] - (CT)S.ws[0][wi + 1];
            const CT dyx = (CT)S.ws[1][wi - 1] - (CT)S.ws[1][wi + 1];
            const CT dzx = (CT)S.ws[2][wi - 1] - (CT)S.ws[2][wi + 1];
            const CT dxy = (CT)S.ws[3][wi - WX] - (CT)S.ws[3][wi + WX];
            const CT dyy = (CT)S.ws[4][wi - WX] - (CT)S.ws[4][wi + WX];
            const CT dzy = (CT)S.ws[5][wi - WX] - (CT)S.ws[5][wi + WX];
            g0 = cix * dxx - ciy * dxy;
            g1 = -cix * dyx + ciy * dyy;
            g2 = cix * dzx - ciy * dzy;
        }
        const float ux = col[1][0], uy = col[1][1], uz = col[1][2];   // trial at plane zb

        // a four-plane ring: the slot of zz + 1 is not read by a thread still on zz - 1
        L.commit(zz + 1, &S.xs[mod4(zz + 1)][0][0], pa, qa);
        __syncthreads();
        L.fetch(zz + 2, pa, qa);
        fetch_w(zz + 1, wt_n, wh_n);
#pragma unroll
        for (int c = 0; c < 3; ++c) { col[0][c] = col[1][c]; col[1][c] = col[2][c]; }
        own(zz + 1, col[2]);   // now planes zz-1, zz, zz+1

        // weighted minors on plane zz
#pragma unroll
        for (int k = 0; k < 3; ++k) { wz_m[k] = wz_c[k]; wz_c[k] = wz_p[k]; wz_p[k] = 0.0f; }
        if (zz >= 0 && zz < nz) {
            const float *Q = S.xs[mod4(zz)][0];
            if (col_in && (gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2 && zz >= 1 && zz <= nz - 2)) {
                CT d[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float *R = Q + c * PL + c0i;
                    d[c][0] = (CT)(R[1] - R[-1]) * (CT)h0;
                    d[c][1] = (CT)(R[HX] - R[-HX]) * (CT)h1;
                    d[c][2] = (CT)(col[2][c] - col[0][c]) * (CT)h2;
                }
                d[0][0] += (CT)1.0; d[1][1] += (CT)1.0; d[2][2] += (CT)1.0;
                float o[3];
                weighted_minors<CT>(0, d, (CT)wt_c, o);
                S.ws[0][wi] = o[0]; S.ws[1][wi] = o[1]; S.ws[2][wi] = o[2];
                weighted_minors<CT>(1, d, (CT)wt_c, o);
                S.ws[3][wi] = o[0]; S.ws[4][wi] = o[1]; S.ws[5][wi] = o[2];
                weighted_minors<CT>(2, d, (CT)wt_c, wz_p);
            } else if (col_in) {
                // points on the outer face: only the minors whose derivatives
                // stay inside the volume are ever used; compute those, zero else
                CT d[3][3];
                const bool okx = gx >= 1 && gx <= nx - 2, oky = gy >= 1 && gy <= ny - 2;
                const bool okz = zz >= 1 && zz <= nz - 2;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float *R = Q + c * PL + c0i;
                    d[c][0] = okx ? (CT)(R[1] - R[-1]) * (CT)h0 : (CT)0;
                    d[c][1] = oky ? (CT)(R[HX] - R[-HX]) * (CT)h1 : (CT)0;
                    d[c][2] = okz ? (CT)(col[2][c] - col[0][c]) * (CT)h2 : (CT)0;
                }
                d[0][0] += (CT)1.0; d[1][1] += (CT)1.0; d[2][2] += (CT)1.0;
                float o[3];
                weighted_minors<CT>(0, d, (CT)wt_c, o);
                S.ws[0][wi] = o[0]; S.ws[1][wi] = o[1]; S.ws[2][wi] = o[2];
                weighted_minors<CT>(1, d, (CT)wt_c, o);
                S.ws[3][wi] = o[0]; S.ws[4][wi] = o[1]; S.ws[5][wi] = o[2];
                weighted_minors<CT>(2, d, (CT)wt_c, wz_p);
            }
            if (halo_in && zz >= 1 && zz <= nz - 2) {
                // ring point: derivatives along the two axes of its relation
                CT d[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float *R = Q + c * PL + h0i;
                    d[c][0] = hrel == 1 ? (CT)(R[1] - R[-1]) * (CT)h0 : (CT)0;
                    d[c][1] = hrel == 0 ? (CT)(R[HX] - R[-HX]) * (CT)h1 : (CT)0;
                    d[c][2] = (CT)(S.xs[mod4(zz + 1)][c][h0i] - S.xs[mod4(zz - 1)][c][h0i]) * (CT)h2;
                }
                d[0][0] += (CT)1.0; d[1][1] += (CT)1.0; d[2][2] += (CT)1.0;
                float o[3];
                weighted_minors<CT>(hrel, d, (CT)wh_c, o);
                const int base = hrel == 0 ? 0 : 3;
                S.ws[base + 0][hwi] = o[0]; S.ws[base + 1][hwi] = o[1]; S.ws[base + 2][hwi] = o[2];
            }
        }

        // finish the gradient of plane zb and write the candidate (and the trial)
        if (do_b) {
            float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
            if (grad_b) {
                f0 = (float)(g0 + ciz * ((CT)wz_m[0] - (CT)wz_p[0]));
                f1 = (float)(g1 - ciz * ((CT)wz_m[1] - (CT)wz_p[1]));
                f2 = (float)(g2 + ciz * ((CT)wz_m[2] - (CT)wz_p[2]));
                Gsum += (double)f0 * f0 + (double)f1 * f1 + (double)f2 * f2;
            }
            const unsigned so = 4u * (unsigned)zb * plane;
            const unsigned v = 4u * col_off;
            if (B.has_next) {
                bstore(B.r_next, ux - B.eta_next * f0, v, so);
                bstore(B.r_next, uy - B.eta_next * f1, v + 4u * n, so);
                bstore(B.r_next, uz - B.eta_next * f2, v + 8u * n, so);
            }
            if (B.has_trial) {
                bstore(B.r_trial, ux, v, so);
                bstore(B.r_trial, uy, v + 4u * n, so);
                bstore(B.r_trial, uz, v + 8u * n, so);
            }
        }
        wt_c = wt_n;
        wh_c = wh_n;
        __syncthreads();
    }
}

template <typename CT, int TX, int TY, int MINW>
__global__ void __launch_bounds__(TX * TY, MINW)
k_update(const Inst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    constexpr int NT = TX * TY;
    static_assert(2 * TX + 2 * TY <= NT, "one ring point per thread at most");
    __shared__ UpdSmem<TX, TY> S;
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    Ctrl *C = I.ctrl;
    if (!uni(C->active)) return;
    TileCoord tc = decode_tile<TX, TY>(I, tl.tile);
    tc.x0 = uni(tc.x0); tc.y0 = uni(tc.y0); tc.z0 = uni(tc.z0); tc.z1 = uni(tc.z1);

    double Gsum = 0.0, unused = 0.0;
    float unused_m = 0.0f;
    // passes that form no candidate (initial step-length search) need only
    // the energy; their gradient is never used
    if (uni(C->out_next) >= 0) {
        if (uni(C->mode) == TRIAL_INTERP)
            update_body<CT, TX, TY, true>(S, I, C, tc, Gsum);
        else
            update_body<CT, TX, TY, false>(S, I, C, tc, Gsum);
    }
    block_reduce3<NT>(unused, unused_m, Gsum, S.red);
    if (threadIdx.x == 0) {
        double *p = I.partials + 4 * (long long)tl.tile;
        __hip_atomic_store(p + 2, Gsum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned prev = __hip_atomic_fetch_add(I.counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        S.flag = (prev + 1 == (unsigned)(I.ntx * I.nty * I.ntz)) ? 1 : 0;
    }
    __syncthreads();
    if (S.flag) {
        double E, G;
        float m;
        fold_partials<NT>(I, E, m, G, S.red);
        if (threadIdx.x == 0) {
            grad_controller(I, C, E, m, G);
            publish_status(I, C);
            __hip_atomic_store(I.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <typename CT, int TX, int TY, int MA, int MB>
static hipError_t launch_split_variant(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL((k_eval<CT, TX, TY, MA>), dim3(nblocks), dim3(TX * TY), 0, st, insts, tiles);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_update<CT, TX, TY, MB>), dim3(nblocks), dim3(TX * TY), 0, st, insts, tiles);
    return hipGetLastError();
}

template <typename CT, int TX, int TY, int MA, int MB>
static int split_occupancy()
{
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_eval<CT, TX, TY, MA>, TX * TY, 0) != hipSuccess) a = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_update<CT, TX, TY, MB>, TX * TY, 0) != hipSuccess) b = 1;
    return std::max(1, std::min(a, b));
}

template <typename CT>
static hipError_t launch_split_ct(int variant, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    switch (variant) {
    case SPLIT_BASE + 1: return launch_split_variant<CT, 32, 8, 4, 4>(insts, tiles, nblocks, st);
    case SPLIT_BASE + 2: return launch_split_variant<CT, 64, 8, 2, 2>(insts, tiles, nblocks, st);
    default: return launch_split_variant<CT, 32, 16, 2, 2>(insts, tiles, nblocks, st);
    }
}

template <typename CT>
static int split_occupancy_ct(int variant)
{
    switch (variant) {
    case SPLIT_BASE + 1: return split_occupancy<CT, 32, 8, 4, 4>();
    case SPLIT_BASE + 2: return split_occupancy<CT, 64, 8, 2, 2>();
    default: return split_occupancy<CT, 32, 16, 2, 2>();
    }
}

hipError_t launch_split_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                             hipStream_t st)
{
    return precision == 64 ? launch_split_ct<double>(variant, insts, tiles, nblocks, st)
                           : launch_split_ct<float>(variant, insts, tiles, nblocks, st);
}

int split_pass_blocks_per_cu(int precision, int variant)
{
    return precision == 64 ? split_occupancy_ct<double>(variant) : split_occupancy_ct<float>(variant);
}

}  // namespace dt
