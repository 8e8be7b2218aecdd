// dt_grad_pass.hip -- the fused pass of the gradient solver (one launch per trial).
//
// A block owns a TX x TY column tile and streams the planes z0..z1-1 of it.
// Per plane zz the pass
//   (A) evaluates on plane zz, over the tile plus a 1-voxel ring, the Jacobian
//       J (region map of src/jacobian.c:70-242), the voxel error e
//       (src/headers/error.h:128-142), the penalised weight
//       w = e + zeta (J - delta)[J < delta] (displacement_field_gradient.c:71-88)
//       and the weighted cofactors W_ab = w * M_ab, where M_ab are the 2x2 minors
//       of I + D (D = 2nd-order central derivatives) used by the adjoint stencil;
//   (B) forms, on plane zz-1, the gradient as the difference of W across each axis
//         g_x = +ix dW_xx - iy dW_xy + iz dW_xz      dW_ab = W_ab(p - e_b) - W_ab(p + e_b)
//         g_y = -ix dW_yx + iy dW_yy - iz dW_yz
//         g_z = +ix dW_zx - iy dW_zy + iz dW_zz
//       which is displacement_field_gradient.c:26-145 reorganised so that each
//       minor is computed once per voxel instead of once per neighbour, and
//       writes the speculative candidate u - eta_next g (gradient.c:150-175).
// Energy sums are FP64; the block's partials are folded by the last block of
// the instance, which runs the line-search controller (gradient.c:250-421).
//
// Memory: trial planes are read with buffer loads (SGPR plane offset, fixed
// per-thread VGPR offsets) two planes ahead of use into registers, and staged
// in a ring of five LDS planes with a 3-voxel halo (the 4th-order J at the
// ring of the tile reaches 3 voxels out).  One LDS plane holds W_{a,x} and
// W_{a,y}; W_{a,z} and the thread's own column of the trial field live in
// registers across planes.  Stencil arithmetic is FP32 with contraction (this
// file is the solver's; the bit-exact standalone jacobian is separate).
#include "dt_device.hpp"
#include "dt_kernels.hpp"
#include "dt_pass_common.hpp"

// Planes of trial-field prefetch in registers (1 or 2).
#ifndef DT_PREFETCH_DIST
#define DT_PREFETCH_DIST 1
#endif

namespace dt {

template <int TX, int TY>
struct GradSmem {
    static constexpr int HX = TX + 6, HY = TY + 6, PL = HX * HY;
    static constexpr int WX = TX + 2, WY = TY + 2, WPL = WX * WY;
    float xs[5][3][PL];   // trial planes, halo 3
    float ws[6][WPL];     // W_{a,x} (0..2) and W_{a,y} (3..5) of one plane, halo 1
    double red[3 * (TX * TY / 64)];
    int flag;
};

__device__ __forceinline__ int ring5(int z) { return (z + 10) % 5; }

// Stage A at one point: samples s[c][axis][m2, m1, p1, p2] and centre v0[c].
// what: 0 = tile point (all W), 1 = x-ring (W_{a,x} only), 2 = y-ring (W_{a,y} only).
template <typename CT, int TX, int TY>
__device__ __forceinline__ void stage_a(GradSmem<TX, TY> &S, const PassParams &P, int zz, int gx, int gy,
                                        const float (&sm)[3][3][4], const float (&v0)[3], int wi, float T,
                                        bool in_mask, int what, float wz[3], bool count, double &Esum, float &mx)
{
    // J - 1 and J - T without the rounding of J near 1.0
    const float dJ = jacobian_delta<CT>(P, gx, gy, zz, sm, v0);
    const float J = 1.0f + dJ;
    const float e = voxel_error_delta(dJ, T - 1.0f, in_mask, P.tol);
    if (count) {
        Esum += (double)e * (double)e;
        if (in_mask) mx = fmaxf(mx, fabsf(e));
    }
    const CT cw = (CT)penalised_weight(e, J, P);

    // 2nd-order central derivatives (gradient.c:13-21), identity on the diagonal
    CT d2[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            d2[c][a] = (CT)(sm[c][a][2] - sm[c][a][1]) * (CT)P.c2[a] + (c == a ? (CT)1.0 : (CT)0.0);

    if (what != 2) {
        // relation axis x: minors from derivatives along y and z
        S.ws[0][wi] = (float)(cw * (d2[1][1] * d2[2][2] - d2[1][2] * d2[2][1]));
        S.ws[1][wi] = (float)(cw * (d2[0][1] * d2[2][2] - d2[0][2] * d2[2][1]));
        S.ws[2][wi] = (float)(cw * (d2[0][1] * d2[1][2] - d2[0][2] * d2[1][1]));
    }
    if (what != 1) {
        // relation axis y: minors from derivatives along x and z
        S.ws[3][wi] = (float)(cw * (d2[1][0] * d2[2][2] - d2[1][2] * d2[2][0]));
        S.ws[4][wi] = (float)(cw * (d2[0][0] * d2[2][2] - d2[0][2] * d2[2][0]));
        S.ws[5][wi] = (float)(cw * (d2[0][0] * d2[1][2] - d2[0][2] * d2[1][0]));
    }
    if (what == 0) {
        // relation axis z: minors from derivatives along x and y
        wz[0] = (float)(cw * (d2[1][0] * d2[2][1] - d2[1][1] * d2[2][0]));
        wz[1] = (float)(cw * (d2[0][0] * d2[2][1] - d2[0][1] * d2[2][0]));
        wz[2] = (float)(cw * (d2[0][0] * d2[1][1] - d2[0][1] * d2[1][0]));
    }
}

// The body of a pass, specialised on the trial form (INTERP: the trial is
// base + s (cand - base), formed when it enters LDS; it is also stored).
//
// Every buffer load in the plane loop is unconditional (out-of-volume slots
// read offset 0 and are zeroed by a select at commit), so the wait counters
// stay exact; loaded values stay raw in registers until their plane is
// committed.  DIRECT passes prefetch two planes ahead (two register sets,
// loop unrolled by two); INTERP passes load two fields and prefetch one ahead.
template <typename CT, int TX, int TY, bool INTERP>
__device__ __forceinline__ void grad_pass_body(GradSmem<TX, TY> &S, const Inst &I, const Ctrl *C,
                                               const TileCoord tc, double &Esum, float &mx, double &Gsum)
{
    constexpr int NT = TX * TY;
    using SM = GradSmem<TX, TY>;
    constexpr int HX = SM::HX, PL = SM::PL, WX = SM::WX;
    constexpr int NHALO = 2 * TX + 2 * TY;
    constexpr int PER = (3 * PL + NT - 1) / NT;

    const PassParams P = load_params(I);
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const unsigned n = uni((unsigned)I.n);
    const unsigned plane = (unsigned)nx * (unsigned)ny;
    const unsigned fbytes = 12u * n;
    const int b_src = uni(C->src), b_cnd = uni(INTERP ? C->cand : C->src);
    const int b_next = uni(C->out_next), b_trial = uni(C->out_trial);
    const bool has_next = b_next >= 0, has_trial = INTERP && b_trial >= 0;
    float *p_src = uni(I.buf[b_src]);
    const __amdgpu_buffer_rsrc_t r_src = make_rsrc(p_src, fbytes);
    const __amdgpu_buffer_rsrc_t r_cnd = make_rsrc(uni(I.buf[b_cnd]), fbytes);
    const __amdgpu_buffer_rsrc_t r_tgt = make_rsrc(uni(I.target), 4u * n);
    const __amdgpu_buffer_rsrc_t r_msk = make_rsrc(uni(I.mask), n);
    const __amdgpu_buffer_rsrc_t r_next = make_rsrc(has_next ? uni(I.buf[b_next]) : p_src, fbytes);
    const __amdgpu_buffer_rsrc_t r_trial = make_rsrc(has_trial ? uni(I.buf[b_trial]) : p_src, fbytes);
    const float s = uni(C->s);
    const float eta_next = uni(C->eta_next);
    const int z0 = tc.z0, z1 = tc.z1;

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int gx = tc.x0 + lx, gy = tc.y0 + ly;
    const bool col_in = gx < nx && gy < ny;
    const bool col_interior = gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2;
    const unsigned col_off = col_in ? (unsigned)(gy * nx + gx) : 0u;
    const int c0i = (ly + 3) * HX + (lx + 3);
    const int wi = (ly + 1) * WX + (lx + 1);

    // ring point of this thread (at most one)
    int hx = 0, hy = 0, hwhat = 0;
    if (tid < NHALO) {
        const int h = tid;
        if (h < TY) { hx = -1; hy = h; hwhat = 1; }
        else if (h < 2 * TY) { hx = TX; hy = h - TY; hwhat = 1; }
        else if (h < 2 * TY + TX) { hx = h - 2 * TY; hy = -1; hwhat = 2; }
        else { hx = h - 2 * TY - TX; hy = TY; hwhat = 2; }
    }
    const int qx = tc.x0 + hx, qy = tc.y0 + hy;
    const bool halo_in = tid < NHALO && qx >= 0 && qx < nx && qy >= 0 && qy < ny;
    const unsigned halo_off = halo_in ? (unsigned)(qy * nx + qx) : 0u;
    const int h0i = (hy + 3) * HX + (hx + 3);
    const int hwi = (hy + 1) * WX + (hx + 1);

    // byte offsets of this thread's load slots inside a halo-3 plane
    unsigned vo[PER];
    unsigned ok = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        const int c = i / PL, r = i - c * PL;
        const int yy = r / HX, xx = r - yy * HX;
        const int x = tc.x0 - 3 + xx, y = tc.y0 - 3 + yy;
        const bool v = i < 3 * PL && x >= 0 && x < nx && y >= 0 && y < ny;
        vo[k] = v ? 4u * ((unsigned)c * n + (unsigned)(y * nx + x)) : 0u;
        ok |= (v ? 1u : 0u) << k;
    }

    auto plane_soff = [&](int gz) -> unsigned {
        const int gzc = min(max(gz, 0), nz - 1);
        return 4u * (unsigned)gzc * plane;
    };
    auto fetch = [&](int gz, float (&a)[PER], float (&b)[PER]) {
        const unsigned so = plane_soff(gz);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            a[k] = bload(r_src, vo[k], so);
            if (INTERP) b[k] = bload(r_cnd, vo[k], so);
        }
    };
    auto commit = [&](int gz, const float (&a)[PER], const float (&b)[PER]) {
        const bool zin = gz >= 0 && gz < nz;
        float *dst = &S.xs[ring5(gz)][0][0];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * NT;
            float v = a[k];
            if (INTERP) v = v + s * (b[k] - v);
            v = (zin && ((ok >> k) & 1u)) ? v : 0.0f;
            if (i < 3 * PL) dst[i] = v;
        }
    };
    struct TM { float ti, th; unsigned mi, mh; };
    auto fetch_tm = [&](int gz, TM &o) {
        const int gzc = min(max(gz, 0), nz - 1);
        const unsigned so4 = 4u * (unsigned)gzc * plane, so1 = (unsigned)gzc * plane;
        o.ti = bload(r_tgt, 4u * col_off, so4);
        o.th = bload(r_tgt, 4u * halo_off, so4);
        o.mi = (unsigned)__builtin_amdgcn_raw_buffer_load_b8(r_msk, (int)col_off, (int)so1, 0);
        o.mh = (unsigned)__builtin_amdgcn_raw_buffer_load_b8(r_msk, (int)halo_off, (int)so1, 0);
    };

    // own column of the trial field: planes zz-3 .. zz+1 at the start of step
    // zz, planes zz-2 .. zz+2 once plane zz+2 is in LDS
    float col[5][3];
    auto own = [&](int gz, float (&v)[3]) {
        const float *Q = S.xs[ring5(gz)][0];
        v[0] = Q[c0i]; v[1] = Q[PL + c0i]; v[2] = Q[2 * PL + c0i];
    };

    // prologue: planes z0-3 .. z0 in LDS, the next planes in registers; all
    // loads are issued before the first plane is committed
    constexpr int DIST = DT_PREFETCH_DIST;
    float pa[PER], pb[PER], qa[PER], qb[PER];
    TM ta, tb;
    {
        float a0[PER], a1[PER], a2[PER], a3[PER], b0[PER], b1[PER], b2[PER], b3[PER];
        fetch(z0 - 3, a0, b0);
        fetch(z0 - 2, a1, b1);
        fetch(z0 - 1, a2, b2);
        fetch(z0, a3, b3);
        fetch(z0 + 1, pa, qa);
        if (DIST == 2) fetch(z0 + 2, pb, qb);
        fetch_tm(z0 - 1, ta);
        fetch_tm(z0, tb);
        commit(z0 - 3, a0, b0);
        commit(z0 - 2, a1, b1);
        commit(z0 - 1, a2, b2);
        commit(z0, a3, b3);
    }
    __syncthreads();
    col[0][0] = col[0][1] = col[0][2] = 0.0f;
    own(z0 - 3, col[1]);
    own(z0 - 2, col[2]);
    own(z0 - 1, col[3]);
    own(z0, col[4]);

    float wz_m[3] = {0, 0, 0}, wz_c[3] = {0, 0, 0}, wz_p[3] = {0, 0, 0};
    const CT cix = (CT)P.inv[0], ciy = (CT)P.inv[1], ciz = (CT)P.inv[2];

    auto step = [&](int zz, float (&pre)[PER], float (&pre2)[PER], TM &tm) {
        const int zb = zz - 1;
        const bool do_b = zb >= z0 && zb < z1 && col_in;
        const bool grad_b = do_b && col_interior && zb >= 1 && zb <= nz - 2;

        // (B, x/y part) gradient on plane zb from the W plane of zb
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
        // the trial value of this thread's voxel on plane zb (= zz - 1)
        const float ux = col[2][0], uy = col[2][1], uz = col[2][2];

        commit(zz + 2, pre, pre2);
        __syncthreads();
        fetch(zz + 2 + DIST, pre, pre2);
        const TM cur = tm;
        fetch_tm(zz + 2, tm);
        // slide the own-column window to planes zz-2 .. zz+2
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[0][c] = col[1][c]; col[1][c] = col[2][c]; col[2][c] = col[3][c]; col[3][c] = col[4][c];
        }
        own(zz + 2, col[4]);

        // (A) plane zz
#pragma unroll
        for (int k = 0; k < 3; ++k) { wz_m[k] = wz_c[k]; wz_c[k] = wz_p[k]; wz_p[k] = 0.0f; }
        if (zz >= 0 && zz < nz) {
            const bool count = zz >= z0 && zz < z1;
            if (col_in) {
                const float *Q = S.xs[ring5(zz)][0];
                float sm[3][3][4];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float *R = Q + c * PL + c0i;
                    sm[c][0][0] = R[-2]; sm[c][0][1] = R[-1]; sm[c][0][2] = R[1]; sm[c][0][3] = R[2];
                    sm[c][1][0] = R[-2 * HX]; sm[c][1][1] = R[-HX]; sm[c][1][2] = R[HX]; sm[c][1][3] = R[2 * HX];
                    sm[c][2][0] = col[0][c]; sm[c][2][1] = col[1][c]; sm[c][2][2] = col[3][c]; sm[c][2][3] = col[4][c];
                }
                stage_a<CT, TX, TY>(S, P, zz, gx, gy, sm, col[2], wi, cur.ti, cur.mi != 0, 0, wz_p, count, Esum, mx);
            }
            if (halo_in) {
                const float *Q = S.xs[ring5(zz)][0];
                const int s_m2 = ring5(zz - 2), s_m1 = ring5(zz - 1), s_p1 = ring5(zz + 1), s_p2 = ring5(zz + 2);
                float sm[3][3][4], v0[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float *R = Q + c * PL + h0i;
                    v0[c] = R[0];
                    sm[c][0][0] = R[-2]; sm[c][0][1] = R[-1]; sm[c][0][2] = R[1]; sm[c][0][3] = R[2];
                    sm[c][1][0] = R[-2 * HX]; sm[c][1][1] = R[-HX]; sm[c][1][2] = R[HX]; sm[c][1][3] = R[2 * HX];
                    sm[c][2][0] = S.xs[s_m2][c][h0i]; sm[c][2][1] = S.xs[s_m1][c][h0i];
                    sm[c][2][2] = S.xs[s_p1][c][h0i]; sm[c][2][3] = S.xs[s_p2][c][h0i];
                }
                float dummy[3];
                double e0 = 0.0;
                float m0 = 0.0f;
                stage_a<CT, TX, TY>(S, P, zz, qx, qy, sm, v0, hwi, cur.th, cur.mh != 0, hwhat, dummy, false, e0, m0);
            }
        }

        // (B, z part) finish the gradient of plane zb and write the candidate
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
            if (has_next) {
                bstore(r_next, ux - eta_next * f0, v, so);
                bstore(r_next, uy - eta_next * f1, v + 4u * n, so);
                bstore(r_next, uz - eta_next * f2, v + 8u * n, so);
            }
            if (has_trial) {
                bstore(r_trial, ux, v, so);
                bstore(r_trial, uy, v + 4u * n, so);
                bstore(r_trial, uz, v + 8u * n, so);
            }
        }
        __syncthreads();
    };

    if (DIST == 2) {
        for (int zz = z0 - 1; zz <= z1; zz += 2) {
            step(zz, pa, qa, ta);
            if (zz + 1 <= z1) step(zz + 1, pb, qb, tb);
        }
    } else {
        for (int zz = z0 - 1; zz <= z1; zz += 2) {
            step(zz, pa, qa, ta);
            if (zz + 1 <= z1) step(zz + 1, pa, qa, tb);
        }
    }
}

template <typename CT, int TX, int TY, int MINW>
__global__ void __launch_bounds__(TX * TY, MINW)
k_grad_pass(const Inst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    constexpr int NT = TX * TY;
    static_assert(2 * TX + 2 * TY <= NT, "one ring point per thread at most");
    __shared__ GradSmem<TX, TY> S;

    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    Ctrl *C = I.ctrl;
    if (!uni(C->active)) return;
    TileCoord tc = decode_tile<TX, TY>(I, tl.tile);
    tc.x0 = uni(tc.x0); tc.y0 = uni(tc.y0); tc.z0 = uni(tc.z0); tc.z1 = uni(tc.z1);

    double Esum = 0.0, Gsum = 0.0;
    float mx = 0.0f;
    if (uni(C->mode) == TRIAL_INTERP)
        grad_pass_body<CT, TX, TY, true>(S, I, C, tc, Esum, mx, Gsum);
    else
        grad_pass_body<CT, TX, TY, false>(S, I, C, tc, Esum, mx, Gsum);

    block_reduce3<NT>(Esum, mx, Gsum, S.red);
    if (arrive<NT>(I, tl.tile, Esum, mx, Gsum, &S.flag)) {
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

template <typename CT, int TX, int TY, int MINW>
static hipError_t launch_variant(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL((k_grad_pass<CT, TX, TY, MINW>), dim3(nblocks), dim3(TX * TY), 0, st, insts, tiles);
    return hipGetLastError();
}

template <typename CT, int TX, int TY, int MINW>
static int occupancy_of()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_grad_pass<CT, TX, TY, MINW>, TX * TY, 0) !=
        hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

// Variant table: column tile TX x TY and the waves per SIMD the register
// allocation is capped for (see grad_variant_dims).
template <typename CT>
static hipError_t launch_ct(int variant, const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    switch (variant) {
    case 1: return launch_variant<CT, 32, 16, 3>(insts, tiles, nblocks, st);
    case 2: return launch_variant<CT, 64, 4, 3>(insts, tiles, nblocks, st);
    case 3: return launch_variant<CT, 64, 8, 3>(insts, tiles, nblocks, st);
    case 4: return launch_variant<CT, 32, 8, 2>(insts, tiles, nblocks, st);
    case 5: return launch_variant<CT, 32, 16, 2>(insts, tiles, nblocks, st);
    case 6: return launch_variant<CT, 32, 16, 4>(insts, tiles, nblocks, st);
    case 7: return launch_variant<CT, 32, 8, 4>(insts, tiles, nblocks, st);
    default: return launch_variant<CT, 32, 8, 3>(insts, tiles, nblocks, st);
    }
}

template <typename CT>
static int occupancy_ct(int variant)
{
    switch (variant) {
    case 1: return occupancy_of<CT, 32, 16, 3>();
    case 2: return occupancy_of<CT, 64, 4, 3>();
    case 3: return occupancy_of<CT, 64, 8, 3>();
    case 4: return occupancy_of<CT, 32, 8, 2>();
    case 5: return occupancy_of<CT, 32, 16, 2>();
    case 6: return occupancy_of<CT, 32, 16, 4>();
    case 7: return occupancy_of<CT, 32, 8, 4>();
    default: return occupancy_of<CT, 32, 8, 3>();
    }
}

// Streaming variants, one object per (precision, tile, register cap)
// (dt_grad_stream.hip); variant STREAM_BASE + k is entry k of this list.
#define DT_STREAM_VARIANTS(X) X(0, 32x16_w3) X(1, 32x20_w3) X(2, 64x8_w3) X(3, 32x24_w3) X(4, 64x12_w3)
#define DT_DECLARE_STREAM(k, name)                                                                          \
    hipError_t stream_launch_f32_##name(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st); \
    hipError_t stream_launch_f64_##name(const Inst *insts, const Tile *tiles, int nblocks, hipStream_t st); \
    int stream_occupancy_f32_##name();                                                                      \
    int stream_occupancy_f64_##name();
DT_STREAM_VARIANTS(DT_DECLARE_STREAM)

hipError_t launch_stream_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                              hipStream_t st)
{
    const bool d = precision == 64;
#define DT_LAUNCH_CASE(k, name)                                                                 \
    case k:                                                                                     \
        return d ? stream_launch_f64_##name(insts, tiles, nblocks, st)                          \
                 : stream_launch_f32_##name(insts, tiles, nblocks, st);
    switch (variant - STREAM_BASE) {
        DT_STREAM_VARIANTS(DT_LAUNCH_CASE)
    default: return hipErrorInvalidValue;
    }
#undef DT_LAUNCH_CASE
}

int stream_pass_blocks_per_cu(int precision, int variant)
{
    const bool d = precision == 64;
#define DT_OCC_CASE(k, name) \
    case k: return d ? stream_occupancy_f64_##name() : stream_occupancy_f32_##name();
    switch (variant - STREAM_BASE) {
        DT_STREAM_VARIANTS(DT_OCC_CASE)
    default: return 1;
    }
#undef DT_OCC_CASE
}

hipError_t launch_grad_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                            hipStream_t st)
{
    if (variant >= STREAM_BASE) return launch_stream_pass(precision, variant, insts, tiles, nblocks, st);
    if (variant >= SPLIT_BASE) return launch_split_pass(precision, variant, insts, tiles, nblocks, st);
    return precision == 64 ? launch_ct<double>(variant, insts, tiles, nblocks, st)
                           : launch_ct<float>(variant, insts, tiles, nblocks, st);
}

int grad_pass_blocks_per_cu(int precision, int variant)
{
    if (variant >= STREAM_BASE) return stream_pass_blocks_per_cu(precision, variant);
    if (variant >= SPLIT_BASE) return split_pass_blocks_per_cu(precision, variant);
    return precision == 64 ? occupancy_ct<double>(variant) : occupancy_ct<float>(variant);
}

}  // namespace dt
