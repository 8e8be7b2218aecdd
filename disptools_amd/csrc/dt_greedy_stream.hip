// dt_greedy_stream.hip -- the greedy solver's pass as one streaming kernel.
//
// One launch per trial: it forms the trial field from the accepted iterate
// and the current error map (src/displacement_field_greedy.c:23-67),
//   u_new = u_base + (eta d_c / 2) (e(p + e_c) - e(p - e_c))   on the interior,
// where the outer face keeps the output buffer's contents (the reference never
// writes it), evaluates the Jacobian of the trial (region map of
// src/jacobian.c:70-242), its voxel error (src/headers/error.h:128-142) into
// the new error map, the exact energy sum and masked max; the greedy
// controller (displacement_field_greedy.c:85-247) advances at the start of
// the next launch, in every block, on these sums (pass_begin).  With the
// step off (the evaluation of the initial guess) the trial is u_base itself.
//
// A block owns a TX x TY column tile and streams z.  Per plane step zz:
//   * the trial is formed on plane zz+2 over the tile and a 2-voxel ring (the
//     reach of the 4th-order stencil), from the error map's plane zz+2 in LDS
//     (halo 3) and its planes zz+1, zz+3 at the thread's own points
//     (registers), into a ring of three LDS trial planes;
//   * the Jacobian and error are evaluated on plane zz at the tile's columns,
//     in-plane neighbours from LDS and z-neighbours from the thread's own
//     column of trial values (registers);
// one barrier per plane separates every LDS write from its reads.  The step
// arithmetic is not contracted (the reference's C rounds each operation).
// Algorithmic traffic: read u_base 12 + e 4 + target 4 + mask 1, write u_new
// 12 + e 4 = 37 bytes per voxel (the two-kernel form: 49).
#include "dt_device.hpp"
#include "dt_kernels.hpp"
#include "dt_pass_common.hpp"

namespace dt {

// u + dd (ep - em), rounded as the reference's C: three operations, no FMA
// (this file is compiled with -ffp-contract=fast-honor-pragmas).
__device__ __forceinline__ float greedy_update(float u, float dd, float ep, float em)
{
#pragma clang fp contract(off)
    return u + dd * (ep - em);
}

template <int TX, int TY>
struct GreedyGeom {
    static constexpr int NT = TX * TY;
    static constexpr int RX = TX + 4, RY = TY + 4, RP = RX * RY;      // trial region (halo 2)
    static constexpr int EX = TX + 6, EY = TY + 6, EP = EX * EY;      // error-map plane (halo 3)
    static constexpr int NRING2 = RP - NT;                             // region points outside the tile
    static constexpr int NRING3 = EP - RP;                             // error points outside the region
    static_assert(NRING2 <= NT && NRING3 <= NT, "one extra point per thread at most");
};

template <int TX, int TY>
struct GreedySmem {
    using G = GreedyGeom<TX, TY>;
    float es[2][G::EP];            // error-map planes by parity
    float us[3][3][G::RP];         // trial planes, ring of three
    SumsWave red[G::NT / 64];
    Ctrl ctrl;                     // this pass's control state (pass_begin)
};

// Region coordinates (rx, ry) of the r-th point of the region ring (outside
// the tile), and error-plane coordinates (ex, ey) of the r-th point of the
// error ring (outside the region).
template <int TX, int TY>
__device__ __forceinline__ void ring2_point(int r, int &rx, int &ry)
{
    constexpr int RX = TX + 4;
    if (r < 2 * RX) { ry = r / RX; rx = r % RX; }
    else if (r < 4 * RX) { const int q = r - 2 * RX; ry = TY + 2 + q / RX; rx = q % RX; }
    else { const int q = r - 4 * RX; ry = 2 + q / 4; const int k = q & 3; rx = k < 2 ? k : TX + k; }
}

template <int TX, int TY>
__device__ __forceinline__ void ring3_point(int r, int &ex, int &ey)
{
    constexpr int EX = TX + 6;
    if (r < 2 * EX) { ey = r < EX ? 0 : TY + 5; ex = r % EX; }
    else { const int q = r - 2 * EX; ey = 1 + q / 2; ex = (q & 1) ? TX + 5 : 0; }
}

// The energy is summed exactly (fixed point, dt_device.hpp): its terms are
// the double sums of e^2 over groups of four planes of one column, starting
// at z = 0 mod 4, so the sum does not depend on the tiling.
template <typename CT, int TX, int TY>
__device__ __forceinline__ void greedy_body(GreedySmem<TX, TY> &S, const Inst &I, const Ctrl *C, const TileCoord tc,
                                            u128 &Esum, unsigned &flags, float &mx)
{
    double Egrp = 0.0;
    using G = GreedyGeom<TX, TY>;
    constexpr int RX = G::RX, RP = G::RP, EX = G::EX, NRING2 = G::NRING2, NRING3 = G::NRING3;

    const PassParams P = load_params(I);
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const unsigned n = uni((unsigned)I.n);
    const unsigned plane = (unsigned)nx * (unsigned)ny;
    const unsigned fbytes = 12u * n;
    const bool step_on = uni(C->step_on) != 0;
    const int in_phase = uni(C->phase);
    const int b_base = uni(C->base);
    const int b_new = uni(C->g_new);
    const float *p_base = uni(I.buf[b_base]);
    float *p_new = uni(I.buf[b_new]);
    const __amdgpu_buffer_rsrc_t r_base = make_rsrc(p_base, fbytes);
    const __amdgpu_buffer_rsrc_t r_new = make_rsrc(p_new, fbytes);
    const __amdgpu_buffer_rsrc_t r_ecur = make_rsrc(uni(I.err[uni(C->e_cur)]), 4u * n);
    const __amdgpu_buffer_rsrc_t r_enew = make_rsrc(uni(I.err[uni(C->e_new)]), 4u * n);
    const __amdgpu_buffer_rsrc_t r_tgt = make_rsrc(uni(I.target), 4u * n);
    const __amdgpu_buffer_rsrc_t r_msk = make_rsrc(uni(I.mask), n);
    // the trial of the initial-guess evaluation is buf[base]
    const bool eval_base = in_phase == PH_INIT;
    const float eta = uni(C->eta);
    const float ddx = eta * uni(I.dx) * 0.5f;
    const float ddy = eta * uni(I.dy) * 0.5f;
    const float ddz = eta * uni(I.dz) * 0.5f;
    const int z0 = tc.z0, z1 = tc.z1;
    const unsigned cn4 = 4u * n;

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    // region slots: 0 = own column, 1 = a region-ring point (threads < NRING2)
    int rxs[2], rys[2];
    rxs[0] = lx + 2; rys[0] = ly + 2;
    const bool has1 = tid < NRING2;
    if (has1) ring2_point<TX, TY>(tid, rxs[1], rys[1]);
    else { rxs[1] = 2; rys[1] = 2; }
    int gxs[2], gys[2];
    bool ins[2], ints[2];
    unsigned vs[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        gxs[k] = tc.x0 - 2 + rxs[k];
        gys[k] = tc.y0 - 2 + rys[k];
        ins[k] = (k == 0 || has1) && gxs[k] >= 0 && gxs[k] < nx && gys[k] >= 0 && gys[k] < ny;
        ints[k] = gxs[k] >= 1 && gxs[k] <= nx - 2 && gys[k] >= 1 && gys[k] <= ny - 2;
        vs[k] = ins[k] ? (unsigned)(gys[k] * nx + gxs[k]) : 0u;
    }
    // error-ring slot (threads < NRING3)
    const bool has3 = tid < NRING3;
    int e3x = 0, e3y = 0;
    if (has3) ring3_point<TX, TY>(tid, e3x, e3y);
    const int g3x = tc.x0 - 3 + e3x, g3y = tc.y0 - 3 + e3y;
    const bool in3 = has3 && g3x >= 0 && g3x < nx && g3y >= 0 && g3y < ny;
    const unsigned v3 = in3 ? (unsigned)(g3y * nx + g3x) : 0u;
    const int d3 = has3 ? e3y * EX + e3x : 0;
    // error-plane and region LDS indices of the region slots
    int eis[2], ris[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        eis[k] = (rys[k] + 1) * EX + rxs[k] + 1;
        ris[k] = rys[k] * RX + rxs[k];
    }
    const bool col_in = ins[0];
    const StencilCoef cf = load_coef(I);
    const ColCoef K = column_coefficients(P, cf, gxs[0], gys[0]);
    // region points on the volume's x / y outer face keep the output
    // buffer's contents in a step (the loads below skip the other lanes)
    bool xyf[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) xyf[k] = ins[k] && !ints[k];

    auto zsoff = [&](int gz) -> unsigned { return (unsigned)min(max(gz, 0), nz - 1) * plane; };

    // raw loads of one plane: error map at the two region slots and the ring
    // slot; trial sources at the two region slots
    struct ERaw { float e[3]; };
    struct URaw { float b[2][3]; float g[2][3]; };
    auto fetch_e = [&](int gz, ERaw &r) {
        const unsigned so = 4u * zsoff(gz);
        r.e[0] = bload(r_ecur, 4u * vs[0], so);
        r.e[1] = has1 ? bload(r_ecur, 4u * vs[1], so) : 0.0f;
        r.e[2] = has3 ? bload(r_ecur, 4u * v3, so) : 0.0f;
    };
    auto fetch_u = [&](int gz, URaw &r) {
        const unsigned so = 4u * zsoff(gz);
        const bool zface = gz <= 0 || gz >= nz - 1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool load = k == 0 || has1;
            const bool face = step_on && ins[k] && (zface || xyf[k]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r.b[k][c] = load ? bload(r_base, 4u * vs[k], so + (unsigned)c * cn4) : 0.0f;
                r.g[k][c] = face ? bload(r_new, 4u * vs[k], so + (unsigned)c * cn4) : 0.0f;
            }
        }
    };
    struct TM { float t; unsigned m; };
    auto fetch_tm = [&](int gz, TM &o) {
        const unsigned so = zsoff(gz);
        o.t = bload(r_tgt, 4u * vs[0], 4u * so);
        o.m = (unsigned)__builtin_amdgcn_raw_buffer_load_b8(r_msk, (int)vs[0], (int)so, 0);
    };

    // error z-window at the region slots: planes p-1, p, p+1 around the trial plane
    float ez[2][3];
    float col[5][3];   // own column of trial values, planes zz-2 .. zz+2
    ERaw er;           // raw error loads of plane zz+4 in flight
    URaw ur;           // raw trial-source loads of plane zz+3 in flight
    TM tm;             // target / mask of plane zz

    // commit error plane gz into LDS (zero outside the volume)
    auto commit_e = [&](int gz, const ERaw &r, float (&eslot)[2]) {
        const bool zin = gz >= 0 && gz < nz;
        float *dst = S.es[gz & 1];
        const float e0 = (zin && ins[0]) ? r.e[0] : 0.0f;
        const float e1 = (zin && ins[1]) ? r.e[1] : 0.0f;
        const float e2 = (zin && in3) ? r.e[2] : 0.0f;
        dst[eis[0]] = e0;
        if (has1) dst[eis[1]] = e1;
        if (has3) dst[d3] = e2;
        eslot[0] = e0;
        eslot[1] = e1;
    };
    // trial values at the region slots on plane gz (needs LDS error plane gz
    // and the error window ez = planes gz-1, gz, gz+1)
    auto form = [&](int gz, const URaw &r, float (&u)[2][3]) {
        const bool zin = gz >= 0 && gz < nz;
        const bool zint = gz >= 1 && gz <= nz - 2;
        const float *E = S.es[gz & 1];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool in = zin && ins[k];
            float v0 = r.b[k][0], v1 = r.b[k][1], v2 = r.b[k][2];
            if (step_on && !eval_base) {
                if (zint && ints[k]) {
                    const int ei = eis[k];
                    v0 = greedy_update(v0, ddx, E[ei + 1], E[ei - 1]);
                    v1 = greedy_update(v1, ddy, E[ei + EX], E[ei - EX]);
                    v2 = greedy_update(v2, ddz, ez[k][2], ez[k][0]);
                } else {
                    v0 = r.g[k][0]; v1 = r.g[k][1]; v2 = r.g[k][2];
                }
            }
            u[k][0] = in ? v0 : 0.0f;
            u[k][1] = in ? v1 : 0.0f;
            u[k][2] = in ? v2 : 0.0f;
        }
    };
    auto store_u = [&](int gz, const float (&u)[2][3]) {
        float *dst = &S.us[(gz + 3) % 3][0][0];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dst[c * RP + ris[0]] = u[0][c];
            if (has1) dst[c * RP + ris[1]] = u[1][c];
        }
    };

    // Steps run zz = z0-4 .. z1-1: step zz forms the trial of plane zz+2 and
    // evaluates plane zz when zz >= z0.  Before the first step: error planes
    // z0-3, z0-2 in the window (z0-2 also in LDS), plane z0-1 in flight, the
    // trial sources of plane z0-2 in flight.
    {
        ERaw ea, eb;
        fetch_e(z0 - 3, ea);
        fetch_e(z0 - 2, eb);
        fetch_e(z0 - 1, er);
        fetch_u(z0 - 2, ur);
        fetch_tm(z0 - 4, tm);
        float sa[2], sb[2];
        commit_e(z0 - 3, ea, sa);
        commit_e(z0 - 2, eb, sb);
#pragma unroll
        for (int k = 0; k < 2; ++k) { ez[k][0] = sa[k]; ez[k][1] = sb[k]; ez[k][2] = 0.0f; }
#pragma unroll
        for (int p = 0; p < 5; ++p) col[p][0] = col[p][1] = col[p][2] = 0.0f;
    }
    __syncthreads();

    // step zz: error plane zz+3 arrives, trial plane zz+2 is formed (window
    // planes zz+1 .. zz+3, LDS error plane zz+2), plane zz is evaluated
    auto step = [&](int zz) {
        // the error plane zz+3 joins LDS and the window
        float snew[2];
        commit_e(zz + 3, er, snew);
        fetch_e(zz + 4, er);
#pragma unroll
        for (int k = 0; k < 2; ++k) ez[k][2] = snew[k];
        // trial plane zz+2 (sources in flight as ur)
        float u[2][3];
        form(zz + 2, ur, u);
        fetch_u(zz + 3, ur);
        store_u(zz + 2, u);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[0][c] = col[1][c]; col[1][c] = col[2][c]; col[2][c] = col[3][c]; col[3][c] = col[4][c];
            col[4][c] = u[0][c];
        }
        // write the trial of the tile's interior columns (the face keeps its contents)
        const int zq = zz + 2;
        if (step_on && !eval_base && zq >= z0 && zq < z1 && zq >= 1 && zq <= nz - 2 && col_in && ints[0]) {
            const unsigned so = 4u * (unsigned)zq * plane;
            bstore(r_new, u[0][0], 4u * vs[0], so);
            bstore(r_new, u[0][1], 4u * vs[0], so + cn4);
            bstore(r_new, u[0][2], 4u * vs[0], so + 2u * cn4);
        }
        const TM cur = tm;
        fetch_tm(zz + 1, tm);
        // evaluate plane zz: window col[0..4] = planes zz-2 .. zz+2
        if (zz >= z0 && zz < z1 && col_in) {
            // LDS trial plane zz at the own column (written two steps ago)
            const float *Q = &S.us[(zz + 3) % 3][0][0] + ris[0];
            float sm[3][3][4];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float *R = Q + c * RP;
                sm[c][0][0] = R[-2]; sm[c][0][1] = R[-1]; sm[c][0][2] = R[1]; sm[c][0][3] = R[2];
                sm[c][1][0] = R[-2 * RX]; sm[c][1][1] = R[-RX]; sm[c][1][2] = R[RX]; sm[c][1][3] = R[2 * RX];
                sm[c][2][0] = col[0][c]; sm[c][2][1] = col[1][c]; sm[c][2][2] = col[3][c]; sm[c][2][3] = col[4][c];
            }
            const float dJ = (zz >= 2 && zz <= nz - 3) ? jacobian_delta_coef<CT>(K, sm, col[2])
                                                       : jacobian_delta<CT>(P, cf, gxs[0], gys[0], zz, sm, col[2]);
            const bool in_mask = cur.m != 0;
            const float e = voxel_error_delta(dJ, cur.t - 1.0f, in_mask, P.tol);
            Egrp += (double)e * (double)e;
            if (in_mask) mx = fmaxf(mx, fabsf(e));
            bstore(r_enew, e, 4u * vs[0], 4u * (unsigned)zz * plane);
        }
        if (zz >= z0 && zz < z1 && ((zz & 3) == 3 || zz == nz - 1)) {
            fix_add(Esum, Egrp, flags, ACC_E_HUGE, ACC_E_NAN);
            Egrp = 0.0;
        }
        // slide the error window: planes zz+2, zz+3 (zz+4 arrives next step)
#pragma unroll
        for (int k = 0; k < 2; ++k) { ez[k][0] = ez[k][1]; ez[k][1] = ez[k][2]; }
        __syncthreads();
    };

#pragma unroll 1
    for (int zz = z0 - 4; zz < z1; ++zz) step(zz);
    fix_add(Esum, Egrp, flags, ACC_E_HUGE, ACC_E_NAN);   // a chunk not ending on a group boundary
}

template <typename CT, int TX, int TY>
__global__ void __launch_bounds__(TX * TY, 2)
k_greedy_stream(const Inst *__restrict__ insts, const Tile *__restrict__ tiles, int parity)
{
    constexpr int NT = TX * TY;
    __shared__ GreedySmem<TX, TY> S;
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    if (threadIdx.x < 64) pass_begin<DT_GREEDY>(I, parity, leader_tile(tl), &S.ctrl);
    __syncthreads();
    const Ctrl *C = &S.ctrl;
    if (!uni(C->active)) {
        if (uni(C->copy_now)) copy_result_tile<TX, TY, NT>(I, C, decode_tile<TX, TY>(I, tl));
        return;
    }
    TileCoord tc = decode_tile<TX, TY>(I, tl);
    tc.x0 = uni(tc.x0); tc.y0 = uni(tc.y0); tc.z0 = uni(tc.z0); tc.z1 = uni(tc.z1);
    u128 Esum = 0;
    unsigned flags = 0;
    float mx = 0.0f;
    greedy_body<CT, TX, TY>(S, I, C, tc, Esum, flags, mx);
    add_pass_sums<NT>(I, uni(C->acc_cur), Esum, (u128)0, mx, flags, S.red);
}

template <int TX, int TY>
static hipError_t launch_greedy_fused(int precision, const Inst *insts, const Tile *tiles, int nblocks, int parity,
                                      hipStream_t st)
{
    if (precision == 64)
        hipLaunchKernelGGL((k_greedy_stream<double, TX, TY>), dim3(nblocks), dim3(TX * TY), 0, st, insts, tiles,
                           parity);
    else
        hipLaunchKernelGGL((k_greedy_stream<float, TX, TY>), dim3(nblocks), dim3(TX * TY), 0, st, insts, tiles,
                           parity);
    return hipGetLastError();
}

template <int TX, int TY>
static int greedy_fused_blocks_per_cu(int precision)
{
    int blocks = 0;
    const hipError_t r = precision == 64
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_greedy_stream<double, TX, TY>, TX * TY, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_greedy_stream<float, TX, TY>, TX * TY, 0);
    return (r == hipSuccess && blocks > 0) ? blocks : 1;
}

hipError_t launch_greedy_pass(int precision, int variant, const Inst *insts, const Tile *tiles, int nblocks,
                              int parity, hipStream_t st)
{
    if (variant == GREEDY_FUSED_WIDE)
        return launch_greedy_fused<32, 32>(precision, insts, tiles, nblocks, parity, st);
    return launch_greedy_fused<32, 16>(precision, insts, tiles, nblocks, parity, st);
}

int greedy_pass_blocks_per_cu(int precision, int variant)
{
    if (variant == GREEDY_FUSED_WIDE) return greedy_fused_blocks_per_cu<32, 32>(precision);
    return greedy_fused_blocks_per_cu<32, 16>(precision);
}

}  // namespace dt
