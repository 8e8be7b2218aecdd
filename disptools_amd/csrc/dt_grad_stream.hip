// dt_grad_stream.hip -- the fused gradient pass, single-barrier streaming form.
//
// One launch per trial: Jacobian, voxel error, penalised weight and weighted
// cofactors on plane zz; adjoint-stencil gradient and speculative candidate
// on plane zz-1; exact sums of e^2 and |g|^2 (src/displacement_field_gradient.c:26-175,
// src/jacobian.c:70-242).  The controller step that chose the trial runs at
// the start of the launch, in every block, on the previous launch's sums
// (pass_begin); nothing synchronises the blocks within a launch.  Work layout,
// balanced over the four SIMDs of a CU:
//
//  * every thread has one primary point: the first TX*TY threads own the
//    tile's columns, the next 2(TX+TY) threads (in waves of their own) the
//    ring around the tile, where W_{a,x} / W_{a,y} are needed across the tile
//    edge.  Each wave therefore evaluates the Jacobian and cofactors once per
//    plane; with 10 tile waves and 2 ring waves (32 x 20) every SIMD carries
//    three of them;
//  * a thread's first three load slots are its primary point's components,
//    so the z-neighbours of its point come from registers;
//  * LDS holds a ring of three trial planes (halo 3) and two planes of
//    W_{a,x} / W_{a,y} (halo 1 along their own axis), indexed by plane
//    parity: one barrier per plane separates every write from its reads.
//    A point's three values share one 16-byte cell, so a stencil neighbour
//    is one 96-bit LDS read; the W planes alternate by a step's place in the
//    two-step trip, so their offsets are constants;
//  * the rest of the halo (rings 2 and 3, ring-1 corners) is spread over all
//    threads;
//  * trips whose two planes lie inside the z-chunk and the volume and off
//    the z faces -- all but the first and last few of a chunk -- run a step
//    with those tests folded away (fewer scalar instructions and branches);
//  * no register that a load writes is carried across the loop's back edge:
//    a plane's loads are resolved and committed to the LDS ring at the top of
//    the next step (the ring's third plane is what lets the commit happen at
//    once), and target / mask alternate between two register sets (the loop
//    body is two steps).  A loaded register copied at the back edge makes the
//    step wait for the load it has just issued, which turns the one-plane
//    prefetch into a full memory round trip per step;
//  * the tile threads' exact sums stay in LDS to the end of the pass; after
//    one barrier wave 0 alone folds them and commits the block's sums, and
//    the other waves leave.
#include <type_traits>

#include "dt_device.hpp"
#include "dt_kernels.hpp"
#include "dt_pass_common.hpp"

namespace dt {

template <int TX, int TY>
struct StreamGeom {
    static constexpr int NOWN = TX * TY;                          // tile threads
    static constexpr int NRING = 2 * TX + 2 * TY;                 // ring-1 points without corners
    static constexpr int NT = NOWN + (NRING + 63) / 64 * 64;      // threads with a point
    // The controller step gets a wave of its own, launched after the point
    // waves, where one more wave does not raise the waves per SIMD (and so the
    // register budget): its registers then never meet the plane loop's, and
    // the point waves issue their first loads while it runs.
    static constexpr int NW = NT / 64;
    static constexpr bool CW = (NW + 4) / 4 == (NW + 3) / 4;
    static constexpr int NL = NT + (CW ? 64 : 0);                 // threads launched
    static constexpr int HX = TX + 6, HY = TY + 6, PL = HX * HY;
    // the rest of the stencils' reach: rings 2-3 along the tile's sides and
    // 12 cells at the corners (the rings' outer corner cells are never read)
    static constexpr int NOTHER = 4 * TX + 4 * TY + 12;
    static constexpr int SX = (3 * NOTHER + NT - 1) / NT;         // halo slots per thread
    static constexpr int SLOTS = 3 + SX;
    static constexpr int WPL = (TY + 2) * HX;                     // cells of a W plane: rows -1 .. TY, pitch HX
    static_assert(NOWN % 64 == 0, "tile threads fill whole waves");
    static_assert(NOTHER + 24 == PL - NOWN - NRING, "halo decomposition");
};

template <int TX, int TY>
struct StreamSmem {
    using G = StreamGeom<TX, TY>;
    u128 accE[G::NOWN], accG[G::NOWN];   // exact sums of the tile threads (kept here, not in registers)
    f3 xs[3][G::PL];                  // ring of three trial planes, halo 3 (components x, y, z of a point)
    f3 w[2][2][G::WPL];               // [plane parity][relation axis x, y]: W_{a,x} / W_{a,y}, a = 0..2
    SumsWave red[G::NT / 64];
    Ctrl ctrl;                        // this pass's control state (pass_begin)
};

// (xx, yy) in the halo-3 plane of the r-th halo point a stencil reaches beyond
// the tile and its ring (points of the tile and ring 1 without corners are
// stencil centres; a centre reads +-1 and +-2 along x and along y): rows
// -3, -2 above and TY + 1, TY + 2 below the tile, columns -3, -2 left and
// TX + 1, TX + 2 right of it, and at each corner the three cells the ring
// centres next to it reach (e.g. (-1, -1), (-2, -1), (-1, -2)).  Coordinates
// are tile-relative; the plane index adds 3.
template <int TX, int TY>
__device__ __forceinline__ void other_point(int r, int &xx, int &yy)
{
    int x, y;
    if (r < 4 * TX) {                       // two rows above, two below
        const int q = r / TX;
        x = r - q * TX;
        y = q < 2 ? q - 3 : TY + q - 1;     // -3, -2, TY + 1, TY + 2
    } else if (r < 4 * TX + 4 * TY) {       // two columns left, two right
        const int t = r - 4 * TX;
        const int q = t / TY;
        y = t - q * TY;
        x = q < 2 ? q - 3 : TX + q - 1;     // -3, -2, TX + 1, TX + 2
    } else {                                // corners: (-1,-1) (-2,-1) (-1,-2) and mirrored
        const int t = r - 4 * TX - 4 * TY;  // 0 .. 11
        const int corner = t / 3, k = t - 3 * corner;
        const int dx = k == 1 ? 2 : 1, dy = k == 2 ? 2 : 1;
        x = (corner & 1) ? TX - 1 + dx : -dx;
        y = (corner & 2) ? TY - 1 + dy : -dy;
    }
    xx = x + 3;
    yy = y + 3;
}

// Stage A at one point: J - 1, error, weight and the nine weighted cofactors
// W[3 r + a] = w * M_{a, r} (r: relation axis x, y, z).
template <typename CT, bool ZFACE>
__device__ __forceinline__ void stream_stage_a(const PassParams &P, const ColCoef &K, const float (&sm)[3][3][4],
                                               const float (&v0)[3], float T, bool in_mask, bool count,
                                               double &Egrp, float &mx, CT (&W)[9])
{
    // the stencil of every region as per-lane coefficients (no wave-wide
    // branch on boundary lanes)
    const float dJ = jacobian_delta_coef<CT, ZFACE>(K, sm, v0);
    const float J = 1.0f + dJ;
    const float e = voxel_error_delta(dJ, T - 1.0f, in_mask, P.tol);
    const float ec = count ? e : 0.0f;
    Egrp = fma_ct((double)ec, (double)ec, Egrp);
    mx = (count && in_mask) ? fmaxf(mx, fabsf(e)) : mx;
    const CT cw = (CT)penalised_weight(e, J, P);
    CT d2[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            d2[c][a] = fma_ct((CT)(sm[c][a][2] - sm[c][a][1]), (CT)P.c2[a], c == a ? (CT)1.0 : (CT)0.0);
    // relation axis x: minors from derivatives along y and z
    W[0] = cw * fms2(d2[1][1], d2[2][2], d2[1][2], d2[2][1]);
    W[1] = cw * fms2(d2[0][1], d2[2][2], d2[0][2], d2[2][1]);
    W[2] = cw * fms2(d2[0][1], d2[1][2], d2[0][2], d2[1][1]);
    // relation axis y: minors from derivatives along x and z
    W[3] = cw * fms2(d2[1][0], d2[2][2], d2[1][2], d2[2][0]);
    W[4] = cw * fms2(d2[0][0], d2[2][2], d2[0][2], d2[2][0]);
    W[5] = cw * fms2(d2[0][0], d2[1][2], d2[0][2], d2[1][0]);
    // relation axis z: minors from derivatives along x and y
    W[6] = cw * fms2(d2[1][0], d2[2][1], d2[1][1], d2[2][0]);
    W[7] = cw * fms2(d2[0][0], d2[2][1], d2[0][1], d2[2][0]);
    W[8] = cw * fms2(d2[0][0], d2[1][1], d2[0][1], d2[1][0]);
}

// The sums are exact (fixed point, dt_device.hpp); their terms are the double
// sums of e^2 and |g|^2 over groups of four planes of one column, the groups
// starting at z = 0 mod 4 (z-chunks start on group boundaries), so the sums
// do not depend on the tiling.
//
// A direct trial's body (INTERP false) starts before the controller step has
// chosen the trial: it issues the prologue's loads from buffer `guess`, then
// calls begin() (the controller step and a block barrier) and reads the
// state.  It returns false, having done nothing more, when the pass turns out
// inactive or interpolated; it reloads the prologue when the trial is another
// buffer.  An interpolated trial's body runs once the state is known (begin
// does nothing there).
template <typename CT, int TX, int TY, bool INTERP, typename Begin>
__device__ __forceinline__ bool stream_pass_body(StreamSmem<TX, TY> &S, const Inst &I, const TileCoord tc, int guess,
                                                 Begin &&begin, unsigned &flags, float &mx)
{
    const Ctrl *C = &S.ctrl;
    double Egrp = 0.0, Ggrp = 0.0;   // sums of the current group of planes
    using G = StreamGeom<TX, TY>;
    constexpr int NT = G::NT, NOWN = G::NOWN, HX = G::HX, NRING = G::NRING;
    constexpr int NOTHER = G::NOTHER, SX = G::SX, SLOTS = G::SLOTS;

    const PassParams P = load_params(I);
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const unsigned n = uni((unsigned)I.n);
    const unsigned plane = (unsigned)nx * (unsigned)ny;
    const unsigned fbytes = 12u * n;
    int b_src = INTERP ? uni(C->src) : guess;
    __amdgpu_buffer_rsrc_t r_src = make_rsrc(uni(I.buf[b_src]), fbytes);
    const __amdgpu_buffer_rsrc_t r_cnd = INTERP ? make_rsrc(uni(I.buf[uni(C->cand)]), fbytes) : r_src;
    const __amdgpu_buffer_rsrc_t r_tgt = make_rsrc(uni(I.target), 4u * n);
    const __amdgpu_buffer_rsrc_t r_msk = make_rsrc(uni(I.mask), n);
    const float s = INTERP ? uni(C->s) : 0.0f;
    // outputs, set once the state is known
    __amdgpu_buffer_rsrc_t r_next = r_src, r_trial = r_src;
    bool has_next = false, has_trial = false;
    float eta_next = 0.0f;
    const int z0 = tc.z0, z1 = tc.z1;
    const unsigned cn4 = 4u * n;   // bytes between components

    // primary point of this thread: a tile column (tid < NOWN) or a ring point.
    // Tile threads fill whole waves, so `own` is uniform over a wave.  The
    // ring waves' spare lanes repeat ring points (the same values land at the
    // same places), so every lane has a point and no lane's work is skipped
    // by a branch.
    const int tid = threadIdx.x;
    const bool own = (int)uni((unsigned)tid) < NOWN;
    const int rr = (tid - NOWN) % NRING;
    int lx = 0, ly = 0;
    if (own) {
        lx = tid % TX; ly = tid / TX;
    } else {
        if (rr < TY) { lx = -1; ly = rr; }
        else if (rr < 2 * TY) { lx = TX; ly = rr - TY; }
        else if (rr < 2 * TY + TX) { lx = rr - 2 * TY; ly = -1; }
        else { lx = rr - 2 * TY - TX; ly = TY; }
    }
    const int gx = tc.x0 + lx, gy = tc.y0 + ly;
    if (own) { S.accE[tid] = 0; S.accG[tid] = 0; }
    // a finished group of planes into the thread's exact sums (tile threads
    // only; `own` is uniform over a wave)
    auto flush = [&](u128 *acc, double &grp, unsigned huge, unsigned nan) {
        if (own) {
            u128 a = acc[tid];
            fix_add(a, grp, flags, huge, nan);
            acc[tid] = a;
        }
        grp = 0.0;
    };
    const bool pin = gx >= 0 && gx < nx && gy >= 0 && gy < ny;   // point in the volume
    const bool col_interior = own && gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2;
    const unsigned pv = pin ? 4u * (unsigned)(gy * nx + gx) : 0u;
    // Every LDS access of a step is the point's cell in one of the planes plus
    // a constant: the W planes share the trial planes' pitch (their rows start
    // at ly = -1), so one address register serves them all.  A ring point
    // writes both relation axes into its own cells; the axis it does not serve
    // is never read (stage B reads W_{a,x} at rows 0 .. TY-1 and W_{a,y} at
    // columns 0 .. TX-1).
    const int pi = (ly + 3) * HX + (lx + 3);
    const int wi = pi - 2 * HX;

    // halo slots: a share of rings 2-3 and the ring-1 corners (slots past the
    // last point repeat points from the start)
    unsigned xv[SX];   // voffset
    int xd[SX];        // float index in a plane of cells
    unsigned xok = 0;
#pragma unroll
    for (int j = 0; j < SX; ++j) {
        const int o = (tid + j * NT) % (3 * NOTHER);
        const int c = o / NOTHER;
        int xx = 0, yy = 0;
        other_point<TX, TY>(o - c * NOTHER, xx, yy);
        const int x = tc.x0 - 3 + xx, y = tc.y0 - 3 + yy;
        const bool in = x >= 0 && x < nx && y >= 0 && y < ny;
        xv[j] = in ? 4u * ((unsigned)c * n + (unsigned)(y * nx + x)) : 0u;
        xd[j] = 4 * (yy * HX + xx) + c;
        xok |= (in ? 1u : 0u) << j;
    }

    // raw loads of one plane: the primary point's components (3) and the
    // halo slots; INTERP passes load base and candidate
    struct Raw { float a[SLOTS]; float b[INTERP ? SLOTS : 1]; };
    // (PART 1: the base buffer, PART 2: the candidate of an interpolated trial)
    auto fetch_part = [&](auto part, int gz, Raw &r) {
        constexpr int PART = decltype(part)::value;
        const int gzc = min(max(gz, 0), nz - 1);
        const unsigned so = 4u * (unsigned)gzc * plane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (PART == 1) r.a[c] = bload(r_src, pv, so + (unsigned)c * cn4);
            if (INTERP && PART == 2) r.b[c] = bload(r_cnd, pv, so + (unsigned)c * cn4);
        }
#pragma unroll
        for (int j = 0; j < SX; ++j) {
            if (PART == 1) r.a[3 + j] = bload(r_src, xv[j], so);
            if (INTERP && PART == 2) r.b[3 + j] = bload(r_cnd, xv[j], so);
        }
    };
    auto fetch = [&](int gz, Raw &r) {
        fetch_part(std::integral_constant<int, 1>{}, gz, r);
        fetch_part(std::integral_constant<int, 2>{}, gz, r);
    };
    // trial values of a plane (zero outside the volume)
    struct Val { float v[SLOTS]; };
    auto resolve = [&](int gz, const Raw &r, Val &o, bool known_in) {
        const bool zin = known_in || (gz >= 0 && gz < nz);
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            float v = r.a[k];
            if (INTERP) v = fma_ct(s, r.b[k] - v, v);
            const bool ok = zin && (k < 3 ? pin : ((xok >> (k - 3)) & 1u) != 0);
            o.v[k] = ok ? v : 0.0f;
        }
    };
    auto commit = [&](int slot, const Val &o) {
        S.xs[slot][pi] = f3{o.v[0], o.v[1], o.v[2]};
        float *dst = (float *)&S.xs[slot][0];
#pragma unroll
        for (int j = 0; j < SX; ++j) dst[xd[j]] = o.v[3 + j];
    };
    struct TM { float t; unsigned m; };
    auto fetch_tm = [&](int gz, TM &o) {
        const int gzc = min(max(gz, 0), nz - 1);
        o.t = bload(r_tgt, pv, 4u * (unsigned)gzc * plane);
        o.m = (unsigned)__builtin_amdgcn_raw_buffer_load_b8(r_msk, (int)(pv >> 2), (int)((unsigned)gzc * plane), 0);
    };

    // z-window of the primary point: planes zz-2 .. zz+2 at stage A of plane zz
    float col[5][3];
    Raw rn;          // raw loads of plane zz+2 in flight at the start of step zz
    TM tma, tmb;     // target / mask: plane zz in one set, plane zz+1 loading into the other

    // Steps run from zs to z1 in pairs: zs = z0 - 1, or z0 - 2 when that count
    // is odd (the extra first step's results are never used: stage A below z0
    // only feeds stage B of planes below z0, which is skipped).
    const int zs = ((z1 - z0) & 1) ? z0 - 2 : z0 - 1;
    // LDS ring slot of plane z: (z - zs) mod 3
    // prologue: planes zs and zs+1 committed (slots 0 and 1), plane zs+2 in
    // flight.  The loads of all five planes are issued at once, so the
    // prologue waits for one memory round trip, not one per plane; for a
    // direct trial that round trip overlaps the controller step.
    {
        Raw r0, r1, r2, r3;
        fetch(zs - 2, r0);
        fetch(zs - 1, r1);
        fetch(zs, r2);
        fetch(zs + 1, r3);
        fetch_tm(zs, tma);
        fetch(zs + 2, rn);
        begin();
        if (!INTERP) {
            if (!uni(C->active) || uni(C->mode) != TRIAL_DIRECT) return false;
            const int b = uni(C->src);
            if (b != b_src) {
                b_src = b;
                r_src = make_rsrc(uni(I.buf[b]), fbytes);
                fetch(zs - 2, r0);
                fetch(zs - 1, r1);
                fetch(zs, r2);
                fetch(zs + 1, r3);
                fetch(zs + 2, rn);
            }
        }
        {
            const int b_next = uni(C->out_next), b_trial = uni(C->out_trial);
            has_next = b_next >= 0;
            has_trial = INTERP && b_trial >= 0;
            if (has_next) r_next = make_rsrc(uni(I.buf[b_next]), fbytes);
            if (has_trial) r_trial = make_rsrc(uni(I.buf[b_trial]), fbytes);
            eta_next = uni(C->eta_next);
        }
        Val v;
        resolve(zs - 2, r0, v, false);
        col[1][0] = v.v[0]; col[1][1] = v.v[1]; col[1][2] = v.v[2];
        resolve(zs - 1, r1, v, false);
        col[2][0] = v.v[0]; col[2][1] = v.v[1]; col[2][2] = v.v[2];
        resolve(zs, r2, v, false);
        col[3][0] = v.v[0]; col[3][1] = v.v[1]; col[3][2] = v.v[2];
        commit(0, v);
        resolve(zs + 1, r3, v, false);
        col[4][0] = v.v[0]; col[4][1] = v.v[1]; col[4][2] = v.v[2];
        commit(1, v);
        col[0][0] = col[0][1] = col[0][2] = 0.0f;
    }
    __syncthreads();

    CT wz_m[3] = {0, 0, 0}, wz_c[3] = {0, 0, 0};
    const CT cix = (CT)P.inv[0], ciy = (CT)P.inv[1], ciz = (CT)P.inv[2];

    // Stage B's outputs (speculative candidate, and the trial itself after a
    // rejection) are stored at the start of the next step, before that step's
    // loads: the waits for a step's loads then find the stores long done
    // (loads and stores share one counter, which the compiler can only wait
    // to zero when both kinds are pending).
    // (the trial values of the pending plane are window slot 0 by then)
    struct Pending { bool on; float n[3]; };
    Pending pend;
    pend.on = false;
    // so: byte offset of the pending plane (wave-uniform)
    auto store_pending = [&](unsigned so) {
        if (pend.on) {
            if (has_next) {
                bstore(r_next, pend.n[0], pv, so);
                bstore(r_next, pend.n[1], pv, so + cn4);
                bstore(r_next, pend.n[2], pv, so + 2u * cn4);
            }
            if (has_trial) {
                bstore(r_trial, col[0][0], pv, so);
                bstore(r_trial, col[0][1], pv, so + cn4);
                bstore(r_trial, col[0][2], pv, so + 2u * cn4);
            }
        }
        pend.on = false;
    };

    // step zz: plane zz+2 joins the window and the LDS ring, stage A on plane
    // zz (ring slot sa), stage B on plane zz-1.  The W planes alternate by the
    // step's place in its trip: PZ receives plane zz, 1 - PZ holds plane zz-1.
    // MODE 0: planes inside the z-chunk and off the z faces (no tests);
    // MODE 1: stage A's plane off the z faces (2 <= zz <= nz - 3), chunk tests
    // kept; MODE 2: any plane (the z faces take the region map).  Keeping the
    // region map out of the first two keeps its registers (and their spills)
    // out of the steps most chunks run.
    auto step = [&](auto mode, auto pz, int zz, int sa, TM &cur, TM &nxt, const ColCoef &K) {
        constexpr int MODE = decltype(mode)::value;
        constexpr bool FAST = MODE == 0, OFFZ = MODE <= 1;
        constexpr int PZ = decltype(pz)::value;
        // Steps at the z faces rebuild their stencil coefficients from the
        // column (the column's own ones are not kept live through them), here,
        // before the step's loads are issued: a register the compiler reloads
        // from scratch for this then waits only for loads already due.
        const bool zint = OFFZ || (zz >= 2 && zz <= nz - 3);
        ColCoef KF;
        if (!OFFZ) {
            const StencilCoef cf = load_coef(I);
            KF = zint ? column_coefficients(P, cf, gx, gy) : face_coefficients(P, cf, gx, gy, zz);
        }
        // The loads of plane zz+2 and of this plane's target / mask were issued
        // one step ago.  An empty asm that takes them here keeps the compiler
        // from hoisting their first use above the barrier that ended the last
        // step, which would make that step wait for its own loads.
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            asm volatile("" : "+v"(rn.a[k]));
            if (INTERP) asm volatile("" : "+v"(rn.b[k]));
        }
        asm volatile("" : "+v"(cur.t));
        // stage A's neighbours on plane zz (ring slot sa), read before this
        // step's commit writes another slot: the reads then overlap the
        // commit, the stores and the loads below
        const f3 *R = &S.xs[sa][pi];
        const f3 xm2 = R[-2], xm1 = R[-1], xp1 = R[1], xp2 = R[2];
        const f3 ym2 = R[-2 * HX], ym1 = R[-HX], yp1 = R[HX], yp2 = R[2 * HX];
        {
            Val vz;
            resolve(zz + 2, rn, vz, OFFZ);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                col[0][c] = col[1][c]; col[1][c] = col[2][c]; col[2][c] = col[3][c]; col[3][c] = col[4][c];
                col[4][c] = vz.v[c];
            }
            commit(sa == 0 ? 2 : sa - 1, vz);
        }
        store_pending(4u * (unsigned)max(zz - 2, 0) * plane);   // plane zz-2, stage B of the last step
        // an interpolated trial's candidate plane is requested after stage A,
        // where the most registers are live
        fetch_part(std::integral_constant<int, 1>{}, zz + 3, rn);
        fetch_tm(zz + 1, nxt);

        // (A) plane zz at the primary point (computed at points outside the
        // volume too: nothing reads their W, their error is not counted)
        CT wz_p[3] = {0, 0, 0};
        if (OFFZ || (zz >= 0 && zz < nz)) {
            float sm[3][3][4];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                sm[c][0][0] = xm2[c]; sm[c][0][1] = xm1[c]; sm[c][0][2] = xp1[c]; sm[c][0][3] = xp2[c];
                sm[c][1][0] = ym2[c]; sm[c][1][1] = ym1[c]; sm[c][1][2] = yp1[c]; sm[c][1][3] = yp2[c];
                sm[c][2][0] = col[0][c]; sm[c][2][1] = col[1][c]; sm[c][2][2] = col[3][c]; sm[c][2][3] = col[4][c];
            }
            const bool count = own && pin && (FAST || (zz >= z0 && zz < z1));
            CT W[9];
            if (zint) stream_stage_a<CT, false>(P, OFFZ ? K : KF, sm, col[2], cur.t, cur.m != 0, count, Egrp, mx, W);
            else stream_stage_a<CT, true>(P, KF, sm, col[2], cur.t, cur.m != 0, count, Egrp, mx, W);
            // (a ring point also writes the relation axis it does not serve: never read)
            S.w[PZ][0][wi] = f3{(float)W[0], (float)W[1], (float)W[2]};
            S.w[PZ][1][wi] = f3{(float)W[3], (float)W[4], (float)W[5]};
#pragma unroll
            for (int a = 0; a < 3; ++a) wz_p[a] = (CT)(float)W[6 + a];
        }
        if (FAST ? (zz & 3) == 3 : (zz >= z0 && zz < z1 && ((zz & 3) == 3 || zz == nz - 1)))
            flush(S.accE, Egrp, ACC_E_HUGE, ACC_E_NAN);

        if (INTERP) fetch_part(std::integral_constant<int, 2>{}, zz + 3, rn);

        // (B) plane zb = zz-1 at tile columns: W_{a,x}, W_{a,y} of plane zb
        // (written last step), W_{a,z} of planes zb-1 and zb+1 (registers)
        const int zb = zz - 1;
        if (own && (FAST || (zb >= z0 && zb < z1))) {
            float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
            if (OFFZ || (zb >= 1 && zb <= nz - 2)) {
                const f3 xl = S.w[1 - PZ][0][wi - 1], xr = S.w[1 - PZ][0][wi + 1];
                const f3 yl = S.w[1 - PZ][1][wi - HX], yr = S.w[1 - PZ][1][wi + HX];
                const CT dxx = (CT)xl[0] - (CT)xr[0];
                const CT dyx = (CT)xl[1] - (CT)xr[1];
                const CT dzx = (CT)xl[2] - (CT)xr[2];
                const CT dxy = (CT)yl[0] - (CT)yr[0];
                const CT dyy = (CT)yl[1] - (CT)yr[1];
                const CT dzy = (CT)yl[2] - (CT)yr[2];
                // the gradient is zero on the outer face (and unused outside the volume)
                f0 = col_interior ? (float)fma_ct(cix, dxx, fma_ct(-ciy, dxy, ciz * (wz_m[0] - wz_p[0]))) : 0.0f;
                f1 = col_interior ? (float)fma_ct(-cix, dyx, fma_ct(ciy, dyy, -ciz * (wz_m[1] - wz_p[1]))) : 0.0f;
                f2 = col_interior ? (float)fma_ct(cix, dzx, fma_ct(-ciy, dzy, ciz * (wz_m[2] - wz_p[2]))) : 0.0f;
                Ggrp += fma_ct((double)f0, (double)f0, fma_ct((double)f1, (double)f1, (double)f2 * f2));
            }
            // trial value of this voxel on plane zb: window slot 1 (planes zz-2 .. zz+2);
            // stored at the start of the next step
            const float ux = col[1][0], uy = col[1][1], uz = col[1][2];
            pend.on = pin;
            pend.n[0] = fma_ct(-eta_next, f0, ux);
            pend.n[1] = fma_ct(-eta_next, f1, uy);
            pend.n[2] = fma_ct(-eta_next, f2, uz);
        }
        if (FAST ? (zb & 3) == 3 : (zb >= z0 && zb < z1 && ((zb & 3) == 3 || zb == nz - 1)))
            flush(S.accG, Ggrp, ACC_G_HUGE, ACC_G_NAN);
#pragma unroll
        for (int a = 0; a < 3; ++a) { wz_m[a] = wz_c[a]; wz_c[a] = wz_p[a]; }
        __syncthreads();
    };

    // two steps per trip, so the target / mask sets swap roles without a copy;
    // the trip is not unrolled further (more steps in flight would raise the
    // live registers)
    int sa = 0;
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    auto trip = [&](auto mode, int zz, const ColCoef &K) {
        step(mode, P0{}, zz, sa, tma, tmb, K);
        sa = sa == 2 ? 0 : sa + 1;
        step(mode, P1{}, zz + 1, sa, tmb, tma, K);
        sa = sa == 2 ? 0 : sa + 1;
    };
    // Three phases: trips touching the z faces below, trips off the z faces
    // (those inside the chunk with their tests folded away, the others with
    // the chunk tests only), trips touching the z faces above.  The column's
    // coefficients live through the middle phase alone.
    int zz = zs;
#pragma unroll 1
    for (; zz < z1 && zz < 2; zz += 2) trip(std::integral_constant<int, 2>{}, zz, ColCoef{});
    {
        const ColCoef K = column_coefficients(P, load_coef(I), gx, gy);
        const int flo = max(z0 + 1, 2), fhi = min(z1 - 1, nz - 3);
#pragma unroll 1
        for (; zz < z1 && zz + 1 <= nz - 3; zz += 2) {
            if (zz >= flo && zz + 1 <= fhi) trip(std::integral_constant<int, 0>{}, zz, K);
            else trip(std::integral_constant<int, 1>{}, zz, K);
        }
    }
#pragma unroll 1
    for (; zz < z1; zz += 2) trip(std::integral_constant<int, 2>{}, zz, ColCoef{});
    // the last step's outputs: shift the window as a next step would
#pragma unroll
    for (int c = 0; c < 3; ++c) col[0][c] = col[1][c];
    store_pending(4u * (unsigned)max(z1 - 1, 0) * plane);
    // a chunk that does not end on a group boundary (not planned so) keeps
    // its remainder (adding a zero group changes nothing)
    if (Egrp != 0.0) flush(S.accE, Egrp, ACC_E_HUGE, ACC_E_NAN);
    if (Ggrp != 0.0) flush(S.accG, Ggrp, ACC_G_HUGE, ACC_G_NAN);
    return true;
}

template <typename CT, int TX, int TY, int MINW>
__global__ void __launch_bounds__((StreamGeom<TX, TY>::NL), MINW)
k_grad_stream(const Inst *__restrict__ insts, const Tile *__restrict__ tiles, int parity)
{
    constexpr int NT = StreamGeom<TX, TY>::NT;
    __shared__ StreamSmem<TX, TY> S;

    using G = StreamGeom<TX, TY>;
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const Inst &I = insts[tl.inst];
    TileCoord tc = decode_tile<TX, TY>(I, tl);
    tc.x0 = uni(tc.x0); tc.y0 = uni(tc.y0); tc.z0 = uni(tc.z0); tc.z1 = uni(tc.z1);
    unsigned flags = 0;
    float mx = 0.0f;
    bool done = false;
    const Ctrl *C = &S.ctrl;
    if (G::CW) {
        if (threadIdx.x >= NT) {   // the controller wave: its step, the barrier that publishes it, done
            pass_begin<DT_GRADIENT>(I, parity, leader_tile(tl), &S.ctrl);
            __syncthreads();
            return;
        }
        // This pass's control copy before the controller step (no block
        // writes it during the pass).  The trial a pass evaluates is most
        // often the candidate the previous pass wrote, the next trial after an
        // accepted one (the first pass: the initial field).
        const Ctrl *cg = I.ctrl + parity;
        int guess = uni(cg->evaluated) ? uni(cg->out_next) : uni(cg->src);
        guess = (guess >= 0 && guess < NBUF) ? guess : 0;
        if (uni(cg->active)) done = stream_pass_body<CT, TX, TY, false>(S, I, tc, guess, [] { __syncthreads(); },
                                                                         flags, mx);
        else __syncthreads();
    } else {
        if (threadIdx.x < 64) pass_begin<DT_GRADIENT>(I, parity, leader_tile(tl), &S.ctrl);
        __syncthreads();
        if (uni(C->active) && uni(C->mode) == TRIAL_DIRECT)
            done = stream_pass_body<CT, TX, TY, false>(S, I, tc, uni(C->src), [] {}, flags, mx);
    }
    if (!done) {
        if (!uni(C->active)) {
            if (uni(C->copy_now)) copy_result_tile<TX, TY, NT>(I, C, tc);
            return;
        }
        stream_pass_body<CT, TX, TY, true>(S, I, tc, 0, [] {}, flags, mx);
    }

    add_pass_sums_lds<NT, StreamGeom<TX, TY>::NOWN>(I, uni(C->acc_cur), S.accE, S.accG, mx, flags, S.red);
}

// One (precision, tile) variant per compilation of this file: the register
// allocation of a kernel changes when other instantiations share its
// translation unit, so each variant is built alone (DT_STREAM_CT, DT_STREAM_TX,
// DT_STREAM_TY, DT_STREAM_MINW -- waves per SIMD the registers are capped
// for -- and DT_STREAM_NAME set by the build).
#if defined(DT_STREAM_CT) && defined(DT_STREAM_TX) && defined(DT_STREAM_TY) && defined(DT_STREAM_NAME) && \
    defined(DT_STREAM_MINW)
#define DT_PASTE2(a, b) a##b
#define DT_PASTE(a, b) DT_PASTE2(a, b)

hipError_t DT_PASTE(stream_launch_, DT_STREAM_NAME)(const Inst *insts, const Tile *tiles, int nblocks, int parity,
                                                     hipStream_t st)
{
    constexpr int NL = StreamGeom<DT_STREAM_TX, DT_STREAM_TY>::NL;
    hipLaunchKernelGGL((k_grad_stream<DT_STREAM_CT, DT_STREAM_TX, DT_STREAM_TY, DT_STREAM_MINW>), dim3(nblocks),
                       dim3(NL), 0, st, insts, tiles, parity);
    return hipGetLastError();
}

int DT_PASTE(stream_occupancy_, DT_STREAM_NAME)()
{
    constexpr int NL = StreamGeom<DT_STREAM_TX, DT_STREAM_TY>::NL;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, k_grad_stream<DT_STREAM_CT, DT_STREAM_TX, DT_STREAM_TY, DT_STREAM_MINW>, NL, 0) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}
#else
#error "dt_grad_stream.hip is compiled once per variant (see disptools_amd/build.py)"
#endif

}  // namespace dt
