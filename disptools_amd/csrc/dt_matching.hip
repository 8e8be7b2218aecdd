// dt_matching.hip -- the `matching` algorithm (Karacali & Davatzikos volume
// matching, src/VolumeMatching3D.c) on the GPU.
//
// Nodes (i, j, k) carry their deformed position in voxel units, F = y + 1 +
// u_y / dy, G = x + 1 + u_x / dx, E = z + 1 + u_z / dz (volume_matching_3d,
// :689-761).  Each cell (the cube of nodes i..i+1, j..j+1, k..k+1) has eight
// corner Jacobians -- at corner q the determinant of the matrix whose rows are
// F, G, E and whose columns are their forward differences along the cell
// edges through q, in the order y, x, z (:119-145) -- and the target V of its
// lowest node.  Cost (VolumeMatchingCost3D, :81-189):
//     sum over cells of 0.5 ([V > 0] (S - 8 V)^2 + sum_q [J_q < delta] zeta (J_q - delta)^2),
// S = sum of the corner Jacobians; max voxel error = max over cells of
// |S - 8 V| / 8.  Gradient (VolumeMatchingGrads3D, :191-406): each cell gives
// each of its nodes W_q times the cofactor of the matrix entry the node enters,
// W_q = [V > 0] (S - 8 V) + [J_q < delta] zeta (J_q - delta), for the corner
// q at the node and its three edge neighbours.
//
// One kernel, k_match_pass, runs either pass over z-chunked column tiles; the
// control block says which.  A cost pass evaluates the trial X - alpha D (or
// X itself for the initial evaluation), stores it (it becomes the accepted
// point if the step is taken) with its cell weights [V > 0](S - 8 V), and
// sums the cost in FP64 with the max voxel error; a gradient pass forms D at
// the accepted point from those weights.  The last block of every pass folds
// the per-tile partials in tile order and runs the step-length control of
// VolumeMatching3D (:441-545): start at step 1; on the first iteration
// double the step while the cost keeps falling and take the last improving
// trial; later iterations take an improving step and grow it by 1.05;
// otherwise halve it until the cost no longer rises or the step falls to
// 1e-9.  The loop runs while iterations < it_max, step > 1e-9 and max voxel
// error > epsilon, the error being the one of the last evaluated trial (as in
// the reference, where it is a global set by every cost call).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "dt_host.hpp"
#include "dt_match.hpp"

namespace dt {
namespace {

constexpr int MX = 32, MY = 32, MNT = MX * MY;    // cell tile, one thread per cell
constexpr int NPX = MX + 1, NPP = NPX * (MY + 1);  // nodes of a cell tile's plane
constexpr int NEXTRA = NPP - MNT;                  // node points loaded by a second slot
static_assert(NEXTRA <= MNT, "one extra point per thread at most");

enum { MM_FIRST = 0, MM_GROW = 1, MM_SHRINK = 2 };

// Instance record, control block and arrival counter of a matching solve,
// staged by one copy.
struct alignas(256) MHead {
    MInst inst;
    MCtrl ctrl;
    unsigned counter;
};

// Matrix columns: 0 = difference along y, 1 = along x, 2 = along z; corner
// bits: 1 = +x, 2 = +y, 4 = +z.
constexpr int COL_BIT[3] = {2, 1, 4};
// the reference sums the corner terms in the order of its names J_abc,
// a = z, b = x, c = y (VolumeMatching3D.c:151-160)
constexpr int REF_ORDER[8] = {0, 2, 1, 3, 4, 6, 5, 7};

__device__ __forceinline__ void corner_matrix(const float (&v)[3][8], int q, float (&m)[3][3])
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) m[r][c] = v[r][q | COL_BIT[c]] - v[r][q & ~COL_BIT[c]];
}

__device__ __forceinline__ float det3(const float (&m)[3][3])
{
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// cofactor (r, c) of m
__device__ __forceinline__ float cof(const float (&m)[3][3], int r, int c)
{
    const int r1 = r == 0 ? 1 : 0, r2 = r == 2 ? 1 : 2;
    const int c1 = c == 0 ? 1 : 0, c2 = c == 2 ? 1 : 2;
    const float minor = m[r1][c1] * m[r2][c2] - m[r1][c2] * m[r2][c1];
    return ((r + c) & 1) ? -minor : minor;
}

// x - a d, rounded as the reference's C (no FMA; this file is compiled with
// -ffp-contract=fast-honor-pragmas)
__device__ __forceinline__ float step_value(float x, float a, float d)
{
#pragma clang fp contract(off)
    return x - a * d;
}

__device__ __forceinline__ long long node_index(int nx, int ny, int i, int j, int k)
{
    return ((long long)k * ny + j) * nx + i;
}

// ---------------------------------------------------------------------------
// Passes
// ---------------------------------------------------------------------------

struct CostSmem {
    float s[3][3][NPP];        // trial node planes, ring of three
};
struct GradSmem {
    float x[2][3][NPP];        // accepted node planes, ring of two
    float c[8][3][MNT];        // each cell's contributions to its 8 nodes
};
union PassSmem {
    CostSmem c;
    GradSmem g;
};

// Node points of a thread: its own (lx, ly) and, for the first NEXTRA
// threads, one of the row MY / column MX that completes the (MX+1) x (MY+1)
// node plane of the cell tile.
__device__ __forceinline__ void node_points(int tid, int (&px)[2], int (&py)[2], bool &has1)
{
    px[0] = tid % MX;
    py[0] = tid / MX;
    has1 = tid < NEXTRA;
    if (tid < MX + 1) { px[1] = tid; py[1] = MY; }
    else { px[1] = MX; py[1] = tid - (MX + 1); }
}

// Cost pass over the cells of a tile whose lowest cell is (x0, y0); stores the
// trial at the tile's nodes [x0, x0+MX) x [y0, y0+MY) and planes [z0, z1).
__device__ void cost_pass(const MInst &I, const MCtrl &C, int x0, int y0, int z0, int z1, PassSmem &S,
                          double &Hs, float &mx)
{
    const int nx = I.nx, ny = I.ny, nz = I.nz;
    const long long n = I.n;
    const bool base = C.phase == MPH_INIT;
    const float *X = I.tri[C.xb];
    float *T = base ? nullptr : I.tri[C.tb];
    float *W = I.w[base ? C.xb : C.tb];
    const float *D = I.D;
    const float alpha = C.alpha;
    const int tid = threadIdx.x;
    const int lx = tid % MX, ly = tid / MX;
    int px[2], py[2];
    bool has1;
    node_points(tid, px, py, has1);
    bool ok[2];
    long long nid[2];
    int li[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int gx = x0 + px[k], gy = y0 + py[k];
        ok[k] = (k == 0 || has1) && gx < nx && gy < ny;
        nid[k] = ok[k] ? node_index(nx, ny, gx, gy, 0) : 0;
        li[k] = py[k] * NPX + px[k];
    }
    const int gx = x0 + lx, gy = y0 + ly;
    const bool cell_ok = gx < nx - 1 && gy < ny - 1;
    const long long plane = (long long)nx * ny;
    const int pend = min(z1, nz - 1);
    // raw loads of the next node plane
    float xr[2][3], dr[2][3];
    auto fetch = [&](int p) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const long long id = c * n + nid[k] + p * plane;
                xr[k][c] = ok[k] ? X[id] : 0.0f;
                dr[k][c] = (ok[k] && !base) ? D[id] : 0.0f;
            }
    };
    fetch(z0);
    for (int p = z0; p <= pend; ++p) {
        float(*dst)[NPP] = S.c.s[p % 3];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (!(k == 0 || has1)) continue;
            float v[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = base ? xr[k][c] : step_value(xr[k][c], alpha, dr[k][c]);
            if (k == 0 && T && ok[0] && p < z1) {
                const long long id = nid[0] + p * plane;
#pragma unroll
                for (int c = 0; c < 3; ++c) T[c * n + id] = v[c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) dst[c][li[k]] = v[c];
        }
        if (p < pend) fetch(p + 1);
        __syncthreads();
        const int kc = p - 1;
        if (kc >= z0 && cell_ok) {
            const float(*lo)[NPP] = S.c.s[kc % 3];
            const float(*hi)[NPP] = S.c.s[p % 3];
            float v[3][8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int idx = (ly + ((q >> 1) & 1)) * NPX + lx + (q & 1);
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c][q] = (q & 4) ? hi[c][idx] : lo[c][idx];
            }
            float J[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                float m[3][3];
                corner_matrix(v, q, m);
                J[q] = det3(m);
            }
            float Ssum = 0.0f;
#pragma unroll
            for (int t = 0; t < 8; ++t) Ssum += J[REF_ORDER[t]];
            const long long cid = node_index(nx, ny, gx, gy, kc);
            const float cv = I.V[cid];
            const float r = Ssum - 8.0f * cv;
            float cres = cv > 0.0f ? r * r : 0.0f;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const float Jq = J[REF_ORDER[t]];
                if (Jq < I.delta) cres += I.zeta * (Jq - I.delta) * (Jq - I.delta);
            }
            Hs += 0.5 * (double)cres;
            mx = fmaxf(mx, fabsf(r) * 0.125f);
            W[cid] = cv > 0.0f ? (float)((double)Ssum - 8.0 * (double)cv) : 0.0f;
        }
    }
}

// Gradient pass.  The tile's cells start at (cx0, cy0) = (tile * (MX-1) - 1,
// ...), one thread per cell; a thread with lx, ly >= 1 owns node (cx0+lx,
// cy0+ly), whose in-plane cells are the threads (lx-1..lx, ly-1..ly).  Per
// cell plane p the cell's corners give their weighted cofactors to the cell's
// 8 nodes (each corner's matrix and cofactors computed once); the node on
// plane p sums what the cells of planes p-1 (kept in a register) and p give
// it, in cell order.
__device__ void grad_pass(const MInst &I, const MCtrl &C, int cx0, int cy0, int z0, int z1, PassSmem &S)
{
    const int nx = I.nx, ny = I.ny, nz = I.nz;
    const long long n = I.n;
    const float *X = I.tri[C.xb];
    const float *Wc = I.w[C.xb];
    float *D = I.D;
    const float delta = I.delta, zeta = I.zeta;
    const int tid = threadIdx.x;
    const int lx = tid % MX, ly = tid / MX;
    const long long plane = (long long)nx * ny;
    int px[2], py[2];
    bool has1;
    node_points(tid, px, py, has1);
    bool ok[2];
    long long nid[2];
    int li[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int gx = cx0 + px[k], gy = cy0 + py[k];
        ok[k] = (k == 0 || has1) && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        nid[k] = ok[k] ? node_index(nx, ny, gx, gy, 0) : 0;
        li[k] = py[k] * NPX + px[k];
    }
    const int ci = cx0 + lx, cj = cy0 + ly;
    const bool cell_in = ci >= 0 && ci < nx - 1 && cj >= 0 && cj < ny - 1;
    const long long cid = cell_in ? node_index(nx, ny, ci, cj, 0) : 0;
    const bool owner = lx >= 1 && ly >= 1 && ci < nx && cj < ny;

    float xr[2][3];
    auto fetch_x = [&](int p) {
        const bool zin = p >= 0 && p < nz;
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                xr[k][c] = (zin && ok[k]) ? X[c * n + nid[k] + (long long)min(max(p, 0), nz - 1) * plane] : 0.0f;
    };
    auto fetch_w = [&](int p) -> float {
        return (cell_in && p >= 0 && p < nz - 1) ? Wc[cid + (long long)min(max(p, 0), nz - 1) * plane] : 0.0f;
    };
    auto put_x = [&](int p) {
        float(*dst)[NPP] = S.g.x[(p + 2) & 1];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (!(k == 0 || has1)) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c) dst[c][li[k]] = xr[k][c];
        }
    };
    // node plane z0-1 into LDS; node plane z0 and the weights of cell plane
    // z0-1 in flight
    fetch_x(z0 - 1);
    put_x(z0 - 1);
    fetch_x(z0);
    float wnext = fetch_w(z0 - 1);
    float up[3] = {0.0f, 0.0f, 0.0f};
    for (int p = z0 - 1; p < z1; ++p) {
        put_x(p + 1);
        const float wcell = wnext;
        if (p + 1 < z1) {
            fetch_x(p + 2);
            wnext = fetch_w(p + 1);
        }
        __syncthreads();
        // contributions of cell (ci, cj, p) to its 8 nodes
        float acc[8][3];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q][0] = acc[q][1] = acc[q][2] = 0.0f;
        if (cell_in && p >= 0 && p < nz - 1) {
            float v[3][8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float(*xs)[NPP] = S.g.x[(p + 2 + ((q >> 2) & 1)) & 1];
                const int idx = (ly + ((q >> 1) & 1)) * NPX + lx + (q & 1);
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c][q] = xs[c][idx];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                float m[3][3], cf[3][3];
                corner_matrix(v, q, m);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) cf[r][c] = cof(m, r, c);
                const float J = m[0][0] * cf[0][0] + m[0][1] * cf[0][1] + m[0][2] * cf[0][2];
                const float Wq = wcell + (J < delta ? zeta * (J - delta) : 0.0f);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int b = COL_BIT[c];
                    const float sgn = (q & b) ? 1.0f : -1.0f;
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        const float t = sgn * (Wq * cf[r][c]);
                        acc[q][r] += t;
                        acc[q ^ b][r] -= t;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int r = 0; r < 3; ++r) S.g.c[q][r][tid] = acc[q][r];
        __syncthreads();
        if (owner) {
            // cells (lx-1, ly-1), (lx, ly-1), (lx-1, ly), (lx, ly): the node is
            // their corner 3, 2, 1, 0 (+4 for the plane above)
            const int cell[4] = {tid - MX - 1, tid - MX, tid - 1, tid};
            if (p >= z0) {
                float d[3] = {up[0], up[1], up[2]};
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int r = 0; r < 3; ++r) d[r] += S.g.c[3 - t][r][cell[t]];
                const long long id = node_index(nx, ny, ci, cj, p);
#pragma unroll
                for (int r = 0; r < 3; ++r) D[r * n + id] = d[r];
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) up[r] = 0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 3; ++r) up[r] += S.g.c[7 - t][r][cell[t]];
        }
    }
}

// ---------------------------------------------------------------------------
// Control (one thread of the last block)
// ---------------------------------------------------------------------------

__device__ void m_finish(const MInst &I, MCtrl *C)
{
    C->phase = MPH_DONE;
    C->active = 0;
    C->reason = C->it >= I.it_max ? DT_STOP_MAX_ITERATIONS
                                  : (!((double)C->alpha > 1e-9) ? DT_STOP_ETA_SMALL : DT_STOP_TOLERANCE);
}

__device__ void m_loop_check(const MInst &I, MCtrl *C)
{
    if (C->it < I.it_max && (double)C->alpha > 1e-9 && C->m > I.epsilon) C->phase = MPH_GRAD;
    else m_finish(I, C);
}

__device__ void m_accept(const MInst &I, MCtrl *C, int t, double Hn)
{
    if (!((double)C->alpha > 1e-9)) { m_finish(I, C); return; }
    C->xb = t;
    C->H = Hn;
    C->it++;
    if (I.trace && C->trace_len < I.trace_cap) {
        double *row = I.trace + 4 * C->trace_len;
        row[0] = (double)C->it; row[1] = C->H; row[2] = C->m; row[3] = C->alpha;
    }
    C->trace_len++;
    m_loop_check(I, C);
}

__device__ __forceinline__ int third(int a, int b)
{
    for (int t = 0; t < 3; ++t)
        if (t != a && t != b) return t;
    return 0;
}

__device__ void match_controller(const MInst &I, MCtrl *C, double nH, float nm)
{
    if (C->phase == MPH_GRAD) {
        C->grad_passes++;
        C->Hmark = C->H;
        C->mode = MM_FIRST;
        C->tb = third(C->xb, C->xb == 0 ? 1 : 0);
        C->phase = MPH_COST;
        return;
    }
    C->cost_passes++;
    C->m = nm;
    if (C->phase == MPH_INIT) {
        C->H = nH;
        m_loop_check(I, C);
        return;
    }
    switch (C->mode) {
    case MM_FIRST:
        if (nH < C->Hmark) {
            if (C->it == 0) {
                // grow the step while the cost keeps falling (:459-471)
                C->Hmark = nH;
                C->alpha = C->alpha * 2.0f;
                C->pb = C->tb;
                C->tb = third(C->xb, C->pb);
                C->mode = MM_GROW;
                return;
            }
            C->alpha = (float)((double)C->alpha * 1.05);
            m_accept(I, C, C->tb, nH);
            return;
        }
        if (nH > C->Hmark && (double)C->alpha > 1e-9) {
            C->alpha = C->alpha * 0.5f;
            C->mode = MM_SHRINK;
            return;
        }
        m_accept(I, C, C->tb, nH);
        return;
    case MM_GROW:
        if (nH < C->Hmark) {
            C->Hmark = nH;
            C->alpha = C->alpha * 2.0f;
            const int t = C->pb;
            C->pb = C->tb;
            C->tb = t;
            return;
        }
        // the previous trial was the last improving one (:472-480)
        C->alpha = C->alpha * 0.5f;
        m_accept(I, C, C->pb, C->Hmark);
        return;
    default:   // MM_SHRINK (:487-499)
        if (nH > C->Hmark && (double)C->alpha > 1e-9) {
            C->alpha = C->alpha * 0.5f;
            return;
        }
        m_accept(I, C, C->tb, nH);
        return;
    }
}

// One block of a pass over tile t of instance I, whose passes launch
// `nblocks` blocks: the tile's part of the cost or gradient pass, its partial
// sums, and in the last block to arrive the fold of all partials (tile order)
// and the step control.
__device__ __forceinline__ void match_block(const MInst &I, int t, int nblocks, PassSmem &S, double *red, int &last)
{
    MCtrl *C = I.ctrl;
    if (!C->active) return;
    const MCtrl ctl = *C;
    const bool grad = ctl.phase == MPH_GRAD;
    const PassGeom G = grad ? I.gg : I.cg;
    double Hs = 0.0;
    float mx = 0.0f;
    if (t < G.ntx * G.nty * G.ntz) {
        const int tx = t % G.ntx, ty = (t / G.ntx) % G.nty, tz = t / (G.ntx * G.nty);
        const int z0 = tz * G.tz, z1 = min(z0 + G.tz, I.nz);
        if (grad) grad_pass(I, ctl, tx * (MX - 1) - 1, ty * (MY - 1) - 1, z0, z1, S);
        else cost_pass(I, ctl, tx * MX, ty * MY, z0, z1, S, Hs, mx);
    }

    // block partials, then the last block folds all tiles in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Hs += __shfl_xor(Hs, o, 64);
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[2 * wv] = Hs; red[2 * wv + 1] = (double)mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < MNT / 64; ++k) { a += red[2 * k]; b = fmax(b, red[2 * k + 1]); }
        __hip_atomic_store(&I.partials[2 * t], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&I.partials[2 * t + 1], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned prev = __hip_atomic_fetch_add(I.counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = prev + 1 == (unsigned)nblocks;
    }
    __syncthreads();
    if (!last) return;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += MNT) {
        a += __hip_atomic_load(&I.partials[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b = fmax(b, __hip_atomic_load(&I.partials[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b = fmax(b, __shfl_xor(b, o, 64));
    }
    __syncthreads();
    if (lane == 0) { red[2 * wv] = a; red[2 * wv + 1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int k = 0; k < MNT / 64; ++k) { sa += red[2 * k]; sb = fmax(sb, red[2 * k + 1]); }
        match_controller(I, C, sa, (float)sb);
        const bool done = C->phase == MPH_DONE;
        if (done && I.final_ctrl) {
            // the control block first, ordered before the word
            const unsigned long long *src = (const unsigned long long *)C;
            unsigned long long *dst = (unsigned long long *)I.final_ctrl;
            for (int k = 0; k < (int)(sizeof(MCtrl) / 8); ++k)
                __hip_atomic_store(dst + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __atomic_thread_fence(__ATOMIC_RELEASE);
        }
        const unsigned long long passes = (unsigned long long)(C->cost_passes + C->grad_passes);
        if (I.status)
            __hip_atomic_store(I.status, (passes << 1) | (done ? 1ull : 0ull), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(I.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// A pass of one instance: blocks are dispatched to the 8 XCDs round robin,
// so each XCD gets a contiguous range of tiles (neighbouring tiles share an
// L2 and their halo loads hit it; the grid is a multiple of 8).
__global__ void __launch_bounds__(MNT)
k_match_pass(const MInst *__restrict__ pinst)
{
    __shared__ PassSmem S;
    __shared__ double red[2 * (MNT / 64)];
    __shared__ int last;
    const int per_xcd = (int)gridDim.x / 8;
    const int t = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    match_block(*pinst, t, (int)gridDim.x, S, red, last);
}

// A pass over the tiles of several instances (the engine's tile table: one
// entry per block, the instance's tiles 0 .. nblocks-1 each).
__global__ void __launch_bounds__(MNT)
k_match_pass_batch(const MInst *__restrict__ insts, const Tile *__restrict__ tiles)
{
    __shared__ PassSmem S;
    __shared__ double red[2 * (MNT / 64)];
    __shared__ int last;
    const Tile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;
    const MInst &I = insts[tl.inst];
    match_block(I, tl.tile, I.nblocks, S, red, last);
}

// field -> node positions (volume_matching_3d, :746-757)
__global__ void __launch_bounds__(256)
k_match_in(const float *__restrict__ field, float *__restrict__ X, int nx, int ny, int nz, float dx, float dy,
           float dz)
{
    const long long n = (long long)nx * ny * nz;
    for (long long v = blockIdx.x * 256ll + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long long)nx * ny));
        X[v] = (float)(j + 1) + field[n + v] / dy;
        X[n + v] = (float)(i + 1) + field[v] / dx;
        X[2 * n + v] = (float)(k + 1) + field[2 * n + v] / dz;
    }
}

// node positions -> field (:769-781)
__global__ void __launch_bounds__(256)
k_match_out(const MInst *__restrict__ inst, float *__restrict__ field, int nx, int ny, int nz, float dx, float dy,
            float dz)
{
    const long long n = (long long)nx * ny * nz;
    const float *__restrict__ X = inst->tri[__builtin_amdgcn_readfirstlane(inst->ctrl->xb)];
    for (long long v = blockIdx.x * 256ll + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long long)nx * ny));
        field[v] = (X[n + v] - (float)(i + 1)) * dx;
        field[n + v] = (X[v] - (float)(j + 1)) * dy;
        field[2 * n + v] = (X[2 * n + v] - (float)(k + 1)) * dz;
    }
}

std::vector<MHead *> g_mstage;   // pinned staging per device

unsigned grid_of(long long work)
{
    long long b = (work + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536));
}

// Tiling of a pass whose tiles cover sx x sy nodes in the plane: z-chunks as
// deep as keep all tiles within one round of resident blocks (measured at
// 128^3 against deeper chunks with one block per CU: 18-19 G vs 17.5 G
// voxel-iters/s).  DISPTOOLS_MATCH_TZ fixes the depth.
PassGeom pass_geom(int nx, int ny, int nz, int sx, int sy, int resident)
{
    PassGeom g;
    g.ntx = (nx + sx - 1) / sx;
    g.nty = (ny + sy - 1) / sy;
    const long long cols = (long long)g.ntx * g.nty;
    const char *env = std::getenv("DISPTOOLS_MATCH_TZ");
    if (env && std::atoi(env) > 0) {
        g.tz = std::min(nz, std::atoi(env));
    } else {
        const long long chunks = std::max<long long>(1, std::min<long long>(nz, resident / cols));
        g.tz = (int)((nz + chunks - 1) / chunks);
    }
    g.ntz = (nz + g.tz - 1) / g.tz;
    return g;
}

// Per-device state of the matching entry points (used under solve_lock()):
// the occupancy of the pass kernel (queried once) and the timing events.
struct MatchDevice {
    int per_cu = 0;
    hipEvent_t kev[2] = {nullptr, nullptr};
};
std::vector<MatchDevice> g_mdev;

MatchDevice &match_device_state()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if ((int)g_mdev.size() <= dev) g_mdev.resize(dev + 1);
    return g_mdev[dev];
}

int resident_match_blocks()
{
    MatchDevice &md = match_device_state();
    if (md.per_cu < 1) {
        int per = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_match_pass, MNT, 0) != hipSuccess || per < 1)
            per = 1;
        md.per_cu = per;
    }
    return device_cus() * md.per_cu;
}

}  // namespace

// Matching solve on device buffers; `field` [3][nz][ny][nx] in/out.
static int matching_device(int nx, int ny, int nz, float dx, float dy, float dz, const float *d_target,
                           float epsilon, float delta, float zeta, long long it_max, float *d_field,
                           dt_matching_info *info, double *trace, long long trace_cap, hipStream_t st)
{
    const auto t_start = std::chrono::steady_clock::now();
    const MarkerRange range("dt_matching %dx%dx%d", nx, ny, nz);
    const long long n = (long long)nx * ny * nz;
    MInst h{};
    h.nx = nx; h.ny = ny; h.nz = nz;
    const int resident = resident_match_blocks();
    h.cg = pass_geom(nx, ny, nz, MX, MY, resident);
    h.gg = pass_geom(nx, ny, nz, MX - 1, MY - 1, resident);
    h.delta = delta; h.zeta = zeta; h.epsilon = epsilon;
    h.it_max = it_max < 0 ? 0 : it_max;
    h.n = n;
    const int ntiles = (std::max(h.cg.ntx * h.cg.nty * h.cg.ntz, h.gg.ntx * h.gg.nty * h.gg.ntz) + 7) / 8 * 8;
    h.nblocks = ntiles;
    h.final_ctrl = nullptr;
    if (trace_cap < 0 || !trace) trace_cap = 0;

    size_t bytes = 0;
    auto add = [&](size_t b) { bytes = align_up(bytes, 256) + b; };
    for (int t = 0; t < 3; ++t) add(12 * (size_t)n);
    for (int t = 0; t < 3; ++t) add(4 * (size_t)n);
    add(12 * (size_t)n);
    add(16 * (size_t)ntiles);
    add(sizeof(MHead));
    add(32 * (size_t)std::max<long long>(trace_cap, 1));
    void *base = nullptr;
    if (arena_device_buffer(bytes + 4096, &base)) return -1;
    Carver cv(base);
    for (int t = 0; t < 3; ++t) h.tri[t] = cv.take<float>(3 * (size_t)n);
    for (int t = 0; t < 3; ++t) h.w[t] = cv.take<float>((size_t)n);
    h.D = cv.take<float>(3 * (size_t)n);
    h.partials = cv.take<double>(2 * (size_t)ntiles);
    MHead *d_head = cv.take<MHead>(1);
    h.ctrl = &d_head->ctrl;
    h.counter = &d_head->counter;
    MInst *d_inst = &d_head->inst;
    h.trace = trace_cap ? cv.take<double>(4 * (size_t)trace_cap) : nullptr;
    h.trace_cap = trace_cap;
    h.V = d_target;

    unsigned long long *status = nullptr;
    if (status_word(&status, &h.status)) return -1;
    if (!status_words_from_env()) h.status = nullptr;

    int dev = 0;
    DT_HIP(hipGetDevice(&dev));
    if ((int)g_mstage.size() <= dev) g_mstage.resize(dev + 1, nullptr);
    MHead *&sh = g_mstage[dev];
    if (!sh) DT_HIP(hipHostMalloc((void **)&sh, sizeof(MHead), hipHostMallocDefault));
    // one copy from pinned staging (the previous solve on this device has
    // finished with it: solves end synchronised)
    sh->inst = h;
    sh->ctrl = MCtrl{};
    sh->ctrl.phase = MPH_INIT;
    sh->ctrl.active = 1;
    sh->ctrl.xb = 0;
    sh->ctrl.alpha = 1.0f;
    sh->counter = 0;
    DT_HIP(hipMemcpyAsync(d_head, sh, sizeof(MHead), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_match_in, dim3(grid_of(n)), dim3(256), 0, st, d_field, h.tri[0], nx, ny, nz, dx, dy, dz);
    DT_HIP(hipGetLastError());

    // timing events around the back-to-back pass launches
    MatchDevice &md = match_device_state();
    for (auto &e : md.kev)
        if (!e) DT_HIP(hipEventCreate(&e));
    hipEvent_t *kev = md.kev;
    // every iteration takes one gradient pass and at most ~160 cost passes
    // (the step doubles from 1 at most 128 times and halves at most 31 times
    // before falling under 1e-9)
    const long long pass_cap = 256 + 192 * (h.it_max + 1);
    long long launched = 0;
    const int chunk = 8;
    auto enqueue = [&]() -> int {
        for (int i = 0; i < chunk; ++i) {
            hipLaunchKernelGGL(k_match_pass, dim3(ntiles), dim3(MNT), 0, st, d_inst);
            DT_HIP(hipGetLastError());
        }
        launched += chunk;
        return 0;
    };
    // two chunks in flight, the host following the completion word that the
    // controller stores in host memory (no copy or event between the passes)
    int rc = hipEventRecord(kev[0], st) == hipSuccess ? 0 : -1;
    if (!rc) rc = enqueue();
    long long target = launched;
    if (!rc) rc = enqueue();
    while (!rc) {
        bool done = false;
        const int w = follow_status(status, target, st, done);
        if (w < 0) {
            set_error(__LINE__, __FILE__, __func__, "device error during the matching solve");
            rc = -1;
            break;
        }
        if (w > 0) {
            // the stream drained but the word did not follow: read the control block
            MCtrl c;
            if (hipMemcpy(&c, h.ctrl, sizeof(MCtrl), hipMemcpyDeviceToHost) != hipSuccess) {
                set_error(__LINE__, __FILE__, __func__, "device error during the matching solve");
                rc = -1;
                break;
            }
            done = c.phase == MPH_DONE;
        }
        if (done) break;
        if (launched > pass_cap) {
            set_error(__LINE__, __FILE__, __func__, "matching solve did not terminate");
            rc = -1;
            break;
        }
        target = launched;
        rc = enqueue();
    }
    const bool timed = hipEventRecord(kev[1], st) == hipSuccess;
    // positions -> field and the final control block, behind the passes
    hipError_t se = hipSuccess;
    if (!rc) {
        hipLaunchKernelGGL(k_match_out, dim3(grid_of(n)), dim3(256), 0, st, d_inst, d_field, nx, ny, nz, dx, dy, dz);
        se = hipGetLastError();
        if (se == hipSuccess) se = hipMemcpyAsync(&sh->ctrl, h.ctrl, sizeof(MCtrl), hipMemcpyDeviceToHost, st);
    }
    const hipError_t ss = hipStreamSynchronize(st);
    if (se == hipSuccess) se = ss;
    float kernel_ms = 0.0f;
    if (se != hipSuccess || !timed || hipEventElapsedTime(&kernel_ms, kev[0], kev[1]) != hipSuccess) kernel_ms = 0.0f;
    if (se != hipSuccess && !rc) { set_error(__LINE__, __FILE__, __func__, hipGetErrorString(se)); rc = -1; }
    if (rc) return -1;

    const MCtrl fin = sh->ctrl;
    if (trace && trace_cap > 0) {
        const long long nl = std::min(fin.trace_len, trace_cap);
        if (nl > 0) DT_HIP(hipMemcpy(trace, h.trace, (size_t)nl * 4 * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (info) {
        info->iterations = fin.it;
        info->reason = fin.reason;
        info->cost_passes = fin.cost_passes;
        info->gradient_passes = fin.grad_passes;
        info->launches = launched;
        info->final_cost = fin.H;
        info->final_max_error = fin.m;
        info->trace_len = fin.trace_len;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        info->kernel_seconds = (double)kernel_ms * 1e-3;
    }
    return 0;
}

size_t match_scratch_bytes(int nx, int ny, int nz)
{
    const long long n = (long long)nx * ny * nz;
    const int resident = resident_match_blocks();
    const PassGeom cg = pass_geom(nx, ny, nz, MX, MY, resident);
    const PassGeom gg = pass_geom(nx, ny, nz, MX - 1, MY - 1, resident);
    const int ntiles = (std::max(cg.ntx * cg.nty * cg.ntz, gg.ntx * gg.nty * gg.ntz) + 7) / 8 * 8;
    size_t bytes = 0;
    auto add = [&](size_t b) { bytes = align_up(bytes, 256) + b; };
    for (int t = 0; t < 3; ++t) add(12 * (size_t)n);
    for (int t = 0; t < 3; ++t) add(4 * (size_t)n);
    add(12 * (size_t)n);
    add(16 * (size_t)ntiles);
    return align_up(bytes, 256);
}

MInst match_make_inst(int nx, int ny, int nz, const float *d_target, float epsilon, float delta, float zeta,
                      long long it_max, Carver &cv, MCtrl *ctrl, unsigned *counter)
{
    MInst h{};
    const long long n = (long long)nx * ny * nz;
    h.nx = nx; h.ny = ny; h.nz = nz;
    const int resident = resident_match_blocks();
    h.cg = pass_geom(nx, ny, nz, MX, MY, resident);
    h.gg = pass_geom(nx, ny, nz, MX - 1, MY - 1, resident);
    h.nblocks = (std::max(h.cg.ntx * h.cg.nty * h.cg.ntz, h.gg.ntx * h.gg.nty * h.gg.ntz) + 7) / 8 * 8;
    h.delta = delta; h.zeta = zeta; h.epsilon = epsilon;
    h.it_max = it_max < 0 ? 0 : it_max;
    h.n = n;
    for (int t = 0; t < 3; ++t) h.tri[t] = cv.take<float>(3 * (size_t)n);
    for (int t = 0; t < 3; ++t) h.w[t] = cv.take<float>((size_t)n);
    h.D = cv.take<float>(3 * (size_t)n);
    h.partials = cv.take<double>(2 * (size_t)h.nblocks);
    h.V = d_target;
    h.ctrl = ctrl;
    h.counter = counter;
    h.trace = nullptr;
    h.trace_cap = 0;
    h.status = nullptr;
    h.final_ctrl = nullptr;
    return h;
}

MCtrl match_initial_ctrl()
{
    MCtrl c{};
    c.phase = MPH_INIT;
    c.active = 1;
    c.xb = 0;
    c.alpha = 1.0f;
    return c;
}

long long match_max_passes(long long it_max)
{
    // every iteration takes one gradient pass and at most ~160 cost passes
    // (the step doubles from 1 at most 128 times and halves at most 31 times
    // before falling under 1e-9)
    return 256 + 192 * ((it_max < 0 ? 0 : it_max) + 1);
}

hipError_t launch_match_in(const MInst &h, const float *d_field, float dx, float dy, float dz, hipStream_t st)
{
    hipLaunchKernelGGL(k_match_in, dim3(grid_of(h.n)), dim3(256), 0, st, d_field, h.tri[0], h.nx, h.ny, h.nz, dx,
                       dy, dz);
    return hipGetLastError();
}

hipError_t launch_match_out(const MInst &h, const MInst *d_inst, float *d_field, float dx, float dy, float dz,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_match_out, dim3(grid_of(h.n)), dim3(256), 0, st, d_inst, d_field, h.nx, h.ny, h.nz, dx, dy,
                       dz);
    return hipGetLastError();
}

hipError_t launch_match_batch(const MInst *insts, const Tile *tiles, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_match_pass_batch, dim3(nblocks), dim3(MNT), 0, st, insts, tiles);
    return hipGetLastError();
}

int matching_validate(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float epsilon, float delta,
                      float zeta)
{
    if (nx < 2 || ny < 2 || nz < 2) DT_FAIL("every image dimension must be at least 2 voxels");
    if (nx > (1u << 30) || ny > (1u << 30) || nz > (1u << 30)) DT_FAIL("image too large");
    // voxel count and the byte counts derived from it (16 B/voxel of staging,
    // 40 B/voxel of solver state) in checked arithmetic: no size_t product wraps
    size_t n = 0, bytes = 0;
    if (__builtin_mul_overflow(nx, ny, &n) || __builtin_mul_overflow(n, nz, &n) ||
        __builtin_mul_overflow(n, (size_t)64, &bytes) || n > ((size_t)1 << 40))
        DT_FAIL("image too large: at most 2^40 voxels per matching solve");
    if (!(dx > 0.0f) || !(dy > 0.0f) || !(dz > 0.0f)) DT_FAIL("spacing must be positive");
    if (!(epsilon > 0.0f)) DT_FAIL("epsilon must be positive");
    if (!(delta > 0.0f) || !(zeta > 0.0f)) DT_FAIL("delta and zeta must be positive");
    return 0;
}

}  // namespace dt

using namespace dt;

extern "C" int dt_matching_device(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                                  const float *d_target, float epsilon, float delta, float zeta, long long it_max,
                                  float *d_field, dt_matching_info *info, double *trace, long long trace_cap,
                                  void *stream)
{
    std::lock_guard<std::mutex> g(solve_lock());
    disptools_clear_error();
    if (matching_validate(nx, ny, nz, dx, dy, dz, epsilon, delta, zeta)) return -1;
    if (!d_target || !d_field) DT_FAIL("null buffer");
    return matching_device((int)nx, (int)ny, (int)nz, dx, dy, dz, d_target, epsilon, delta, zeta, it_max, d_field,
                           info, trace, trace_cap, (hipStream_t)stream);
}

extern "C" int dt_matching_host(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, const float *target,
                                float epsilon, float delta, float zeta, long long it_max, float *field,
                                dt_matching_info *info, double *trace, long long trace_cap)
{
    std::lock_guard<std::mutex> g(solve_lock());
    disptools_clear_error();
    if (matching_validate(nx, ny, nz, dx, dy, dz, epsilon, delta, zeta)) return -1;
    if (!target || !field) DT_FAIL("null buffer");
    const size_t n = nx * ny * nz;
    float *d = nullptr;
    if (hipMalloc((void **)&d, 16 * n) != hipSuccess) {
        (void)hipGetLastError();
        DT_FAIL("cannot allocate device memory for the matching inputs");
    }
    float *dT = d, *dF = d + n;
    hipStream_t st = nullptr;
    int rc = 0;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipMemcpyAsync(dT, target, 4 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dF, field, 12 * n, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error(__LINE__, __FILE__, __func__, "cannot stage the matching inputs");
        rc = -1;
    }
    if (!rc)
        rc = matching_device((int)nx, (int)ny, (int)nz, dx, dy, dz, dT, epsilon, delta, zeta, it_max, dF, info,
                             trace, trace_cap, st);
    if (!rc && (hipMemcpyAsync(field, dF, 12 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)) {
        set_error(__LINE__, __FILE__, __func__, "cannot copy the matching result back");
        rc = -1;
    }
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(d);
    return rc;
}

// The reference's matching entry point (src/headers/VolumeMatching3D.h:7-29,
// src/VolumeMatching3D.c:689-796) on the current HIP device: host buffers,
// the field in place.  Uses epsilon, delta, zeta and it_max; the other solver
// parameters and the mask are accepted and unused, as in the reference.
extern "C" void volume_matching_3d(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, const float *J,
                                   const bool *mask, float epsilon, float tolerance, float eta, float eta_max,
                                   float alpha, float beta, float gamma, float delta, float zeta, float theta,
                                   float iota, bool strict, size_t it_max, float *field)
{
    (void)mask; (void)tolerance; (void)eta; (void)eta_max; (void)alpha; (void)beta; (void)gamma;
    (void)theta; (void)iota; (void)strict;
    if (!verbose_from_env()) {
        (void)dt_matching_host(nx, ny, nz, dx, dy, dz, J, epsilon, delta, zeta, (long long)it_max, field, nullptr,
                               nullptr, 0);
        return;
    }
    // the reference prints one line per accepted iteration (VolumeMatching3D.c:538-544)
    const long long cap = std::min<long long>((long long)it_max + 1, 1ll << 21);
    std::vector<double> tr(4 * (size_t)std::max<long long>(cap, 1));
    dt_matching_info info{};
    if (dt_matching_host(nx, ny, nz, dx, dy, dz, J, epsilon, delta, zeta, (long long)it_max, field, &info,
                         tr.data(), cap) == 0) {
        const long long lines = std::min(info.trace_len, cap);
        for (long long k = 0; k < lines; ++k) print_iteration((long long)tr[4 * k], tr[4 * k + 1], tr[4 * k + 2],
                                                              tr[4 * k + 3]);
        std::fflush(stdout);
    }
}
