// dt_jacobian.hip -- the standalone Jacobian determinant det(I + grad u)
// over a volume (reference jacobian(), src/jacobian.c:70-242), streamed
// through z.
//
// A block is one wave and owns a 16 x 4 patch of columns, walking a z-chunk
// of it one plane at a time.  Each lane keeps its column's three field
// components for the planes z-2 .. z+5 in a register queue (the z-stencil
// reads z-2 .. z+2; loads are issued three steps before they are needed);
// the plane z of the patch and a halo of two voxels in x and y sit in the
// wave's LDS cells, from which the x- and y-stencils read.  The lanes load
// their halo cells three planes ahead too.  A wave writes and reads
// only its own LDS cells, so no step waits for another wave: waves of
// boundary patches (the branch-free stencil, about twice the instructions)
// never hold up interior ones, and a CU takes a new patch as soon as a wave
// ends.  The step is unrolled eight times so that the queue and the halo
// ring rotate by renaming registers, never by copying a register whose load
// is in flight.
//
// Arithmetic: samples are differenced in float, stencil weights and the
// determinant in FP64, operation by operation as the reference does
// (dstencil_c4 in waves of interior voxels, the branch-free dstencil_sel in
// the others; det_plus_identity; built with contraction off), so J is
// bit-identical to the reference C build.
//
// HBM traffic per voxel: 12 bytes read (three components), 4 written; the
// halo planes and columns a wave re-reads come mostly from L2 (patches of one
// XCD are neighbours, see jac_tile).
#include <math.h>

#include "dt_device.hpp"
#include "dt_kernels.hpp"
#include "dt_pass_common.hpp"

namespace dt {

namespace {

// A block is one wave; it owns a patch of PX x PY columns.
constexpr int PX = 16, PY = 4;
constexpr int HX = 48, HY = PY + 4;               // the patch with a 2-voxel halo: HX >= PX + 4,
                                                  // HX = 16 mod 32 so a wave's four rows use
                                                  // disjoint LDS banks
static_assert(HX >= PX + 4 && HX % 32 == 16, "LDS row stride");
constexpr int WPLANE = HX * HY;                   // LDS cells of one component
constexpr int WDUMMY = 3 * WPLANE;                // cell written by lanes with no halo value
constexpr int NHALO = 4 * PX + 4 * PY;            // halo cells: 2 rows each side, 2 columns each side
constexpr int HVALS = 3 * NHALO;                  // halo values (three components)
constexpr int HPL = (HVALS + 63) / 64;            // halo values per lane
constexpr int LEAD = 3;                           // loads are issued this many planes ahead
constexpr int JQ = 5 + LEAD;                      // queue: planes z-2 .. z+2+LEAD
constexpr int HR = LEAD + 1;                      // halo planes z .. z+LEAD in registers
static_assert(JQ % HR == 0, "the unrolled step rotates both rings");
static_assert(PX * PY == 64, "one voxel per lane");

struct JacGeom {
    const float *f;
    float *J;
    int nx, ny, nz;
    long long n, nxy;
    unsigned plane_bytes;         // 4 nx ny
    float ix, iy, iz;
    int ntx, cols, tz, ntiles, per;
};

// Tile of block b: blocks b, b + 8, ... run on one XCD and take a contiguous
// run of the tile list (patches fastest, x before y), so the halos they share
// stay in that XCD's L2.  -1: a spare block of the rounded-up grid.
__device__ __forceinline__ int jac_tile(const JacGeom &G, int b)
{
    const int t = (b & 7) * G.per + (b >> 3);
    return t < G.ntiles ? t : -1;
}

struct JacLane {
    int lx, ly, x, y;
    unsigned off;                 // y * nx + x, clamped into the plane
    unsigned voff[3];             // byte offsets of the column, per component
    unsigned hoff[HPL];           // byte offsets of the halo values (component included)
    int hcell[HPL];               // their LDS cells (WDUMMY: none)
};

// Halo value v (component v / NHALO) of the patch at (x0, y0): the element
// loaded for it (clamped into the volume) and its LDS cell.
__device__ __forceinline__ void halo_cell(const JacGeom &G, int v, int x0, int y0, unsigned &off, int &cell)
{
    const int comp = v / NHALO;
    const int r = v - comp * NHALO;
    int lx, ly;
    if (r < 4 * PX) {             // two rows below and two above the patch
        const int k = r / PX;
        lx = r - k * PX;
        ly = k < 2 ? k - 2 : PY + k - 2;
    } else {                      // two columns left and two right of it
        const int r2 = r - 4 * PX;
        const int k = r2 / PY;
        ly = r2 - k * PY;
        lx = k < 2 ? k - 2 : PX + k - 2;
    }
    const int x = min(max(x0 + lx, 0), G.nx - 1), y = min(max(y0 + ly, 0), G.ny - 1);
    off = 4u * (unsigned)(comp * G.n + (long long)y * G.nx + x);
    cell = comp * WPLANE + (ly + 2) * HX + lx + 2;
}

// Byte offset of plane z of a component plane set, z clamped into the
// volume (a field holds fewer than 2^32 bytes, so every offset is 32-bit).
// Cells outside the volume thus hold a copy of a sample inside it: the
// stencil of every voxel leaves them unweighted, and the loads need no
// branch or select (a select right after a load would make the wave wait for
// it at once, which the lead of the loads is there to avoid).
__device__ __forceinline__ unsigned plane_off(const JacGeom &G, int z)
{
    return (unsigned)min(max(z, 0), G.nz - 1) * G.plane_bytes;
}

// One plane step: plane z of the patch, queue slot (S + k) % JQ holding plane
// z - 2 + k, halo registers h[(S + k) % HR] plane z + k's halo.  The wave
// alone writes and reads its LDS cells, in program order: no barrier.
template <int S>
__device__ __forceinline__ void jac_step(const JacGeom &G, const JacLane &T, float *P, float (&q)[3][JQ],
                                         float (&h)[HR][HPL], int z, int z1, __amdgpu_buffer_rsrc_t rf,
                                         __amdgpu_buffer_rsrc_t rj)
{
    // loads LEAD steps ahead: plane z + 2 + LEAD of the column, z + LEAD of the halo
    const unsigned sq = plane_off(G, z + 2 + LEAD), sh = plane_off(G, z + LEAD);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c][(S + JQ - 1) % JQ] = bload(rf, T.voff[c], sq);
#pragma unroll
    for (int k = 0; k < HPL; ++k) h[(S + LEAD) % HR][k] = bload(rf, T.hoff[k], sh);
    // plane z into LDS
#pragma unroll
    for (int c = 0; c < 3; ++c) P[c * WPLANE + (T.ly + 2) * HX + T.lx + 2] = q[c][(S + 2) % JQ];
#pragma unroll
    for (int k = 0; k < HPL; ++k) P[T.hcell[k]] = h[S % HR][k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int kind[3];
    classify(T.x, T.y, z, G.nx, G.ny, G.nz, kind);
    const bool c4 = kind[0] == K_C4 && kind[1] == K_C4 && kind[2] == K_C4;
    double d[3][3];
    if (__builtin_amdgcn_ballot_w64(!c4) == 0) {
        // every lane of the wave is interior: 4th-order central on all axes
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *row = P + c * WPLANE + (T.ly + 2) * HX + T.lx;
            const float *clm = P + c * WPLANE + T.ly * HX + T.lx + 2;
            d[c][0] = dstencil_c4<double>(row[0], row[1], row[3], row[4], G.ix);
            d[c][1] = dstencil_c4<double>(clm[0], clm[HX], clm[3 * HX], clm[4 * HX], G.iy);
            d[c][2] = dstencil_c4<double>(q[c][S % JQ], q[c][(S + 1) % JQ], q[c][(S + 3) % JQ],
                                          q[c][(S + 4) % JQ], G.iz);
        }
    } else if (__builtin_amdgcn_ballot_w64(kind[1] > K_C2 || kind[2] > K_C2) == 0) {
        // along y and z every lane's stencil is central (a patch at the x
        // faces, away from the others): the general form along x only
        const StencilSel sx = stencil_sel(kind[0]);
        const bool cy = kind[1] == K_C4, cz = kind[2] == K_C4;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *row = P + c * WPLANE + (T.ly + 2) * HX + T.lx;
            const float *clm = P + c * WPLANE + T.ly * HX + T.lx + 2;
            d[c][0] = dstencil_sel(sx, row[0], row[1], row[2], row[3], row[4], G.ix);
            d[c][1] = dstencil_c4c2(cy, clm[0], clm[HX], clm[3 * HX], clm[4 * HX], G.iy);
            d[c][2] = dstencil_c4c2(cz, q[c][S % JQ], q[c][(S + 1) % JQ], q[c][(S + 3) % JQ], q[c][(S + 4) % JQ],
                                    G.iz);
        }
    } else {
        const StencilSel sx = stencil_sel(kind[0]), sy = stencil_sel(kind[1]), sz = stencil_sel(kind[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *row = P + c * WPLANE + (T.ly + 2) * HX + T.lx;
            const float *clm = P + c * WPLANE + T.ly * HX + T.lx + 2;
            d[c][0] = dstencil_sel(sx, row[0], row[1], row[2], row[3], row[4], G.ix);
            d[c][1] = dstencil_sel(sy, clm[0], clm[HX], clm[2 * HX], clm[3 * HX], clm[4 * HX], G.iy);
            d[c][2] = dstencil_sel(sz, q[c][S % JQ], q[c][(S + 1) % JQ], q[c][(S + 2) % JQ], q[c][(S + 3) % JQ],
                                   q[c][(S + 4) % JQ], G.iz);
        }
    }
    const float j = (float)det_plus_identity<double>(d);
    if (z < z1 && T.x < G.nx && T.y < G.ny) bstore(rj, j, T.voff[0], (unsigned)z * G.plane_bytes);
}

}  // namespace

__global__ void __launch_bounds__(64) k_jacobian_stream(JacGeom G)
{
    __shared__ float P[3 * WPLANE + 1];
    const int t = jac_tile(G, blockIdx.x);
    if (t < 0) return;
    const int tc = t % G.cols;
    const int x0 = tc % G.ntx * PX, y0 = tc / G.ntx * PY;
    const int z0 = t / G.cols * G.tz, z1 = min(z0 + G.tz, G.nz);

    JacLane T;
    const int l = threadIdx.x;
    T.lx = l % PX;
    T.ly = l / PX;
    T.x = x0 + T.lx;
    T.y = y0 + T.ly;
    T.off = (unsigned)(min(T.y, G.ny - 1) * (long long)G.nx + min(T.x, G.nx - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) T.voff[c] = 4u * (unsigned)(c * G.n + T.off);
#pragma unroll
    for (int k = 0; k < HPL; ++k) {
        const int v = l + 64 * k;
        if (v < HVALS) {
            halo_cell(G, v, x0, y0, T.hoff[k], T.hcell[k]);
        } else {
            T.hoff[k] = 0; T.hcell[k] = WDUMMY;
        }
    }
    float q[3][JQ];
    float h[HR][HPL];
    const __amdgpu_buffer_rsrc_t rf = make_rsrc(G.f, 3u * (unsigned)G.n * 4u);
    const __amdgpu_buffer_rsrc_t rj = make_rsrc(G.J, (unsigned)G.n * 4u);
#pragma unroll
    for (int k = 0; k < JQ - 1; ++k) {
        const unsigned so = plane_off(G, z0 - 2 + k);
#pragma unroll
        for (int c = 0; c < 3; ++c) q[c][k] = bload(rf, T.voff[c], so);
    }
#pragma unroll
    for (int j = 0; j < HR - 1; ++j) {
        const unsigned so = plane_off(G, z0 + j);
#pragma unroll
        for (int k = 0; k < HPL; ++k) h[j][k] = bload(rf, T.hoff[k], so);
    }

    // whole groups of eight steps, every step taken: a step past z1 (only in
    // the last chunk of a volume whose depth is no multiple of 8) reads
    // clamped planes and stores nothing.  Steps under conditions would leave
    // the compiler unsure which loads are in flight, and it would then wait
    // for loads that are steps ahead.
    static_assert(JQ == 8, "the loop below unrolls eight steps");
    for (int z = z0; z < z1; z += JQ) {
        jac_step<0>(G, T, P, q, h, z, z1, rf, rj);
        jac_step<1>(G, T, P, q, h, z + 1, z1, rf, rj);
        jac_step<2>(G, T, P, q, h, z + 2, z1, rf, rj);
        jac_step<3>(G, T, P, q, h, z + 3, z1, rf, rj);
        jac_step<4>(G, T, P, q, h, z + 4, z1, rf, rj);
        jac_step<5>(G, T, P, q, h, z + 5, z1, rf, rj);
        jac_step<6>(G, T, P, q, h, z + 6, z1, rf, rj);
        jac_step<7>(G, T, P, q, h, z + 7, z1, rf, rj);
    }
}

static int jacobian_blocks_per_cu()
{
    static int v = 0;
    if (!v) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_jacobian_stream, 64, 0) != hipSuccess || b < 1) b = 1;
        v = b;
    }
    return v;
}

// z-chunk depth (a multiple of the eight unrolled steps): minimises rounds
// of resident blocks x (depth + the planes of queue fill) -- deep chunks
// re-read fewer halo planes, shallow ones fill the CUs of a small volume.
static int jacobian_tz(int cols, int nz, int cus, int resident)
{
    int best = nz;
    double best_cost = 1e300;
    for (int tz = JQ; tz < nz + JQ; tz += JQ) {
        const long long blocks = (long long)cols * ((nz + tz - 1) / tz);
        const long long rounds = (blocks + (long long)cus * resident - 1) / ((long long)cus * resident);
        const double cost = (double)rounds * (tz + 4 + LEAD);
        if (cost < best_cost - 1e-9) { best_cost = cost; best = tz; }
    }
    return best;
}

hipError_t launch_jacobian(const float *f, float *J, int nx, int ny, int nz, float ix, float iy, float iz,
                           hipStream_t st)
{
    JacGeom G;
    G.f = f; G.J = J;
    G.nx = nx; G.ny = ny; G.nz = nz;
    G.nxy = (long long)nx * ny;
    G.n = G.nxy * nz;
    G.plane_bytes = (unsigned)(4 * G.nxy);
    G.ix = ix; G.iy = iy; G.iz = iz;
    // byte offsets (component included) are 32-bit: at most 357913941 voxels
    if (12 * G.n > 0xffffffffll) return hipErrorInvalidValue;
    G.ntx = (nx + PX - 1) / PX;
    G.cols = G.ntx * ((ny + PY - 1) / PY);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    G.tz = jacobian_tz(G.cols, nz, cus, jacobian_blocks_per_cu());
    const long long ntiles = (long long)G.cols * ((nz + G.tz - 1) / G.tz);
    if (ntiles > (1ll << 30)) return hipErrorInvalidValue;
    G.ntiles = (int)ntiles;
    G.per = (G.ntiles + 7) / 8;
    const unsigned blocks = (unsigned)G.per * 8;
    hipLaunchKernelGGL(k_jacobian_stream, dim3(blocks), dim3(64), 0, st, G);
    return hipGetLastError();
}

}  // namespace dt
