// dt_match.hpp -- records of a matching solve (src/VolumeMatching3D.c:81-796)
// shared by the matching kernels and the batch engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "dt_host.hpp"

namespace dt {

enum { MPH_INIT = 0, MPH_GRAD = 1, MPH_COST = 2, MPH_DONE = 3 };

// Control block of a matching solve (step length, line search, counters).
struct MCtrl {
    int phase, active, mode;
    int xb, tb, pb;            // accepted triple, trial triple, previous trial (growth)
    float alpha, m;            // step; max voxel error of the last cost pass
    double H, Hmark;           // cost of the accepted point; line-search reference
    long long it, cost_passes, grad_passes, trace_len;
    int reason;
};
static_assert(sizeof(MCtrl) % 8 == 0, "the control block is published in 8-byte words");

struct PassGeom {
    int ntx, nty, ntz, tz;
};

// One matching instance: geometry, parameters and its device buffers.
struct MInst {
    int nx, ny, nz;
    PassGeom cg, gg;           // tilings of the cost and gradient passes
    int nblocks;               // blocks of one pass (both tilings fit, a multiple of 8)
    float delta, zeta, epsilon;
    long long it_max;
    long long n;
    float *tri[3];             // triple t: F, G, E at tri[t] + {0, n, 2n}
    float *w[3];               // cell weights of triple t, node-indexed
    float *D;                  // gradient, F, G, E at D + {0, n, 2n}
    const float *V;            // target, node-indexed
    double *partials;          // 2 per block
    unsigned *counter;
    MCtrl *ctrl;
    double *trace;
    long long trace_cap;
    unsigned long long *status;   // host-coherent completion word: (passes << 1) | done (null: none)
    MCtrl *final_ctrl;            // host-coherent copy of the control block once done (null: none)
};

// Checks the arguments of a matching solve (sets the error state, returns -1).
int matching_validate(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float epsilon, float delta,
                      float zeta);
// Device bytes of an instance's solver state (node triples, weights,
// gradient, tile partials), carved by match_make_inst in this order.
size_t match_scratch_bytes(int nx, int ny, int nz);
// Fills the instance record; buffers from `cv`, control block and counter
// given (not carved).
MInst match_make_inst(int nx, int ny, int nz, const float *d_target, float epsilon, float delta, float zeta,
                      long long it_max, Carver &cv, MCtrl *ctrl, unsigned *counter);
// Control block of a solve that starts from the positions in tri[0].
MCtrl match_initial_ctrl();
// Passes after which the host declares a matching solve stuck.
long long match_max_passes(long long it_max);
// field -> node positions tri[0] of the instance at `d_inst` (host copy `h`).
hipError_t launch_match_in(const MInst &h, const float *d_field, float dx, float dy, float dz, hipStream_t st);
// accepted node positions -> field.
hipError_t launch_match_out(const MInst &h, const MInst *d_inst, float *d_field, float dx, float dy, float dz,
                            hipStream_t st);
// One pass over the tiles of several instances (tile table as the engine's).
hipError_t launch_match_batch(const MInst *insts, const Tile *tiles, int nblocks, hipStream_t st);

}  // namespace dt
