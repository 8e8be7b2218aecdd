// dt_internal.hpp -- structures shared by the HIP kernels and the host runtime.
//
// Data layout in HBM (the reference's layout, src/headers/disptools.h:30-41,174-183):
//   field   float [3][nz][ny][nx]   component-major, x fastest
//   target  float [nz][ny][nx]
//   mask    uint8 [nz][ny][nx]      (C bool)
// A solve keeps four field buffers resident (accepted iterate, speculative
// candidate, two scratch) plus target and mask: 53 bytes per voxel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "disptools_hip.h"

namespace dt {

constexpr int NBUF = 4;

// Trial forms of a gradient pass.
enum TrialMode : int {
    TRIAL_DIRECT = 0,   // trial = buf[src]
    TRIAL_INTERP = 1,   // trial = buf[src] + s * (buf[cand] - buf[src])
};

// Controller phases.
enum Phase : int {
    PH_INIT = 0,     // evaluating the initial guess
    PH_SEARCH = 1,   // initial step-length search (gradient only)
    PH_MAIN = 2,     // main loop
    PH_DONE = 3,
};

// Per-solve control block, resident in device memory.  The pass kernels read
// the first group of fields; the controller (one thread of the last block of
// each pass) updates everything.
struct Ctrl {
    // ---- read by the pass kernels ----
    int active;      // 0 once the solve has terminated (passes exit at once)
    int mode;        // TrialMode
    int src;         // DIRECT: trial buffer; INTERP: base buffer
    int cand;        // INTERP: candidate buffer
    int out_trial;   // INTERP: buffer receiving the trial (-1: not stored)
    int out_next;    // buffer receiving the speculative next candidate (-1: none)
    float s;         // INTERP factor
    float eta_next;  // step of the speculative candidate
    int step_on;     // greedy: 1 when the step kernel forms a new trial
    int g_new;       // greedy: buffer receiving the trial
    int e_cur;       // greedy: error buffer the step reads
    int e_new;       // greedy: error buffer the evaluation writes
    float eta;       // step length of the trial being evaluated
    // ---- controller state ----
    int phase;
    int base;        // accepted iterate
    int cand_buf;    // candidate formed from base at eta_c
    float eta_c;
    int reason;      // dt_reason
    int result_buf;  // buffer holding the result once done
    int accepted;    // at least one trial accepted
    double E_last, m_last, G_base;
    double E0, m0;
    double E_result, m_result;
    long long it;
    long long evaluations;
    long long rejections;
    long long trace_len;
    long long field_io;      // 3-component field volumes read + written by evaluated passes
};

static_assert(sizeof(Ctrl) % 8 == 0, "the control block is published in 8-byte words");

// One solve instance (a volume with its parameters and buffers).
struct Inst {
    int nx, ny, nz;
    int ntx, nty, ntz, tz;   // tiles per axis and z-chunk depth
    int variant;             // column-tile shape of the gradient pass
    long long n;             // voxels
    float ix, iy, iz;        // inverse spacing, (float)(1.0 / d)
    float dx, dy, dz;        // spacing (the greedy step uses it)
    dt_params p;
    const float *target;
    const uint8_t *mask;
    float *buf[NBUF];
    float *err[2];           // greedy error maps
    float *wmap;             // penalised weight per voxel (two-kernel gradient pass)
    double *partials;        // [tiles][4]: sum e^2, max |e| (masked), sum |g|^2, spare
    unsigned int *counter;   // blocks of this instance that finished the pass
    Ctrl *ctrl;
    double *trace;           // [trace_cap][4]
    long long trace_cap;
    unsigned long long *status;   // host-coherent completion word (null: none)
    Ctrl *final_ctrl;             // host-coherent copy of the control block once done (null: none)
};

// One entry of a launch's tile table.
struct Tile {
    int inst;   // instance index
    int tile;   // tile index inside the instance
};

}  // namespace dt
