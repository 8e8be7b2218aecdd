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

// Per-solve control block, resident in device memory in two copies (one per
// pass parity).  At the start of a pass, one thread of every block reads the
// copy of the pass's parity and the sums of the previous pass and advances
// the controller (every block reaches the same state: the inputs are the
// same); the instance's first tile writes the new state into the other copy,
// which the next pass reads.  The pass kernels then use the first group of
// fields.
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
    int acc_cur;     // accumulator slot this pass adds its sums into
    int acc_prev;    // slot holding the previous pass's sums
    int acc_zero;    // slot the instance's first tile clears during this pass
    int evaluated;   // a trial has been evaluated since the state was last advanced
    int copy_now;    // this pass copies the result into buf[0] (the pass that ended the solve)
    int result_copied;   // the result has been copied into buf[0] by a pass
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

// Exact sums of one pass.  A sum of squares is kept as an unsigned 128-bit
// fixed-point number with quantum 2^-64, split into four 32-bit limbs, each
// limb added into a 64-bit word: integer addition in any order gives the same
// total, so the sums do not depend on how a volume is tiled.  Blocks add
// into the copy of their XCD (block index mod ACC_COPIES) to spread the
// atomics; a slot is ACC_COPIES copies, an instance has ACC_SLOTS slots
// (the pass adding, the previous pass's sums, the one being cleared).
constexpr int ACC_COPIES = 8;
constexpr int ACC_SLOTS = 3;
struct Acc {
    unsigned long long e[4];   // sum of e^2 over all voxels, limbs
    unsigned long long g[4];   // sum of |g|^2, limbs
    unsigned int mx;           // max |e| over the mask (bits of a non-negative float)
    unsigned int flags;        // ACC_E_* / ACC_G_* below
};
constexpr unsigned ACC_E_HUGE = 1u;   // an e^2 term at or above 2^32, or infinite
constexpr unsigned ACC_E_NAN = 2u;
constexpr unsigned ACC_G_HUGE = 4u;
constexpr unsigned ACC_G_NAN = 8u;

// One stencil along one axis as  d = A (p1 - m1) + B (p2 - m2) + G c  (samples
// at -2, -1, 0, +1, +2; 1/spacing folded into the coefficients).  One row is
// one 16-byte load.
struct alignas(16) CoefRow {
    float A, B, G, unused;
};

// One solve instance (a volume with its parameters and buffers).
struct Inst {
    int nx, ny, nz;
    int ntx, nty, ntz, tz;   // tiles per axis and z-chunk depth
    int variant;             // column-tile shape of the gradient pass
    long long n;             // voxels
    float ix, iy, iz;        // inverse spacing, (float)(1.0 / d)
    float dx, dy, dz;        // spacing (the greedy step uses it)
    // stencil coefficients [axis][kind: 4th-order central, 2nd-order central,
    // forward, backward], float products of the inverse spacing i formed on
    // the host: (2/3 i, -(1/12 i), 0), (i/2, 0, 0), (2 i, -(i/2), -(3/2 i)),
    // (2 i, -(i/2), 3/2 i)
    CoefRow coef[3][4];
    dt_params p;
    const float *target;
    const uint8_t *mask;
    float *buf[NBUF];
    float *err[2];           // greedy error maps
    Acc *acc;                // [ACC_SLOTS][ACC_COPIES] exact sums of the passes
    Ctrl *ctrl;              // [2]: control block by pass parity
    double *trace;           // [trace_cap][4]
    long long trace_cap;
    unsigned long long *status;   // host-coherent completion word (null: none)
    Ctrl *final_ctrl;             // host-coherent copy of the control block once done (null: none)
};

// One entry of a launch's tile table: a column tile of an instance and the
// planes [z0, z1) the block covers (z0 a multiple of 4).  The tiling of an
// instance is only in its table entries, so a table may re-tile an instance
// from one pass to the next.
struct Tile {
    int inst;   // instance index
    int tile;   // column tile tx + ntx * ty (matching: block index)
    int z0, z1;
};

}  // namespace dt
