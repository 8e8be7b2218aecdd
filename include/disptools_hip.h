/*
 * disptools_hip.h -- C ABI of the MI355X (gfx950) displacement-generation library.
 *
 * Plain C types only.  Every pointer argument of the reference-compatible
 * entry points is caller-owned HOST memory, exactly as in the reference;
 * the dt_*_device entry points take DEVICE pointers and a hipStream_t passed
 * as `void *` (NULL = the default stream).
 *
 * FLOATING is float: the library is the FLOAT_SIZE=32 build of the reference
 * ABI (reference CMakeLists.txt:106-111).
 */
#ifndef DISPTOOLS_HIP_H
#define DISPTOOLS_HIP_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Reference-compatible entry points                                        */
/* ------------------------------------------------------------------------ */

/* Error state; replaces src/headers/disptools.h:205-212 (same layout). */
typedef struct disptools_error_state {
    bool error;
    int line;
    char file[512];
    char function[512];
    char message[1024];
    char trace[4096];
} disptools_error_state;

/* Replaces disptools_get_error / disptools_has_error / disptools_clear_error
 * (src/headers/disptools.h:250,257,264; src/disptools.c:26-43). */
disptools_error_state *disptools_get_error(void);
bool disptools_has_error(void);
void disptools_clear_error(void);

/* Replaces get_float_type_size (src/headers/disptools.h:289): returns 32. */
int get_float_type_size(void);

/* Replaces set_device (cuda/headers/generate_displacement.cuh:11,
 * cuda/device.cu:4-15): false if id is not a valid HIP device. */
bool set_device(const int id);

/* Replaces generate_displacement_gradient_cuda
 * (cuda/headers/generate_displacement.cuh:39-62); the semantics are those of
 * the reference C solver generate_displacement_gradient
 * (src/displacement_field_gradient.c:191-431).  `field` is the initial guess
 * on entry and the result on exit, layout [3][nz][ny][nx]. */
void generate_displacement_gradient_cuda(
        const size_t nx, const size_t ny, const size_t nz,
        const float dx, const float dy, const float dz,
        const float *J, const bool *mask,
        const float epsilon, const float tolerance, float eta, const float eta_max,
        const float alpha, const float beta, const float gamma, const float delta,
        const float zeta, const float theta, const float iota,
        const bool strict, const size_t it_max, float *field);

/* Replaces generate_displacement_greedy_cuda
 * (cuda/headers/generate_displacement.cuh:14-37) with the semantics of
 * generate_displacement_greedy (src/displacement_field_greedy.c:85-247). */
void generate_displacement_greedy_cuda(
        const size_t nx, const size_t ny, const size_t nz,
        const float dx, const float dy, const float dz,
        const float *J, const bool *mask,
        const float epsilon, const float tolerance, float eta, const float eta_max,
        const float alpha, const float beta, const float gamma, const float delta,
        const float zeta, const float theta, const float iota,
        const bool strict, const size_t it_max, float *field);

/* Replaces the CPU-named solvers generate_displacement_gradient
 * (src/headers/displacement_field_gradient.h:7-30) and
 * generate_displacement_greedy (src/headers/displacement_field_greedy.h:7-30),
 * which the reference extension calls for gpu_id < 0
 * (python_c_extension/_disptools.c:374-384).  Same arguments, semantics and
 * log lines; they run on the calling thread's current HIP device (this
 * library has no CPU solver). */
void generate_displacement_gradient(
        const size_t nx, const size_t ny, const size_t nz,
        const float dx, const float dy, const float dz,
        const float *J, const bool *mask,
        const float epsilon, const float tolerance, float eta, const float eta_max,
        const float alpha, const float beta, const float gamma, const float delta,
        const float zeta, const float theta, const float iota,
        const bool strict, const size_t it_max, float *field);

void generate_displacement_greedy(
        const size_t nx, const size_t ny, const size_t nz,
        const float dx, const float dy, const float dz,
        const float *J, const bool *mask,
        const float epsilon, const float tolerance, float eta, const float eta_max,
        const float alpha, const float beta, const float gamma, const float delta,
        const float zeta, const float theta, const float iota,
        const bool strict, const size_t it_max, float *field);

/* Replaces jacobian_dynamic (src/headers/jacobian.h:24-33, src/jacobian.c:7-21)
 * as bound by _disptools.jacobian (python_c_extension/_disptools.c:238-274):
 * f is [3][nz][ny][nx], J is [nz][ny][nx]; computed on the current device. */
void jacobian_dynamic(const size_t nx, const size_t ny, const size_t nz,
                      const float dx, const float dy, const float dz,
                      const float *f, float *J);

/* Replaces regularise (src/headers/jacobian.h:13-19, src/jacobian.c:29-55):
 * J[i] = max-threshold at epsilon, in place, on the current device. */
void regularise(const size_t nx, const size_t ny, const size_t nz, float *J, float epsilon);

/* ------------------------------------------------------------------------ */
/* Additive entry points (no counterpart in the reference)                 */
/* ------------------------------------------------------------------------ */

enum dt_algorithm { DT_GRADIENT = 0, DT_GREEDY = 1, DT_MATCHING = 2 /* batch engine only */ };

/* Termination reasons of a solve. */
enum dt_reason {
    DT_RUNNING = 0,
    DT_STOP_ERROR_EXPLODED = 1,   /* !isnormal(total error)               */
    DT_STOP_ETA_SMALL = 2,        /* eta < iota                           */
    DT_STOP_NO_IMPROVEMENT = 3,   /* 1 - E / E_last < theta               */
    DT_STOP_VOXEL_EXPLODED = 4,   /* !isnormal(max voxel error)           */
    DT_STOP_TOLERANCE = 5,        /* max voxel error < epsilon            */
    DT_STOP_MAX_ITERATIONS = 6    /* it_max iterations accepted           */
};

/* Solver parameters, same meaning as the reference arguments. */
typedef struct dt_params {
    float epsilon, tolerance, eta, eta_max, alpha, beta, gamma, delta, zeta, theta, iota;
    int strict;
    long long it_max;
} dt_params;

/* Record of one solve. */
typedef struct dt_solve_info {
    long long iterations;    /* iteration counter at exit                   */
    long long evaluations;   /* stencil passes that evaluated a trial       */
    long long rejections;    /* rejected trials                             */
    long long launches;      /* pass kernels launched (incl. idle ones)     */
    int reason;              /* enum dt_reason                              */
    double final_error;      /* energy of the returned field (FP64 sum)     */
    double final_max_error;  /* masked max |e| of the returned field        */
    double seconds;          /* wall time of the solve on the device        */
    long long trace_len;     /* log lines produced (it, E, max, eta each)   */
    long long field_io;      /* [3][nz][ny][nx] field volumes read+written by
                                the evaluated passes (algorithmic traffic)  */
    double kernel_seconds;   /* device time of the pass launches (events
                                before the first and after the last of the
                                back-to-back launches; 0 in engine results) */
} dt_solve_info;

/* Record of a matching solve (dt_matching_device / dt_matching_host). */
typedef struct dt_matching_info {
    long long iterations;       /* accepted iterations                         */
    int reason;                 /* DT_STOP_MAX_ITERATIONS, DT_STOP_ETA_SMALL
                                   (step below 1e-9) or DT_STOP_TOLERANCE
                                   (max voxel error <= epsilon)                */
    long long cost_passes;      /* cost evaluations (trials + the initial one) */
    long long gradient_passes;  /* gradient evaluations                        */
    long long launches;         /* pass kernels launched (incl. idle ones)     */
    double final_cost;          /* cost of the returned field (FP64 sum)       */
    double final_max_error;     /* max voxel error of the last evaluated trial */
    long long trace_len;        /* log lines (it, cost, max error, step)       */
    double seconds;             /* wall time of the solve                      */
    double kernel_seconds;      /* device time of the pass launches            */
} dt_matching_info;

/* The `matching` algorithm (src/VolumeMatching3D.c) on device buffers:
 * d_target [nz][ny][nx], d_field [3][nz][ny][nx] in/out.  Uses epsilon,
 * delta, zeta and it_max (the reference ignores the other solver parameters
 * and the mask).  `trace` (host, optional) receives up to trace_cap rows of 4
 * doubles.  Runs on `stream` and synchronises it.  Returns 0 or -1. */
int dt_matching_device(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                       const float *d_target, float epsilon, float delta, float zeta,
                       long long it_max, float *d_field, dt_matching_info *info,
                       double *trace, long long trace_cap, void *stream);

/* Same with host buffers (one H2D and one D2H copy). */
int dt_matching_host(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                     const float *target, float epsilon, float delta, float zeta,
                     long long it_max, float *field, dt_matching_info *info,
                     double *trace, long long trace_cap);

/* Reference entry point of the matching algorithm
 * (src/headers/VolumeMatching3D.h:7-29, src/VolumeMatching3D.c:689-796), run on
 * the current HIP device; host buffers, field in place, errors in the error
 * state.  CPU-only in the reference. */
void volume_matching_3d(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                        const float *J, const bool *mask, float epsilon, float tolerance,
                        float eta, float eta_max, float alpha, float beta, float gamma,
                        float delta, float zeta, float theta, float iota, bool strict,
                        size_t it_max, float *field);

/* Jacobian of a device-resident field (device pointers). Returns 0 or -1. */
int dt_jacobian_device(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                       const float *d_field, float *d_J, void *stream);

/* Solve on device-resident data: d_target [nz][ny][nx], d_mask bytes,
 * d_field [3][nz][ny][nx] in/out.  `trace` (host, optional) receives up to
 * trace_cap log lines of 4 doubles (it, E, max, eta).  Returns 0 or -1
 * (error state set).  Runs on `stream` and synchronises it at the end. */
int dt_solve_device(int algorithm, size_t nx, size_t ny, size_t nz,
                    float dx, float dy, float dz,
                    const float *d_target, const unsigned char *d_mask,
                    const dt_params *params, float *d_field,
                    dt_solve_info *info, double *trace, long long trace_cap, void *stream);

/* Size of the device array that dt_solve_device_noalloc keeps its solver
 * state in (80 MiB): a gradient solve of up to ~2.2 M voxels (128^3), a greedy
 * solve of up to ~4.1 M voxels (161^3).  Each GPU the library runs on holds
 * one such array until the process ends. */
#define DT_NOALLOC_SCRATCH_BYTES (80u * 1024u * 1024u)

/* No counterpart in the reference: dt_solve_device for callers that cannot
 * allocate memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Same arguments and results as
 * dt_solve_device without the trace; calls no hipMalloc, hipFree,
 * hipHostMalloc or event creation: the solver's state lives in a static
 * device array of DT_NOALLOC_SCRATCH_BYTES (one per GPU), set up by a kernel.
 * A volume whose state does not fit is rejected (-1, error state set) before
 * any device work.  Calls on one GPU share its array and run one at a time
 * within the library; calls from outside it that overlap on one GPU race.
 * info->seconds and info->kernel_seconds are 0 (no events are created). */
int dt_solve_device_noalloc(int algorithm, size_t nx, size_t ny, size_t nz,
                            float dx, float dy, float dz,
                            const float *d_target, const unsigned char *d_mask,
                            const dt_params *params, float *d_field,
                            dt_solve_info *info, void *stream);

/* Same as dt_solve_device with host pointers (one H2D and one D2H copy). */
int dt_solve_host(int algorithm, size_t nx, size_t ny, size_t nz,
                  float dx, float dy, float dz,
                  const float *target, const bool *mask,
                  const dt_params *params, float *field,
                  dt_solve_info *info, double *trace, long long trace_cap);

/* Scale step of the image pyramid (the reference's drawing.scale_image,
 * disptools/drawing.py:23-69, there with SimpleITK): smooths a
 * [ncomp][nz][ny][nx] float volume with spacing (dx, dy, dz) by the recursive
 * Gaussian of standard deviation `sigma` in physical units along each axis
 * (Deriche's fourth-order filter, which SimpleITK's SmoothingRecursiveGaussian
 * implements; unit DC gain, the signal continued by its end samples), then
 * resamples it onto an mx x my x mz grid with the same origin and physical
 * extent (spacing d * n / m): linear interpolation, or nearest neighbour when
 * `nearest` (masks), 0 outside the input.  Device pointers; runs on
 * `stream`.  Returns 0 or -1. */
int dt_scale_volume_device(size_t nx, size_t ny, size_t nz, size_t ncomp, float dx, float dy, float dz,
                           const float *d_in, size_t mx, size_t my, size_t mz, float sigma, int nearest,
                           float *d_out, void *stream);

/* The multi-resolution solve of displacement() (disptools/displacements.py:
 * 454-483), host buffers in and out, everything between on the device:
 * `levels` grids, each half the previous one (integer division); every
 * coarser target and mask is scale_image of the finest (linear, nearest);
 * the coarsest level starts from the initial guess scaled onto its grid, or
 * zero, every finer one from the previous level's field scaled onto its own;
 * it_max[k] is the iteration limit of the k-th level from the coarsest.
 *   J      [nz][ny][nx] target (padded as displacement() pads it)
 *   mask   [nz][ny][nx] float (the reference casts the mask to float; a
 *          voxel takes part where the scaled mask is non-zero)
 *   guess  [gz][gy][gx][3] initial guess with spacing (dx, dy, dz), or NULL
 *   field  [nz][ny][nx][3] result
 * algorithm: DT_GRADIENT, DT_GREEDY or DT_MATCHING (epsilon, delta, zeta of
 * `params`).  `info` (may be NULL) receives the finest level's record.
 * Returns 0 or -1.  No counterpart in the reference's C library (its Python
 * driver composes SimpleITK and the C solver). */
int dt_displacement_levels(int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                           const float *J, const float *mask, const float *guess, size_t gx, size_t gy, size_t gz,
                           int levels, const long long *it_max, const dt_params *params, float *field,
                           dt_solve_info *info);

/* ------------------------------------------------------------------------ */
/* Autotuner of the gradient pass (no counterpart in the reference).        */
/* ------------------------------------------------------------------------ */

#define DT_TUNE_MAX 32

typedef struct dt_tune_result {
    int best_variant;                 /* fastest variant for the shape          */
    int count;                        /* variants timed                         */
    int variant[DT_TUNE_MAX];
    double us_per_pass[DT_TUNE_MAX];  /* mean time of one pass (HIP events)     */
} dt_tune_result;

/* Times `passes` passes (0: 8) of every fused gradient-pass variant (tile
 * shapes and register budgets) on a synthetic volume of this shape, inputs
 * resident in HBM, and records the fastest for the shape on the current
 * device.  With DISPTOOLS_AUTOTUNE=1 in the environment, dt_solve_* calls
 * use the recorded variant for their shape (tuning it on first use);
 * DISPTOOLS_GRAD_VARIANT overrides both.  The variants agree to the solver's
 * parity bar, not bit for bit.  Returns 0 or -1 (error state set). */
int dt_tune_gradient(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, int passes,
                     dt_tune_result *out);

/* ------------------------------------------------------------------------ */
/* Batch engine: continuous batching of solves of uneven sizes.             */
/* No counterpart in the reference (one solve per call there); for callers  */
/* that serve many volumes: every pass launch covers all admitted volumes,  */
/* each with its own control block (step length, line-search state,         */
/* stopping tests) on the device.  Admission is by device-memory footprint. */
/* One engine is driven from one host thread.                               */
/* ------------------------------------------------------------------------ */

typedef struct dt_engine dt_engine;

typedef struct dt_engine_config {
    size_t hbm_budget;   /* device bytes for instance state (0: half of the free memory) */
    int max_slots;       /* instances on the device at once (0: 256; at most 4096)       */
    int chunk;           /* passes enqueued per chunk (0: 2)                               */
    int use_graphs;      /* 1: each chunk is replayed from a captured hipGraph             */
    int variant;         /* gradient pass variant (-1: 32 x 24 column tiles)               */
    int tz;              /* z-chunk depth in planes: 0 = re-planned for the set of
                            instances whenever it changes (balanced chunks, the bound
                            that fills the CUs best), -1 = as a lone dt_solve_device
                            call, -2 = balanced chunks of at most 16 planes, > 0 =
                            fixed; results are the same for every choice            */
} dt_engine_config;

typedef struct dt_engine_result {
    long long ticket;
    int status;          /* 0: solved; -1: failed (the line search did not terminate) */
    dt_solve_info info;  /* seconds = from submission to the result being in place    */
} dt_engine_result;

typedef struct dt_engine_stats {
    long long chunks;          /* chunks enqueued                          */
    long long passes;          /* pass launches (per kernel kind)          */
    long long graph_builds;    /* hipGraph captures + instantiations       */
    long long admitted;        /* instances admitted so far                */
    long long finished;        /* results produced so far                  */
    int active;                /* instances being solved                   */
    int retiring;              /* instances whose result copy is in flight */
    size_t pool_bytes;         /* size of the device pool                  */
    size_t pool_used;          /* bytes held by instances now              */
    size_t pool_peak;          /* largest pool_used seen                   */
} dt_engine_stats;

/* Returned by dt_engine_submit when the instance does not fit now (no free
 * slot or not enough pool memory): retry after results have retired. */
#define DT_ENGINE_FULL (-2)

/* Creates an engine on the current device.  `stream` (hipStream_t as void*,
 * NULL: a stream of its own) carries the passes.  NULL on error. */
dt_engine *dt_engine_create(const dt_engine_config *cfg, void *stream);
void dt_engine_destroy(dt_engine *e);

/* Pool bytes an instance takes (host_io: target, mask and field are staged
 * in the pool too). */
size_t dt_engine_footprint(const dt_engine *e, int algorithm, size_t nx, size_t ny, size_t nz, int host_io);

/* Admits a solve (DT_GRADIENT, DT_GREEDY or DT_MATCHING; a matching solve
 * reads epsilon, delta, zeta and it_max and no mask, which may be NULL).
 * host_io = 0: target/mask/field are device buffers that
 * stay valid and untouched until the result is polled; the result is written
 * to `field`.  host_io = 1: host buffers (pinned memory lets the copies
 * overlap the passes); inputs are copied at submission and the result is
 * copied into `field` before the ticket is reported.  Returns a ticket >= 0,
 * DT_ENGINE_FULL, or -1 (error state set). */
long long dt_engine_submit(dt_engine *e, int algorithm, size_t nx, size_t ny, size_t nz,
                           float dx, float dy, float dz, const float *target, const unsigned char *mask,
                           const dt_params *params, float *field, int host_io);

/* Advances the engine: keeps two chunks of passes in flight, retires
 * finished instances.  wait = 1 blocks until a chunk has run or an instance
 * has finished.
 * Returns the number of instances not yet reported, or -1 on error. */
int dt_engine_progress(dt_engine *e, int wait);

/* Moves up to `cap` finished results to `out`; returns how many. */
int dt_engine_poll(dt_engine *e, dt_engine_result *out, int cap);

/* Progresses until every admitted instance has finished.  0 or -1. */
int dt_engine_drain(dt_engine *e);

int dt_engine_get_stats(const dt_engine *e, dt_engine_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* DISPTOOLS_HIP_H */
