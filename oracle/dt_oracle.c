/*
 * dt_oracle.c -- CPU restatement of disptools' displacement-generation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  Nothing in the product library links or calls it.
 *
 * It restates, in plain C99, the arithmetic of the reference C library
 * (m-pilia/disptools v0.4.1) so that, built with the same numerics as the
 * reference release build (-std=c99 -O3, no fast-math, FLOATING = float or
 * double), it reproduces the reference bit for bit at one OpenMP thread:
 *
 *   jacobian          src/jacobian.c:70-242, src/headers/jacobian_macros.h:4-195,
 *                     det3/det3j src/headers/disptools.h:186-199
 *   compute_error     src/headers/error.h:87-172
 *   gradient          src/displacement_field_gradient.c:13-145
 *   move_field        src/displacement_field_gradient.c:150-175
 *   gradient solver   src/displacement_field_gradient.c:191-431
 *   greedy_step       src/displacement_field_greedy.c:23-67
 *   greedy solver     src/displacement_field_greedy.c:85-247
 *
 * The C promotion rules are part of the specification: differences of two
 * FLOATING samples are FLOATING, the stencil weights are double literals, so
 * most stencils and the determinant are evaluated in double and rounded to
 * FLOATING when stored; the energy sums are FLOATING accumulators.
 *
 * Pinning (what checks each function, so that it is not "parity unpinned"):
 *   - every function is compared bit-for-bit with the reference C compiled
 *     from its own sources into oracle/_ref (tests/test_oracle_ref.py);
 *   - the one-sided stencils against the reference's gtest known answers
 *     (test/c/test_jacobian_macros.cpp:70-148, restated in tests);
 *   - jacobian against closed forms (u = 0, affine u, quadratic u);
 *   - gradient against a finite-difference derivative of the second-order
 *     energy in double precision;
 *   - solvers against committed golden traces from the compiled reference.
 *
 * Build: see oracle/Makefile (FP32 and FP64 shared libraries).
 */

#include <float.h>
#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_FLOAT_SIZE
#define ORACLE_FLOAT_SIZE 32
#endif

#if ORACLE_FLOAT_SIZE == 32
typedef float real;
#define REAL_HUGE HUGE_VALF
#elif ORACLE_FLOAT_SIZE == 64
typedef double real;
#define REAL_HUGE DBL_MAX
#else
#error "ORACLE_FLOAT_SIZE must be 32 or 64"
#endif

/* Termination reasons (shared numbering with the HIP library). */
enum {
    ORACLE_RUNNING = 0,
    ORACLE_STOP_ERROR_EXPLODED = 1,   /* !isnormal(total error)            */
    ORACLE_STOP_ETA_SMALL = 2,        /* eta < iota                        */
    ORACLE_STOP_NO_IMPROVEMENT = 3,   /* 1 - E/E_last < theta              */
    ORACLE_STOP_VOXEL_EXPLODED = 4,   /* !isnormal(max voxel error)        */
    ORACLE_STOP_TOLERANCE = 5,        /* max voxel error < epsilon         */
    ORACLE_STOP_MAX_ITERATIONS = 6    /* it_max iterations accepted        */
};

/* Per-solve record; trace holds 4 doubles (it, E, max, eta) per log line. */
typedef struct oracle_result {
    long iterations;   /* value of the iteration counter at exit            */
    int reason;        /* one of the ORACLE_STOP_* codes                    */
    long evaluations;  /* number of jacobian + error evaluations            */
    long rejections;   /* line-search / greedy rejections                   */
    double final_error;      /* E of the returned field                     */
    double final_max_error;  /* masked max |e| of the returned field        */
    double *trace;     /* optional caller buffer, may be NULL               */
    long trace_cap;    /* capacity in log lines                             */
    long trace_len;    /* log lines produced (may exceed trace_cap)         */
} oracle_result;

int oracle_float_size(void) { return ORACLE_FLOAT_SIZE; }

/* ------------------------------------------------------------------------ */
/* Geometry                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct grid {
    ptrdiff_t nx, ny, nz;   /* sizes                                        */
    ptrdiff_t sy, sz;       /* strides of y and z in a scalar volume        */
    ptrdiff_t n;            /* voxels in a scalar volume (component stride) */
} grid;

static grid make_grid(size_t nx, size_t ny, size_t nz)
{
    grid g;
    g.nx = (ptrdiff_t) nx;
    g.ny = (ptrdiff_t) ny;
    g.nz = (ptrdiff_t) nz;
    g.sy = g.nx;
    g.sz = g.nx * g.ny;
    g.n = g.nx * g.ny * g.nz;
    return g;
}

static inline ptrdiff_t at(const grid *g, ptrdiff_t x, ptrdiff_t y, ptrdiff_t z)
{
    return (z * g->ny + y) * g->nx + x;
}

/* ------------------------------------------------------------------------ */
/* Finite-difference stencils (jacobian_macros.h:4-148)                     */
/* `c` points at the sample of one component, `s` is the axis stride.      */
/* ------------------------------------------------------------------------ */

#define ORACLE_TWO_THIRDS (2.0 / 3.0)
#define ORACLE_ONE_TWELFTH (1.0 / 12.0)

/* Fourth-order central: sample differences in FLOATING, weights in double. */
static inline double stencil_c4(const real *c, ptrdiff_t s, real inv)
{
    return ((c[s] - c[-s]) * ORACLE_TWO_THIRDS
            - (c[2 * s] - c[-2 * s]) * ORACLE_ONE_TWELFTH) * inv;
}

/* Second-order central: (a - b) * inv in FLOATING, then halved in double. */
static inline double stencil_c2(const real *c, ptrdiff_t s, real inv)
{
    return (c[s] - c[-s]) * inv * .5;
}

/* Second-order one-sided, forward (towards +s) and backward, in double.   */
static inline double stencil_fwd(const real *c, ptrdiff_t s, real inv)
{
    return (-1.5 * c[0] + 2.0 * c[s] - 0.5 * c[2 * s]) * inv;
}

static inline double stencil_bwd(const real *c, ptrdiff_t s, real inv)
{
    return (+1.5 * c[0] - 2.0 * c[-s] + 0.5 * c[-2 * s]) * inv;
}

/* det(D + I) in double with the cofactor expansion along the first row.   */
static inline double det_plus_identity(const double d[3][3])
{
    const double a11 = d[0][0] + 1.0, a12 = d[0][1], a13 = d[0][2];
    const double a21 = d[1][0], a22 = d[1][1] + 1.0, a23 = d[1][2];
    const double a31 = d[2][0], a32 = d[2][1], a33 = d[2][2] + 1.0;
    return a11 * (a22 * a33 - a23 * a32)
         - a12 * (a21 * a33 - a23 * a31)
         + a13 * (a21 * a32 - a22 * a31);
}

/* Stencil kinds along one axis. */
enum { K_C4, K_C2, K_FWD, K_BWD };

static inline double stencil(int kind, const real *c, ptrdiff_t s, real inv)
{
    switch (kind) {
    case K_C4: return stencil_c4(c, s, inv);
    case K_C2: return stencil_c2(c, s, inv);
    case K_FWD: return stencil_fwd(c, s, inv);
    default: return stencil_bwd(c, s, inv);
    }
}

/*
 * Exposed for the known-answer tests of the reference's gtest suite:
 * derivative of component `comp` along `axis` (0=x,1=y,2=z) at (x,y,z).
 * kind: 0 = 4th-order central, 1 = 2nd-order central, 2 = forward, 3 = backward.
 */
double oracle_stencil(int kind, const real *f, size_t nx, size_t ny, size_t nz,
                      size_t x, size_t y, size_t z, int comp, int axis, real inv)
{
    const grid g = make_grid(nx, ny, nz);
    const ptrdiff_t stride[3] = {1, g.sy, g.sz};
    const real *c = f + comp * g.n + at(&g, (ptrdiff_t) x, (ptrdiff_t) y, (ptrdiff_t) z);
    return stencil(kind, c, stride[axis], inv);
}

/* Which stencil the reference's jacobian() leaves at a voxel (jacobian.c):
 * voxels on the outer face use one-sided differences on the axes where they
 * touch the face and 2nd-order central elsewhere (:112-221); voxels on the
 * penultimate layer use 2nd-order central on all axes (:223-241); the rest
 * use 4th-order central (:87-110).  Each voxel's value is the last write. */
static inline void classify(const grid *g, ptrdiff_t x, ptrdiff_t y, ptrdiff_t z, int kind[3])
{
    const ptrdiff_t p[3] = {x, y, z};
    const ptrdiff_t n[3] = {g->nx, g->ny, g->nz};
    bool outer = false, penultimate = false;
    for (int a = 0; a < 3; ++a) {
        if (p[a] == 0 || p[a] == n[a] - 1) outer = true;
        else if (p[a] == 1 || p[a] == n[a] - 2) penultimate = true;
    }
    for (int a = 0; a < 3; ++a) {
        if (outer) {
            if (p[a] == 0) kind[a] = K_FWD;
            else if (p[a] == n[a] - 1) kind[a] = K_BWD;
            else kind[a] = K_C2;
        } else {
            kind[a] = penultimate ? K_C2 : K_C4;
        }
    }
}

static inline real jacobian_at(const grid *g, const real *f, ptrdiff_t x, ptrdiff_t y, ptrdiff_t z,
                               const real inv[3])
{
    const ptrdiff_t stride[3] = {1, g->sy, g->sz};
    int kind[3];
    classify(g, x, y, z, kind);
    const ptrdiff_t i = at(g, x, y, z);
    double d[3][3];
    for (int comp = 0; comp < 3; ++comp) {
        const real *c = f + comp * g->n + i;
        for (int axis = 0; axis < 3; ++axis) {
            d[comp][axis] = stencil(kind[axis], c, stride[axis], inv[axis]);
        }
    }
    return (real) det_plus_identity(d);
}

static void jacobian_volume(const grid *g, const real *f, real *J, real dx, real dy, real dz)
{
    /* inverse spacing computed in double and stored as FLOATING (jacobian.c:76-78) */
    const real inv[3] = {1.0 / dx, 1.0 / dy, 1.0 / dz};
    #pragma omp parallel for collapse(2) schedule(static)
    for (ptrdiff_t z = 0; z < g->nz; ++z) {
        for (ptrdiff_t y = 0; y < g->ny; ++y) {
            for (ptrdiff_t x = 0; x < g->nx; ++x) {
                J[at(g, x, y, z)] = jacobian_at(g, f, x, y, z, inv);
            }
        }
    }
}

/* jacobian_dynamic (jacobian.c:7-21): f is [3][nz][ny][nx], J is [nz][ny][nx]. */
void oracle_jacobian(size_t nx, size_t ny, size_t nz, real dx, real dy, real dz,
                     const real *f, real *J)
{
    const grid g = make_grid(nx, ny, nz);
    jacobian_volume(&g, f, J, dx, dy, dz);
}

/* ------------------------------------------------------------------------ */
/* Error (error.h:87-172)                                                   */
/* ------------------------------------------------------------------------ */

static real error_volume(const grid *g, const real *target, const real *Jf, const bool *mask,
                         real tolerance, real *e, real *max_error_out)
{
    real total = 0.0;
    real max_error = 0.0;
    const ptrdiff_t n = g->n;
    /* Linear order over (z, y, x) equals the reference's collapse(3) order. */
    #pragma omp parallel for reduction(+: total) reduction(max: max_error) schedule(static)
    for (ptrdiff_t i = 0; i < n; ++i) {
        const real err = Jf[i] - target[i];
        real v;
        if (mask[i]) {
            v = err;
        } else {
            const real a = fabs(err);
            v = a < tolerance ? 0.0 : copysign(a - tolerance, err);
        }
        e[i] = v;
        total += v * v;
        if (mask[i] && fabs(v) > max_error) {
            max_error = fabs(v);
        }
    }
    *max_error_out = max_error;
    return total;
}

real oracle_compute_error(size_t nx, size_t ny, size_t nz, const real *target, const real *Jf,
                          const bool *mask, real tolerance, real *e, real *max_error)
{
    const grid g = make_grid(nx, ny, nz);
    return error_volume(&g, target, Jf, mask, tolerance, e, max_error);
}

/* ------------------------------------------------------------------------ */
/* Gradient of the energy (displacement_field_gradient.c:13-145)            */
/* ------------------------------------------------------------------------ */

/* 2nd-order central derivative of component c along an axis, with the
 * identity added on the diagonal (gradient.c:13-21): (a-b)*inv in FLOATING,
 * then "* .5" (and "+ 1.0") in double. */
static inline double dg(const real *c, ptrdiff_t s, real inv)
{
    return (c[s] - c[-s]) * inv * .5;
}

static inline double dg1(const real *c, ptrdiff_t s, real inv)
{
    return (c[s] - c[-s]) * inv * .5 + 1.0;
}

static real gradient_volume(const grid *g, const real *f, real *gr, const real *Jf, const real *e,
                            real delta, real zeta, real dx, real dy, real dz)
{
    const real ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
    const ptrdiff_t n = g->n, sx = 1, sy = g->sy, sz = g->sz;
    const real *ux = f, *uy = f + n, *uz = f + 2 * n;
    real squared_norm = 0.0;

    #pragma omp parallel for collapse(3) reduction(+: squared_norm) schedule(static)
    for (ptrdiff_t z = 1; z < g->nz - 1; ++z) {
        for (ptrdiff_t y = 1; y < g->ny - 1; ++y) {
            for (ptrdiff_t x = 1; x < g->nx - 1; ++x) {
                const ptrdiff_t i = at(g, x, y, z);
                const ptrdiff_t q[6] = {i - sx, i + sx, i - sy, i + sy, i - sz, i + sz};
                real w[6];
                for (int k = 0; k < 6; ++k) {
                    /* voxel error plus the penalty for J below delta (:71-88) */
                    w[k] = e[q[k]];
                    if (Jf[q[k]] < delta) {
                        w[k] += zeta * (Jf[q[k]] - delta);
                    }
                }
                const ptrdiff_t xb = q[0], xf = q[1], yb = q[2], yf = q[3], zb = q[4], zf = q[5];

                /* minors of D at the neighbours; D_ab = d u_a / d b */
                const double gx =
                    (+ix * (+w[0] * (dg1(uy + xb, sy, iy) * dg1(uz + xb, sz, iz) -
                                     dg(uy + xb, sz, iz) * dg(uz + xb, sy, iy))
                            -w[1] * (dg1(uy + xf, sy, iy) * dg1(uz + xf, sz, iz) -
                                     dg(uy + xf, sz, iz) * dg(uz + xf, sy, iy)))
                     -iy * (+w[2] * (dg(uy + yb, sx, ix) * dg1(uz + yb, sz, iz) -
                                     dg(uy + yb, sz, iz) * dg(uz + yb, sx, ix))
                            -w[3] * (dg(uy + yf, sx, ix) * dg1(uz + yf, sz, iz) -
                                     dg(uy + yf, sz, iz) * dg(uz + yf, sx, ix)))
                     +iz * (+w[4] * (dg(uy + zb, sx, ix) * dg(uz + zb, sy, iy) -
                                     dg1(uy + zb, sy, iy) * dg(uz + zb, sx, ix))
                            -w[5] * (dg(uy + zf, sx, ix) * dg(uz + zf, sy, iy) -
                                     dg1(uy + zf, sy, iy) * dg(uz + zf, sx, ix))));

                const double gy =
                    (-ix * (+w[0] * (dg(ux + xb, sy, iy) * dg1(uz + xb, sz, iz) -
                                     dg(ux + xb, sz, iz) * dg(uz + xb, sy, iy))
                            -w[1] * (dg(ux + xf, sy, iy) * dg1(uz + xf, sz, iz) -
                                     dg(ux + xf, sz, iz) * dg(uz + xf, sy, iy)))
                     +iy * (+w[2] * (dg1(ux + yb, sx, ix) * dg1(uz + yb, sz, iz) -
                                     dg(ux + yb, sz, iz) * dg(uz + yb, sx, ix))
                            -w[3] * (dg1(ux + yf, sx, ix) * dg1(uz + yf, sz, iz) -
                                     dg(ux + yf, sz, iz) * dg(uz + yf, sx, ix)))
                     -iz * (+w[4] * (dg1(ux + zb, sx, ix) * dg(uz + zb, sy, iy) -
                                     dg(ux + zb, sy, iy) * dg(uz + zb, sx, ix))
                            -w[5] * (dg1(ux + zf, sx, ix) * dg(uz + zf, sy, iy) -
                                     dg(ux + zf, sy, iy) * dg(uz + zf, sx, ix))));

                const double gz =
                    (+ix * (+w[0] * (dg(ux + xb, sy, iy) * dg(uy + xb, sz, iz) -
                                     dg(ux + xb, sz, iz) * dg1(uy + xb, sy, iy))
                            -w[1] * (dg(ux + xf, sy, iy) * dg(uy + xf, sz, iz) -
                                     dg(ux + xf, sz, iz) * dg1(uy + xf, sy, iy)))
                     -iy * (+w[2] * (dg1(ux + yb, sx, ix) * dg(uy + yb, sz, iz) -
                                     dg(ux + yb, sz, iz) * dg(uy + yb, sx, ix))
                            -w[3] * (dg1(ux + yf, sx, ix) * dg(uy + yf, sz, iz) -
                                     dg(ux + yf, sz, iz) * dg(uy + yf, sx, ix)))
                     +iz * (+w[4] * (dg1(ux + zb, sx, ix) * dg1(uy + zb, sy, iy) -
                                     dg(ux + zb, sy, iy) * dg(uy + zb, sx, ix))
                            -w[5] * (dg1(ux + zf, sx, ix) * dg1(uy + zf, sy, iy) -
                                     dg(ux + zf, sy, iy) * dg(uy + zf, sx, ix))));

                gr[i] = (real) gx;
                gr[n + i] = (real) gy;
                gr[2 * n + i] = (real) gz;
                squared_norm += gr[i] * gr[i] + gr[n + i] * gr[n + i] + gr[2 * n + i] * gr[2 * n + i];
            }
        }
    }
    return squared_norm;
}

/* g is [3][nz][ny][nx]; its outer face is left untouched (zero on entry in the reference). */
real oracle_gradient(size_t nx, size_t ny, size_t nz, real dx, real dy, real dz,
                     const real *f, const real *Jf, const real *e, real delta, real zeta, real *g)
{
    const grid gr = make_grid(nx, ny, nz);
    return gradient_volume(&gr, f, g, Jf, e, delta, zeta, dx, dy, dz);
}

/* new = old - eta * g over every voxel (gradient.c:150-175) */
static void move_volume(ptrdiff_t n3, const real *old_f, real *new_f, const real *g, real eta)
{
    #pragma omp parallel for schedule(static)
    for (ptrdiff_t i = 0; i < n3; ++i) {
        new_f[i] = old_f[i] - eta * g[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Solvers                                                                  */
/* ------------------------------------------------------------------------ */

static void log_line(oracle_result *r, double it, double E, double m, double eta)
{
    if (r->trace && r->trace_len < r->trace_cap) {
        double *t = r->trace + 4 * r->trace_len;
        t[0] = it; t[1] = E; t[2] = m; t[3] = eta;
    }
    r->trace_len++;
}

static void reset_result(oracle_result *r)
{
    r->iterations = 0;
    r->reason = ORACLE_RUNNING;
    r->evaluations = 0;
    r->rejections = 0;
    r->final_error = 0.0;
    r->final_max_error = 0.0;
    r->trace_len = 0;
}

/*
 * generate_displacement_gradient (gradient.c:191-431).  Returns 0, or -1 if
 * scratch memory could not be allocated.  `field` is the initial guess on
 * entry and the result on exit, [3][nz][ny][nx].
 */
int oracle_solve_gradient(size_t nx, size_t ny, size_t nz, real dx, real dy, real dz,
                          const real *target, const bool *mask,
                          real epsilon, real tolerance, real eta, real eta_max,
                          real alpha, real beta, real gamma, real delta, real zeta,
                          real theta, real iota, bool strict, size_t it_max,
                          real *field, oracle_result *res)
{
    const grid g = make_grid(nx, ny, nz);
    const ptrdiff_t n = g.n;
    oracle_result dummy;
    oracle_result *r = res ? res : &dummy;
    if (!res) { dummy.trace = NULL; dummy.trace_cap = 0; }
    reset_result(r);

    real *buf = (real *) calloc((size_t) (3 * n * 2 + 2 * n + n + 3 * n), sizeof(real));
    if (!buf) return -1;
    real *f_[2] = {buf, buf + 3 * n};
    real *J_[2] = {buf + 6 * n, buf + 7 * n};
    real *e = buf + 8 * n;
    real *gr = buf + 9 * n;
    unsigned o = 0, nw = 1;

    real last_error, error, max_error, last_max_error = REAL_HUGE, g2 = 0.0;
    bool accepted = false;

    memcpy(f_[o], field, (size_t) (3 * n) * sizeof(real));
    jacobian_volume(&g, f_[o], J_[o], dx, dy, dz);
    last_error = error_volume(&g, target, J_[o], mask, tolerance, e, &max_error);
    r->evaluations++;
    log_line(r, 0.0, last_error, max_error, eta);
    const real initial_max = max_error;

    g2 = gradient_volume(&g, f_[o], gr, J_[o], e, delta, zeta, dx, dy, dz);

    /* search for a large initial step (:290-317) */
    do {
        eta *= alpha;
        move_volume(3 * n, f_[o], f_[nw], gr, eta);
        jacobian_volume(&g, f_[nw], J_[nw], dx, dy, dz);
        error = error_volume(&g, target, J_[nw], mask, tolerance, e, &max_error);
        r->evaluations++;
    } while (eta < eta_max && error - last_error > -gamma * eta * g2);
    eta /= alpha * alpha;

    last_error = error_volume(&g, target, J_[o], mask, tolerance, e, &max_error);

    size_t it;
    r->reason = ORACLE_STOP_MAX_ITERATIONS;
    for (it = 1; it <= it_max; ++it) {
        g2 = gradient_volume(&g, f_[o], gr, J_[o], e, delta, zeta, dx, dy, dz);
        eta *= alpha;
        eta = eta > eta_max ? eta_max : eta;
        for (;;) {
            move_volume(3 * n, f_[o], f_[nw], gr, eta);
            jacobian_volume(&g, f_[nw], J_[nw], dx, dy, dz);
            error = error_volume(&g, target, J_[nw], mask, tolerance, e, &max_error);
            r->evaluations++;
            const bool eta_good = eta >= iota;
            const bool armijo_fails = error - last_error > -gamma * eta * g2;
            const bool strict_fails = strict && max_error > last_max_error;
            if (eta_good && (armijo_fails || strict_fails)) {
                eta *= beta;
                r->rejections++;
            } else {
                break;
            }
        }
        log_line(r, (double) it, error, max_error, eta);

        if (!isnormal(error)) { r->reason = ORACLE_STOP_ERROR_EXPLODED; break; }
        if (eta < iota) { r->reason = ORACLE_STOP_ETA_SMALL; break; }
        if (1.0 - error / last_error < theta) { r->reason = ORACLE_STOP_NO_IMPROVEMENT; break; }
        if (!isnormal(max_error)) { r->reason = ORACLE_STOP_VOXEL_EXPLODED; break; }
        if (max_error < epsilon) { r->reason = ORACLE_STOP_TOLERANCE; break; }

        last_error = error;
        last_max_error = max_error;
        accepted = true;
        o ^= 1u; nw ^= 1u;
    }
    r->iterations = (long) it;
    /* the returned field is the last accepted iterate (or the initial guess) */
    r->final_error = last_error;
    r->final_max_error = accepted ? last_max_error : initial_max;

    memcpy(field, f_[o], (size_t) (3 * n) * sizeof(real));
    free(buf);
    return 0;
}

/* Greedy update of the interior (greedy.c:23-67): d_c is the spacing, not its inverse. */
static void greedy_volume(const grid *g, const real *old_f, real *new_f, const real *e, real eta,
                          real dx, real dy, real dz)
{
    const real ddx = eta * dx * .5f;
    const real ddy = eta * dy * .5f;
    const real ddz = eta * dz * .5f;
    const ptrdiff_t n = g->n, sy = g->sy, sz = g->sz;
    #pragma omp parallel for collapse(2) schedule(static)
    for (ptrdiff_t z = 1; z < g->nz - 1; ++z) {
        for (ptrdiff_t y = 1; y < g->ny - 1; ++y) {
            for (ptrdiff_t x = 1; x < g->nx - 1; ++x) {
                const ptrdiff_t i = at(g, x, y, z);
                const real step_x = ddx * (e[i + 1] - e[i - 1]);
                const real step_y = ddy * (e[i + sy] - e[i - sy]);
                const real step_z = ddz * (e[i + sz] - e[i - sz]);
                new_f[i] = old_f[i] + step_x;
                new_f[n + i] = old_f[n + i] + step_y;
                new_f[2 * n + i] = old_f[2 * n + i] + step_z;
            }
        }
    }
}

/* generate_displacement_greedy (greedy.c:85-247). */
int oracle_solve_greedy(size_t nx, size_t ny, size_t nz, real dx, real dy, real dz,
                        const real *target, const bool *mask,
                        real epsilon, real tolerance, real eta, real eta_max,
                        real alpha, real beta, real gamma, real delta, real zeta,
                        real theta, real iota, bool strict, size_t it_max,
                        real *field, oracle_result *res)
{
    (void) eta_max; (void) alpha; (void) gamma; (void) delta; (void) zeta; (void) strict;
    const grid g = make_grid(nx, ny, nz);
    const ptrdiff_t n = g.n;
    oracle_result dummy;
    oracle_result *r = res ? res : &dummy;
    if (!res) { dummy.trace = NULL; dummy.trace_cap = 0; }
    reset_result(r);

    real *buf = (real *) calloc((size_t) (3 * n * 2 + 2 * n + n), sizeof(real));
    if (!buf) return -1;
    real *f_[2] = {buf, buf + 3 * n};
    real *J_[2] = {buf + 6 * n, buf + 7 * n};
    real *e = buf + 8 * n;
    unsigned o = 0, nw = 1;

    real last_error, error, max_error = 0.0, last_max_error;

    memcpy(f_[o], field, (size_t) (3 * n) * sizeof(real));
    jacobian_volume(&g, f_[o], J_[o], dx, dy, dz);
    last_error = error_volume(&g, target, J_[o], mask, tolerance, e, &max_error);
    last_max_error = max_error;
    r->evaluations++;

    size_t it;
    r->reason = ORACLE_STOP_MAX_ITERATIONS;
    for (it = 1; it <= it_max; ++it) {
        greedy_volume(&g, f_[o], f_[nw], e, eta, dx, dy, dz);
        jacobian_volume(&g, f_[nw], J_[nw], dx, dy, dz);
        error = error_volume(&g, target, J_[nw], mask, tolerance, e, &max_error);
        r->evaluations++;
        log_line(r, (double) it, error, max_error, eta);

        if (!isnormal(error)) { r->reason = ORACLE_STOP_ERROR_EXPLODED; break; }
        if (error >= last_error) {
            eta *= beta;
            r->rejections++;
            if (eta < iota) { r->reason = ORACLE_STOP_ETA_SMALL; break; }
            --it;
            continue;
        }
        if (1.0 - error / last_error < theta) { r->reason = ORACLE_STOP_NO_IMPROVEMENT; break; }
        if (!isnormal(max_error)) { r->reason = ORACLE_STOP_VOXEL_EXPLODED; break; }
        if (max_error < epsilon) { r->reason = ORACLE_STOP_TOLERANCE; break; }

        last_error = error;
        last_max_error = max_error;
        o ^= 1u; nw ^= 1u;
    }
    r->iterations = (long) it;
    r->final_error = last_error;
    r->final_max_error = last_max_error;

    memcpy(field, f_[o], (size_t) (3 * n) * sizeof(real));
    free(buf);
    return 0;
}

/* Thread count for the OpenMP loops of this library (and of any library that
 * shares the process's OpenMP runtime, such as the reference build). */
#ifdef _OPENMP
#include <omp.h>
#endif
void oracle_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void) n;
#endif
}
