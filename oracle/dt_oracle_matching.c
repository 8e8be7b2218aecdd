/*
 * dt_oracle_matching.c -- CPU restatement of the reference's `matching`
 * algorithm (Karacali & Davatzikos volume matching).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP matching solver; nothing in
 * the product library links or calls it.
 *
 * Follows src/VolumeMatching3D.c:
 *   cell cost and max voxel error   VolumeMatchingCost3D   :81-189
 *   gradient of the cost            VolumeMatchingGrads3D  :191-406
 *   step-length control             VolumeMatching3D       :408-558
 *   coordinates in / field out      volume_matching_3d     :689-796
 *
 * The problem: nodes (i, j, k) of the nx x ny x nz grid carry the deformed
 * position in voxel units, F = y + 1 + u_y / dy, G = x + 1 + u_x / dx,
 * E = z + 1 + u_z / dz.  Each of the (nx-1)(ny-1)(nz-1) cells has a target
 * V (the target Jacobian at its lowest node) and eight corner Jacobians:
 * at corner q, the determinant of the 3x3 matrix whose rows are F, G, E and
 * whose columns are their forward differences along the cell edges through q
 * in the order y, x, z.  The cost is
 *     sum over cells of 0.5 ( [V > 0] (S - 8 V)^2
 *                             + sum_q [J_q < delta] zeta (J_q - delta)^2 ),
 * S = sum_q J_q, and the max voxel error is max over all cells of
 * |S - 8 V| / 8.  Its gradient with respect to a node value collects, from
 * each cell around the node, W_q times the cofactor of the entry of corner
 * q's matrix that the node enters (with the sign of the difference), where
 * W_q = [V > 0] (S - 8 V) + [J_q < delta] zeta (J_q - delta).
 *
 * The mask is not used (as in the reference).  The cost is summed in cell
 * order in `real` (float: the reference's FP32 sum; double: the FP64
 * trajectory the GPU solver follows).  The gradient is written per node and
 * cofactor, not with the reference's expanded expressions, so the FP32 build
 * agrees with the reference to rounding, not bit for bit.
 *
 * Pinning: the cost of the identity map has a closed form (every J_q = 1);
 * the gradient is checked against a finite-difference derivative of the cost
 * in double precision; whole solves are compared with the reference compiled
 * from its own sources (oracle/_ref).
 */

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
#else
typedef double real;
#endif

/* Result record of the matching solve. */
typedef struct {
    long iterations;
    int reason;            /* 6 iteration limit, 2 step below 1e-9, 5 max error <= epsilon */
    long cost_evaluations;
    long gradient_evaluations;
    double final_cost;
    double final_max_error;
    double *trace;         /* rows (it, cost, max error, step) */
    long trace_cap;
    long trace_len;
} oracle_match_result;

/* Corner q of a cell: bit 0 = +x (i), bit 1 = +y (j), bit 2 = +z (k).
 * Matrix columns: 0 = difference along y, 1 = along x, 2 = along z. */
static const int COL_BIT[3] = {2, 1, 4};

static inline void corner_matrix(const real v[3][8], int q, real m[3][3])
{
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) {
            const int b = COL_BIT[col];
            m[row][col] = v[row][q | b] - v[row][q & ~b];
        }
}

/* det m, expanded along the first row (the reference's corner Jacobian,
 * VolumeMatching3D.c:138-145) */
static inline real corner_det(const real m[3][3])
{
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
           m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

static inline void cofactors(const real m[3][3], real c[3][3])
{
    c[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    c[0][1] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    c[0][2] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    c[1][0] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    c[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    c[1][2] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
    c[2][0] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    c[2][1] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    c[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
}

/* The reference sums corner terms in the order of its names J_abc with
 * a = z, b = x, c = y (VolumeMatching3D.c:151-160): q = 0, 2, 1, 3, 4, 6, 5, 7. */
static const int REF_ORDER[8] = {0, 2, 1, 3, 4, 6, 5, 7};

typedef struct {
    size_t nx, ny, nz;
    real delta, zeta;
} mgrid;

static inline size_t node(const mgrid *g, size_t i, size_t j, size_t k)
{
    return (k * g->ny + j) * g->nx + i;
}

static void load_cell(const mgrid *g, const real *F, const real *G, const real *E, size_t i, size_t j,
                      size_t k, real v[3][8])
{
    for (int q = 0; q < 8; ++q) {
        const size_t n = node(g, i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
        v[0][q] = F[n];
        v[1][q] = G[n];
        v[2][q] = E[n];
    }
}

/* Cost and max voxel error (VolumeMatchingCost3D). */
static real match_cost(const mgrid *g, const real *F, const real *G, const real *E, const real *V,
                       real *max_err)
{
    real res = 0;
    real mx = 0;
    for (size_t k = 0; k + 1 < g->nz; ++k)
        for (size_t j = 0; j + 1 < g->ny; ++j)
            for (size_t i = 0; i + 1 < g->nx; ++i) {
                real v[3][8], m[3][3], J[8];
                load_cell(g, F, G, E, i, j, k, v);
                for (int q = 0; q < 8; ++q) {
                    corner_matrix(v, q, m);
                    J[q] = corner_det(m);
                }
                real S = 0;
                for (int t = 0; t < 8; ++t) S += J[REF_ORDER[t]];
                const real cv = V[node(g, i, j, k)];
                real cres = 0;
                if (cv > 0) cres += (S - 8 * cv) * (S - 8 * cv);
                for (int t = 0; t < 8; ++t) {
                    const real Jq = J[REF_ORDER[t]];
                    if (Jq < g->delta) cres += g->zeta * (Jq - g->delta) * (Jq - g->delta);
                }
                res += (real)(0.5 * cres);
                const real err = (real)(fabs((double)(S - 8 * cv)) / 8.0);
                if (err > mx) mx = err;
            }
    *max_err = mx;
    return res;
}

/* Gradient of the cost (VolumeMatchingGrads3D), accumulated per node in cell
 * order; dF, dG, dE are overwritten. */
static void match_grad(const mgrid *g, const real *F, const real *G, const real *E, const real *V, real *dF,
                       real *dG, real *dE)
{
    const size_t n3 = g->nx * g->ny * g->nz;
    memset(dF, 0, n3 * sizeof(real));
    memset(dG, 0, n3 * sizeof(real));
    memset(dE, 0, n3 * sizeof(real));
    real *d[3] = {dF, dG, dE};
    for (size_t k = 0; k + 1 < g->nz; ++k)
        for (size_t j = 0; j + 1 < g->ny; ++j)
            for (size_t i = 0; i + 1 < g->nx; ++i) {
                real v[3][8], m[3][3], C[8][3][3], J[8], W[8];
                load_cell(g, F, G, E, i, j, k, v);
                for (int q = 0; q < 8; ++q) {
                    corner_matrix(v, q, m);
                    J[q] = corner_det(m);
                    cofactors(m, C[q]);
                }
                real S = 0;
                for (int t = 0; t < 8; ++t) S += J[REF_ORDER[t]];
                const real cv = V[node(g, i, j, k)];
                const real w = cv > 0 ? (real)((double)S - 8.0 * (double)cv) : 0;
                for (int q = 0; q < 8; ++q)
                    W[q] = w + (J[q] < g->delta ? g->zeta * (J[q] - g->delta) : 0);
                for (int p = 0; p < 8; ++p) {
                    const size_t n = node(g, i + (p & 1), j + ((p >> 1) & 1), k + ((p >> 2) & 1));
                    for (int row = 0; row < 3; ++row) {
                        real acc = 0;
                        for (int col = 0; col < 3; ++col) {
                            const int b = COL_BIT[col];
                            const real sgn = (p & b) ? 1 : -1;
                            acc += sgn * (W[p] * C[p][row][col] + W[p ^ b] * C[p ^ b][row][col]);
                        }
                        d[row][n] += acc;
                    }
                }
            }
}

double oracle_match_cost(size_t nx, size_t ny, size_t nz, const real *F, const real *G, const real *E,
                         const real *V, real delta, real zeta, double *max_err)
{
    mgrid g = {nx, ny, nz, delta, zeta};
    real mx;
    const real c = match_cost(&g, F, G, E, V, &mx);
    if (max_err) *max_err = mx;
    return c;
}

void oracle_match_grad(size_t nx, size_t ny, size_t nz, const real *F, const real *G, const real *E,
                       const real *V, real delta, real zeta, real *dF, real *dG, real *dE)
{
    mgrid g = {nx, ny, nz, delta, zeta};
    match_grad(&g, F, G, E, V, dF, dG, dE);
}

static void step_to(size_t n3, real alpha, const real *x, const real *dx, real *out)
{
    for (size_t t = 0; t < n3; ++t) out[t] = x[t] - alpha * dx[t];
}

/* volume_matching_3d (:689-796) with VolumeMatching3D (:408-558): the field
 * [3][nz][ny][nx] is the initial guess on entry and the result on exit. */
int oracle_solve_matching(size_t nx, size_t ny, size_t nz, real dx, real dy, real dz, const real *target,
                          real epsilon, real delta, real zeta, size_t it_max, real *field,
                          oracle_match_result *res)
{
    const size_t n3 = nx * ny * nz;
    mgrid g = {nx, ny, nz, delta, zeta};
    real *buf = (real *)calloc(10 * n3, sizeof(real));
    if (!buf) return -1;
    real *X[3] = {buf, buf + n3, buf + 2 * n3};          /* F, G, E */
    real *D[3] = {buf + 3 * n3, buf + 4 * n3, buf + 5 * n3};
    real *T[3] = {buf + 6 * n3, buf + 7 * n3, buf + 8 * n3};
    real *V = buf + 9 * n3;
    const real *ux = field, *uy = field + n3, *uz = field + 2 * n3;
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t n = node(&g, i, j, k);
                X[0][n] = (real)(j + 1) + uy[n] / dy;
                X[1][n] = (real)(i + 1) + ux[n] / dx;
                X[2][n] = (real)(k + 1) + uz[n] / dz;
                V[n] = target[n];
            }
    res->iterations = 0;
    res->cost_evaluations = 0;
    res->gradient_evaluations = 0;
    res->trace_len = 0;

    real m;
    real H = match_cost(&g, X[0], X[1], X[2], V, &m);
    res->cost_evaluations++;
    real alpha = 1;
    size_t it = 0;
#define COST_OF_TRIAL()                                                                       \
    do {                                                                                      \
        for (int r_ = 0; r_ < 3; ++r_) step_to(n3, alpha, X[r_], D[r_], T[r_]);               \
        nH = match_cost(&g, T[0], T[1], T[2], V, &m);                                         \
        res->cost_evaluations++;                                                              \
    } while (0)
    while (it < it_max && alpha > 1e-9 && m > epsilon) {
        real Hmark = H, nH;
        match_grad(&g, X[0], X[1], X[2], V, D[0], D[1], D[2]);
        res->gradient_evaluations++;
        COST_OF_TRIAL();
        real Hnext;
        if (nH < Hmark) {
            if (it == 0) {
                /* grow the step while the cost keeps falling, then take the last good one */
                while (nH < Hmark) {
                    Hmark = nH;
                    alpha = (real)(alpha * 2.0);
                    COST_OF_TRIAL();
                }
                alpha = (real)(alpha * 0.5);
                for (int r_ = 0; r_ < 3; ++r_) step_to(n3, alpha, X[r_], D[r_], T[r_]);
                Hnext = Hmark;
            } else {
                Hnext = nH;
                alpha = (real)(alpha * 1.05);
            }
        } else {
            while (nH > Hmark && alpha > 1e-9) {
                alpha = (real)(alpha * 0.5);
                COST_OF_TRIAL();
            }
            Hnext = nH;
        }
        if (alpha > 1e-9) {
            for (int r_ = 0; r_ < 3; ++r_) memcpy(X[r_], T[r_], n3 * sizeof(real));
            H = Hnext;
            it++;
            if (res->trace && res->trace_len < res->trace_cap) {
                double *row = res->trace + 4 * res->trace_len;
                row[0] = (double)it; row[1] = H; row[2] = m; row[3] = alpha;
            }
            res->trace_len++;
        }
    }
#undef COST_OF_TRIAL
    res->iterations = (long)it;
    /* the first loop condition that fails, in the loop's order */
    res->reason = it >= it_max ? 6 : (!(alpha > 1e-9) ? 2 : 5);
    res->final_cost = H;
    res->final_max_error = m;
    real *fx = field, *fy = field + n3, *fz = field + 2 * n3;
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t n = node(&g, i, j, k);
                fx[n] = (X[1][n] - (real)(i + 1)) * dx;
                fy[n] = (X[0][n] - (real)(j + 1)) * dy;
                fz[n] = (X[2][n] - (real)(k + 1)) * dz;
            }
    free(buf);
    return 0;
}
