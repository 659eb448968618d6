/*
 * Oracle: cell inverse and the cell/bin neighbour-list builder.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates, in order of operations:
 *   - particles_set_cell / gaussn -> dgesv (partial-pivot LU)
 *       src/python/f90/python_particles.f90:286-346
 *       src/support/f_linearalgebra.f90:559-637
 *   - neighbors_binning_init   src/python/f90/python_neighbors.f90:765-871
 *   - neighbors_binning_update src/python/f90/python_neighbors.f90:904-959
 *   - fill_neighbor_list       src/python/f90/python_neighbors.f90:570-754
 * and an independent brute-force image enumerator (the closed-form set
 * {(j, n): |r_i - r_j + A n|^2 < rc^2, (j, n) != (i, 0)}).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define A_(c, x, k) ((c)[3 * (k) + (x)])

/* dgesv(3, 3, A, 3, ipiv, I, 3) with the unblocked right-looking LU. */
int orc_inverse_cell(const double *cell, double *bbox)
{
    double a[9], b[9];
    int ipiv[3];
    memcpy(a, cell, sizeof a);
    for (int k = 0; k < 9; ++k) b[k] = 0.0;
    b[0] = b[4] = b[8] = 1.0;
    for (int j = 0; j < 3; ++j) {
        int p = j;
        double amax = fabs(A_(a, j, j));
        for (int i = j + 1; i < 3; ++i)
            if (fabs(A_(a, i, j)) > amax) { amax = fabs(A_(a, i, j)); p = i; }
        ipiv[j] = p;
        if (A_(a, p, j) == 0.0) return ORC_ERR_SINGULAR;
        if (p != j)
            for (int k = 0; k < 3; ++k) {
                double t = A_(a, j, k); A_(a, j, k) = A_(a, p, k); A_(a, p, k) = t;
            }
        double rp = 1.0 / A_(a, j, j);
        for (int i = j + 1; i < 3; ++i) A_(a, i, j) = A_(a, i, j) * rp;
        for (int k = j + 1; k < 3; ++k)
            for (int i = j + 1; i < 3; ++i)
                A_(a, i, k) = A_(a, i, k) - A_(a, i, j) * A_(a, j, k);
    }
    /* dgetrs: row swaps, unit-lower forward solve, upper back solve */
    for (int j = 0; j < 3; ++j)
        if (ipiv[j] != j)
            for (int c = 0; c < 3; ++c) {
                double t = A_(b, j, c); A_(b, j, c) = A_(b, ipiv[j], c); A_(b, ipiv[j], c) = t;
            }
    for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < 3; ++k)
            if (A_(b, k, c) != 0.0)
                for (int i = k + 1; i < 3; ++i)
                    A_(b, i, c) = A_(b, i, c) - A_(b, k, c) * A_(a, i, k);
        for (int k = 2; k >= 0; --k)
            if (A_(b, k, c) != 0.0) {
                A_(b, k, c) = A_(b, k, c) / A_(a, k, k);
                for (int i = 0; i < k; ++i)
                    A_(b, i, c) = A_(b, i, c) - A_(b, k, c) * A_(a, i, k);
            }
    }
    memcpy(bbox, b, sizeof b);
    return ORC_OK;
}

static void cross(const double *u, const double *v, double *w)
{
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

static double dot3(const double *u, const double *v)
{
    double s = 0.0;
    s = s + u[0] * v[0];
    s = s + u[1] * v[1];
    s = s + u[2] * v[2];
    return s;
}

/* binning_init: number of bins per axis and stencil half widths. */
int orc_nl_geometry(const double *cell, const int *pbc, double cutoff,
                    double bin_size, int *n_cells, int *n_stencil)
{
    (void)pbc;
    if (bin_size <= 0.0) bin_size = cutoff;
    double cs[3][3];
    for (int x = 0; x < 3; ++x) {
        double len = sqrt(dot3(&cell[3 * x], &cell[3 * x]));
        int n = (int)(len / bin_size);
        if (n < 3) n = 3;
        n_cells[x] = n;
    }
    for (int x = 0; x < 3; ++x)
        for (int c = 0; c < 3; ++c) cs[x][c] = cell[3 * x + c] / n_cells[x];
    double nx[3], ny[3], nz[3];
    cross(cs[1], cs[2], nx);
    cross(cs[2], cs[0], ny);
    cross(cs[0], cs[1], nz);
    double cv = dot3(cs[0], nx);
    double *nn[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        double d2 = dot3(nn[a], nn[a]);
        for (int c = 0; c < 3; ++c) nn[a][c] = cv * nn[a][c] / d2;
        n_stencil[a] = (int)(cutoff / sqrt(dot3(nn[a], nn[a]))) + 1;
    }
    return ORC_OK;
}

/* floor(matmul(rec_cell_size, r)) + 1, summed k = 1..3 without contraction */
static void cell_of(const double *rec, const double *r, int *c)
{
    for (int x = 0; x < 3; ++x) {
        double s = 0.0;
        s = s + rec[3 * 0 + x] * r[0];
        s = s + rec[3 * 1 + x] * r[1];
        s = s + rec[3 * 2 + x] * r[2];
        c[x] = (int)floor(s) + 1;
    }
}

typedef struct {
    const double *r, *cell, *rec;
    const int *pbc, *nc, *nd, *bseed, *next;
    double cutoff_sq;
} nl_ctx;

/* Stencil walk of atom i (the body of fill_neighbor_list's i_loop).  With
 * out == NULL only counts; otherwise writes at most cnt_max entries. */
static int64_t walk_atom(const nl_ctx *c, int i, int *nbr, int *dc, int64_t cnt_max)
{
    const double *r = c->r, *cell = c->cell;
    const int *pbc = c->pbc, *nc = c->nc, *nd = c->nd;
    int ci[3], shift[3] = {0, 0, 0};
    cell_of(c->rec, &r[3 * i], ci);
    for (int k = 0; k < 3; ++k) {
        if (pbc[k]) {
            while (ci[k] < 1) { ci[k] += nc[k]; shift[k] += 1; }
            while (ci[k] > nc[k]) { ci[k] -= nc[k]; shift[k] -= 1; }
        } else {
            if (ci[k] < 1) ci[k] = 1;
            if (ci[k] > nc[k]) ci[k] = nc[k];
        }
    }
    int64_t cnt = 0;
    for (int dx = -nd[0]; dx <= nd[0]; ++dx)
    for (int dy = -nd[1]; dy <= nd[1]; ++dy)
    for (int dz = -nd[2]; dz <= nd[2]; ++dz) {
        int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
        int s1[3] = {shift[0], shift[1], shift[2]};
        for (int k = 0; k < 3; ++k)
            if (pbc[k]) {
                while (cc[k] < 1) { cc[k] += nc[k]; s1[k] += 1; }
                while (cc[k] > nc[k]) { cc[k] -= nc[k]; s1[k] -= 1; }
            }
        if (cc[0] < 1 || cc[1] < 1 || cc[2] < 1 || cc[0] > nc[0] ||
            cc[1] > nc[1] || cc[2] > nc[2])
            continue;
        int64_t lin = ((int64_t)(cc[0] - 1) * nc[1] + (cc[1] - 1)) * nc[2] + (cc[2] - 1);
        for (int j = c->bseed[lin]; j != -1; j = c->next[j]) {
            int cj[3], s2[3] = {s1[0], s1[1], s1[2]};
            cell_of(c->rec, &r[3 * j], cj);
            for (int k = 0; k < 3; ++k)
                if (pbc[k]) {
                    while (cj[k] < 1) { cj[k] += nc[k]; s2[k] -= 1; }
                    while (cj[k] > nc[k]) { cj[k] -= nc[k]; s2[k] += 1; }
                }
            if (i == j && s2[0] == 0 && s2[1] == 0 && s2[2] == 0) continue;
            double d[3];
            for (int x = 0; x < 3; ++x) {
                double as = 0.0;
                as = as + A_(cell, x, 0) * (double)s2[0];
                as = as + A_(cell, x, 1) * (double)s2[1];
                as = as + A_(cell, x, 2) * (double)s2[2];
                d[x] = (r[3 * i + x] - r[3 * j + x]) + as;
            }
            if (dot3(d, d) < c->cutoff_sq) {
                if (nbr && cnt < cnt_max) {
                    nbr[cnt] = j;
                    dc[3 * cnt + 0] = s2[0];
                    dc[3 * cnt + 1] = s2[1];
                    dc[3 * cnt + 2] = s2[2];
                }
                ++cnt;
            }
        }
    }
    return cnt;
}

/* Two-pass OpenMP form (count, scan, fill) of the same list; the output is
 * identical to the serial walk. */
int orc_nl_build_mt(int nat, const double *r, const double *cell, const int *pbc,
                    double cutoff, double bin_size, int64_t cap,
                    int64_t *offsets, int *nbr, int *dc, int nthreads);

int orc_nl_build(int nat, const double *r, const double *cell, const int *pbc,
                 double cutoff, double bin_size, int64_t cap,
                 int64_t *offsets, int *nbr, int *dc)
{
    double bbox[9], rec[9];
    int nc[3], nd[3];
    int err = orc_inverse_cell(cell, bbox);
    if (err) return err;
    orc_nl_geometry(cell, pbc, cutoff, bin_size, nc, nd);
    /* rec_cell_size(x, :) = Bbox(x, :) * n_cells(x); column-major image */
    for (int x = 0; x < 3; ++x)
        for (int k = 0; k < 3; ++k) rec[3 * k + x] = bbox[3 * k + x] * nc[x];
    double cutoff_sq = cutoff * cutoff;

    /* binning_update: linked lists in ascending atom order */
    int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    int *bseed = malloc(sizeof(int) * ncell);
    int *blast = malloc(sizeof(int) * ncell);
    int *next = malloc(sizeof(int) * (nat > 0 ? nat : 1));
    for (int64_t c = 0; c < ncell; ++c) bseed[c] = blast[c] = -1;
    for (int i = 0; i < nat; ++i) next[i] = -1;
    for (int i = 0; i < nat; ++i) {
        int c[3];
        cell_of(rec, &r[3 * i], c);
        for (int k = 0; k < 3; ++k) {
            if (pbc[k]) {
                while (c[k] < 1) c[k] += nc[k];
                while (c[k] > nc[k]) c[k] -= nc[k];
            } else {
                if (c[k] < 1) c[k] = 1;
                if (c[k] > nc[k]) c[k] = nc[k];
            }
        }
        int64_t lin = ((int64_t)(c[0] - 1) * nc[1] + (c[1] - 1)) * nc[2] + (c[2] - 1);
        if (bseed[lin] == -1) bseed[lin] = blast[lin] = i;
        else { next[blast[lin]] = i; blast[lin] = i; }
    }

    int64_t cur = 0;        /* neighbours written */
    int err_out = ORC_OK;
    for (int i = 0; i < nat; ++i) {
        int ci[3], shift[3] = {0, 0, 0};
        cell_of(rec, &r[3 * i], ci);
        for (int k = 0; k < 3; ++k) {
            if (pbc[k]) {
                while (ci[k] < 1) { ci[k] += nc[k]; shift[k] += 1; }
                while (ci[k] > nc[k]) { ci[k] -= nc[k]; shift[k] -= 1; }
            } else {
                if (ci[k] < 1) ci[k] = 1;
                if (ci[k] > nc[k]) ci[k] = nc[k];
            }
        }
        offsets[i] = cur;
        for (int dx = -nd[0]; dx <= nd[0]; ++dx)
        for (int dy = -nd[1]; dy <= nd[1]; ++dy)
        for (int dz = -nd[2]; dz <= nd[2]; ++dz) {
            int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
            int s1[3] = {shift[0], shift[1], shift[2]};
            for (int k = 0; k < 3; ++k)
                if (pbc[k]) {
                    while (cc[k] < 1) { cc[k] += nc[k]; s1[k] += 1; }
                    while (cc[k] > nc[k]) { cc[k] -= nc[k]; s1[k] -= 1; }
                }
            if (cc[0] < 1 || cc[1] < 1 || cc[2] < 1 || cc[0] > nc[0] ||
                cc[1] > nc[1] || cc[2] > nc[2])
                continue;
            int64_t lin = ((int64_t)(cc[0] - 1) * nc[1] + (cc[1] - 1)) * nc[2] + (cc[2] - 1);
            for (int j = bseed[lin]; j != -1; j = next[j]) {
                int cj[3], s2[3] = {s1[0], s1[1], s1[2]};
                cell_of(rec, &r[3 * j], cj);
                for (int k = 0; k < 3; ++k)
                    if (pbc[k]) {
                        while (cj[k] < 1) { cj[k] += nc[k]; s2[k] -= 1; }
                        while (cj[k] > nc[k]) { cj[k] -= nc[k]; s2[k] += 1; }
                    }
                if (i == j && s2[0] == 0 && s2[1] == 0 && s2[2] == 0) continue;
                double d[3];
                for (int x = 0; x < 3; ++x) {
                    double as = 0.0;
                    as = as + A_(cell, x, 0) * (double)s2[0];
                    as = as + A_(cell, x, 1) * (double)s2[1];
                    as = as + A_(cell, x, 2) * (double)s2[2];
                    d[x] = (r[3 * i + x] - r[3 * j + x]) + as;
                }
                double d2 = dot3(d, d);
                if (d2 < cutoff_sq) {
                    /* serial overflow rule: the slot (1-based, counting one
                     * separator per finished atom) must stay below cap */
                    if (cur + i + 1 >= cap) { err_out = ORC_ERR_OVERFLOW; continue; }
                    nbr[cur] = j;
                    dc[3 * cur + 0] = s2[0];
                    dc[3 * cur + 1] = s2[1];
                    dc[3 * cur + 2] = s2[2];
                    ++cur;
                }
            }
        }
    }
    offsets[nat] = cur;
    free(bseed); free(blast); free(next);
    return err_out;
}

/* Closed-form image set, enumerated over a box of lattice translations large
 * enough to contain every image within the cutoff.  Output order: i, then j,
 * then n lexicographic.  Same distance arithmetic as the builder. */
int orc_nl_brute(int nat, const double *r, const double *cell, const int *pbc,
                 double cutoff, int64_t cap, int64_t *offsets, int *nbr,
                 int *dc)
{
    double bbox[9];
    int err = orc_inverse_cell(cell, bbox);
    if (err) return err;
    double cutoff_sq = cutoff * cutoff;
    /* |n_k| range: fractional extent of a cutoff sphere along k is
     * rc * |row k of Bbox| */
    double hw[3];
    for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int x = 0; x < 3; ++x) s += bbox[3 * x + k] * bbox[3 * x + k];
        hw[k] = cutoff * sqrt(s);
    }
    int64_t cur = 0;
    for (int i = 0; i < nat; ++i) {
        offsets[i] = cur;
        for (int j = 0; j < nat; ++j) {
            double fs[3];
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int x = 0; x < 3; ++x) s += bbox[3 * x + k] * (r[3 * i + x] - r[3 * j + x]);
                fs[k] = s;
            }
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                if (pbc[k]) {
                    lo[k] = (int)floor(-fs[k] - hw[k]) - 1;
                    hi[k] = (int)ceil(-fs[k] + hw[k]) + 1;
                } else lo[k] = hi[k] = 0;
            }
            for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
            for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
            for (int n2 = lo[2]; n2 <= hi[2]; ++n2) {
                if (i == j && n0 == 0 && n1 == 0 && n2 == 0) continue;
                int s2[3] = {n0, n1, n2};
                double d[3];
                for (int x = 0; x < 3; ++x) {
                    double as = 0.0;
                    as = as + A_(cell, x, 0) * (double)s2[0];
                    as = as + A_(cell, x, 1) * (double)s2[1];
                    as = as + A_(cell, x, 2) * (double)s2[2];
                    d[x] = (r[3 * i + x] - r[3 * j + x]) + as;
                }
                if (dot3(d, d) < cutoff_sq) {
                    if (cur >= cap) return ORC_ERR_OVERFLOW;
                    nbr[cur] = j;
                    dc[3 * cur] = n0; dc[3 * cur + 1] = n1; dc[3 * cur + 2] = n2;
                    ++cur;
                }
            }
        }
    }
    offsets[nat] = cur;
    return ORC_OK;
}

int orc_nl_build_mt(int nat, const double *r, const double *cell, const int *pbc,
                    double cutoff, double bin_size, int64_t cap,
                    int64_t *offsets, int *nbr, int *dc, int nthreads)
{
    double bbox[9], rec[9];
    int nc[3], nd[3];
    int err = orc_inverse_cell(cell, bbox);
    if (err) return err;
    orc_nl_geometry(cell, pbc, cutoff, bin_size, nc, nd);
    for (int x = 0; x < 3; ++x)
        for (int k = 0; k < 3; ++k) rec[3 * k + x] = bbox[3 * k + x] * nc[x];
    int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    int *bseed = malloc(sizeof(int) * ncell);
    int *blast = malloc(sizeof(int) * ncell);
    int *next = malloc(sizeof(int) * (nat > 0 ? nat : 1));
    for (int64_t c = 0; c < ncell; ++c) bseed[c] = blast[c] = -1;
    for (int i = 0; i < nat; ++i) next[i] = -1;
    for (int i = 0; i < nat; ++i) {
        int c[3];
        cell_of(rec, &r[3 * i], c);
        for (int k = 0; k < 3; ++k) {
            if (pbc[k]) {
                while (c[k] < 1) c[k] += nc[k];
                while (c[k] > nc[k]) c[k] -= nc[k];
            } else {
                if (c[k] < 1) c[k] = 1;
                if (c[k] > nc[k]) c[k] = nc[k];
            }
        }
        int64_t lin = ((int64_t)(c[0] - 1) * nc[1] + (c[1] - 1)) * nc[2] + (c[2] - 1);
        if (bseed[lin] == -1) bseed[lin] = blast[lin] = i;
        else { next[blast[lin]] = i; blast[lin] = i; }
    }
    nl_ctx ctx = {r, cell, rec, pbc, nc, nd, bseed, next, cutoff * cutoff};
    int64_t *cnt = malloc(sizeof(int64_t) * (nat > 0 ? nat : 1));
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < nat; ++i) cnt[i] = walk_atom(&ctx, i, NULL, NULL, 0);
    int64_t tot = 0;
    int err_out = ORC_OK;
    for (int i = 0; i < nat; ++i) {
        offsets[i] = tot;
        tot += cnt[i];
        if (cnt[i] > 0 && tot + i >= cap) err_out = ORC_ERR_OVERFLOW;
    }
    offsets[nat] = tot;
    if (err_out == ORC_OK) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
        for (int i = 0; i < nat; ++i)
            walk_atom(&ctx, i, nbr + offsets[i], dc + 3 * offsets[i], cnt[i]);
    }
    free(cnt); free(bseed); free(blast); free(next);
    return err_out;
}
