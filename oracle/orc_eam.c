/*
 * Oracle: natural cubic spline tables and the tabulated (setfl) EAM kernel.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates:
 *   - simple_spline_init     src/support/simple_spline.f90:127-195
 *   - simple_spline_f / df   src/support/simple_spline.f90:373-528
 *   - simple_spline_f_and_df (extrapolate) src/support/simple_spline.f90:536-615
 *   - tabulated_alloy_eam_energy_and_forces_kernel
 *       src/potentials/eam/tabulated_alloy_eam.f90:423-627
 * The caller (setfl reader, tabulated_alloy_eam.f90:147-259) passes one
 * coefficient block per table; pair tables hold r*phi(r) already scaled by 1/2
 * (scale_y_axis, simple_spline.f90:756-773).  el[] holds database indices
 * (el2db), -1 for atoms the potential ignores.
 *
 * funcfl != 0: the single-element TabulatedEAM kernel
 * (src/potentials/eam/tabulated_eam.f90:334-511) instead: the pair table is
 * the effective charge Z(r) scaled by sqrt(Hartree Bohr / 2) (:178), and the
 * pair energy per directed pair is Z^2 / r with derivative 2 Z Z' / r -
 * Z^2 / r^2 (:446-450), i.e. the r*phi form above with r*phi = Z^2.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc.h"

int orc_spline_init(int n, double x0, double dx, const double *y, double *coef)
{
    (void)x0;
    if (n < 3) return ORC_ERR_ARG;
    const double sig = 0.5;
    double *yy = coef, *d2y = coef + n, *c1 = coef + 2 * n, *c2 = coef + 3 * n,
           *c3 = coef + 4 * n, *dc1 = coef + 5 * n, *dc2 = coef + 6 * n,
           *dc3 = coef + 7 * n;
    double *u = malloc(sizeof(double) * n);
    memcpy(yy, y, sizeof(double) * n);
    d2y[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        double p = sig * d2y[i - 1] + 2;
        d2y[i] = (sig - 1) / p;
        u[i] = (6.0 * ((yy[i + 1] - yy[i]) / dx - (yy[i] - yy[i - 1]) / dx) / (2 * dx) -
                sig * u[i - 1]) / p;
    }
    double qn = 0.0, un = 0.0;
    d2y[n - 1] = (un - qn * u[n - 2]) / (qn * d2y[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; --k) d2y[k] = d2y[k] * d2y[k + 1] + u[k];
    double dx2 = dx * dx;
    for (int k = 0; k < n - 1; ++k) {
        c1[k] = yy[k + 1] - yy[k] - (2 * d2y[k] + d2y[k + 1]) * dx2 / 6;
        c2[k] = d2y[k] * dx2 / 2;
        c3[k] = (d2y[k + 1] - d2y[k]) * dx2 / 6;
    }
    for (int k = 0; k < n - 1; ++k) {
        dc1[k] = c1[k] / dx;
        dc2[k] = 2 * c2[k] / dx;
        dc3[k] = 3 * c3[k] / dx;
    }
    c1[n - 1] = c2[n - 1] = c3[n - 1] = dc1[n - 1] = dc2[n - 1] = dc3[n - 1] = 0.0;
    free(u);
    return ORC_OK;
}

/* simple_spline_f (no extrapolation); returns nonzero outside the table */
static int spl_f(int n, double x0, double dx, const double *coef, double x, double *val)
{
    double cut = x0 + dx * (n - 1);
    double xf;
    int i;
    if (x == cut) { xf = n; i = n - 1; }
    else { xf = (x - x0) / dx + 1; i = (int)floor(xf); }
    if (i < 1 || i >= n) { *val = 1.0; return ORC_ERR_SPLINE; }
    double B = xf - i;
    int k = i - 1;
    *val = coef[k] + B * (coef[2 * n + k] + B * (coef[3 * n + k] + B * coef[4 * n + k]));
    return ORC_OK;
}

static int spl_df(int n, double x0, double dx, const double *coef, double x, double *val)
{
    double cut = x0 + dx * (n - 1);
    double xf;
    int i;
    if (x == cut) { xf = n; i = n - 1; }
    else { xf = (x - x0) / dx + 1; i = (int)floor(xf); }
    if (i < 1 || i >= n) { *val = 1.0; return ORC_ERR_SPLINE; }
    double B = xf - i;
    int k = i - 1;
    *val = coef[5 * n + k] + B * (coef[6 * n + k] + B * coef[7 * n + k]);
    return ORC_OK;
}

/* simple_spline_f_and_df(..., extrapolate=.true.) */
static void spl_f_df_extrap(int n, double x0, double dx, const double *coef, double x,
                            double *f, double *df)
{
    double xf = (x - x0) / dx + 1;
    int i = (int)floor(xf);
    if (i < 1) i = 1;
    else if (i >= n) i = n - 1;
    double B = xf - i;
    int k = i - 1;
    *f = coef[k] + B * (coef[2 * n + k] + B * (coef[3 * n + k] + B * coef[4 * n + k]));
    *df = coef[5 * n + k] + B * (coef[6 * n + k] + B * coef[7 * n + k]);
}

int orc_eam(int nel, int nat, const double *r, const int *el,
            const double *cell, const int64_t *offsets, const int *nbr,
            const int *dc, const int *mask, double cutoff,
            const int *n_F, const double *dF, const double *const *coef_F,
            const int *n_rho, const double *drho,
            const double *const *coef_rho, const int *n_phi,
            const double *dphi, const double *const *coef_phi,
            double *epot, double *f, double *wpot, double *epot_per_at,
            double *wpot_per_at, int funcfl, int nthreads)
{
    double cutoff_sq = cutoff * cutoff;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : 1;
#else
    int nt = 1;
    (void)nthreads;
#endif
    double *fbuf = calloc((size_t)nt * nat * 3 + 1, sizeof(double));
    double *ebuf = calloc((size_t)nt * nat + 1, sizeof(double));
    double *wbuf = calloc((size_t)nt * 9, sizeof(double));
    int err = ORC_OK;
#pragma omp parallel num_threads(nt) reduction(| : err)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *ft = fbuf + (size_t)tid * nat * 3;
        double *et = ebuf + (size_t)tid * nat;
        double *w = wbuf + 9 * tid;
        int cap = 64;
        int *neb = malloc(sizeof(int) * cap);
        double *neb_dr = malloc(sizeof(double) * 4 * cap);
#pragma omp for schedule(static)
        for (int i = 0; i < nat; ++i) {
            if (mask && mask[i] == 0) continue;
            int dbi = el[i];
            if (dbi < 0) continue;
            double rho = 0.0;
            int neb_n = 0;
            for (int64_t ni = offsets[i]; ni < offsets[i + 1]; ++ni) {
                int j = nbr[ni];
                int dbj = el[j];
                if (dbj < 0) continue;
                double d[3];
                for (int x = 0; x < 3; ++x) {
                    double as = 0.0;
                    as = as + cell[3 * 0 + x] * (double)dc[3 * ni + 0];
                    as = as + cell[3 * 1 + x] * (double)dc[3 * ni + 1];
                    as = as + cell[3 * 2 + x] * (double)dc[3 * ni + 2];
                    d[x] = (r[3 * i + x] - r[3 * j + x]) + as;
                }
                double ad = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (ad < cutoff_sq) {
                    ad = sqrt(ad);
                    double dr_;
                    err |= spl_f(n_rho[dbj], 0.0, drho[dbj], coef_rho[dbj], ad, &dr_);
                    rho = rho + dr_;
                    if (neb_n == cap) {
                        cap *= 2;
                        neb = realloc(neb, sizeof(int) * cap);
                        neb_dr = realloc(neb_dr, sizeof(double) * 4 * cap);
                    }
                    neb[neb_n] = j;
                    neb_dr[4 * neb_n] = d[0];
                    neb_dr[4 * neb_n + 1] = d[1];
                    neb_dr[4 * neb_n + 2] = d[2];
                    neb_dr[4 * neb_n + 3] = ad;
                    ++neb_n;
                }
            }
            if (rho < 0.0) rho = 0.0;
            double Fi, dFi;
            spl_f_df_extrap(n_F[dbi], 0.0, dF[dbi], coef_F[dbi], rho, &Fi, &dFi);
            et[i] += Fi;
            double fori[3] = {0, 0, 0};
            for (int n = 0; n < neb_n; ++n) {
                int j = neb[n];
                int dbj = el[j];
                const double *d = &neb_dr[4 * n];
                double ad = d[3];
                int pidx = dbi * nel + dbj;
                double phi, dphi_;
                spl_f_df_extrap(n_phi[pidx], 0.0, dphi[pidx], coef_phi[pidx], ad, &phi, &dphi_);
                if (funcfl) {
                    /* Z(r): r phi = Z^2, (r phi)' = 2 Z Z' */
                    dphi_ = 2 * phi * dphi_;
                    phi = phi * phi;
                }
                double rad = 1.0 / ad;
                et[i] += phi * rad;
                double fac;
                err |= spl_df(n_rho[dbj], 0.0, drho[dbj], coef_rho[dbj], ad, &fac);
                double df[3];
                for (int x = 0; x < 3; ++x) {
                    df[x] = -(dFi * fac + (dphi_ - phi * rad) * rad) * rad * d[x];
                    fori[x] += df[x];
                    ft[3 * j + x] -= df[x];
                }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double wij = -d[a] * df[b];
                        w[a + 3 * b] += wij;
                        if (wpot_per_at) {
#pragma omp atomic
                            wpot_per_at[9 * i + a + 3 * b] += wij / 2;
#pragma omp atomic
                            wpot_per_at[9 * j + a + 3 * b] += wij / 2;
                        }
                    }
            }
            for (int x = 0; x < 3; ++x) ft[3 * i + x] += fori[x];
        }
        free(neb);
        free(neb_dr);
    }
    double e = 0.0;
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < nat; ++i) e += ebuf[(size_t)t * nat + i];
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 3 * nat; ++k) f[k] += fbuf[(size_t)t * nat * 3 + k];
        if (epot_per_at)
            for (int i = 0; i < nat; ++i) epot_per_at[i] += ebuf[(size_t)t * nat + i];
        for (int a = 0; a < 9; ++a) wpot[a] += wbuf[9 * t + a];
    }
    *epot += e;
    free(fbuf); free(ebuf); free(wbuf);
    return err;
}
