/*
 * Oracle: Ewald Coulomb sum in e^2/Angstrom.  TEST INFRASTRUCTURE ONLY.
 *
 * Real space + self term restate pme_energy_and_forces / pme_potential:
 *   src/potentials/coulomb/pme.f90:215-321 (phi = dE/dq, direct part)
 *   src/potentials/coulomb/pme.f90:386-437 (i<j pairs over every image,
 *     i==j images at half weight, erfc(beta r)/r and its radial derivative)
 *   src/potentials/coulomb/pme.f90:467-472 (self energy -beta/sqrt(pi) sum q^2)
 * with beta = sqrt(alpha) the Ewald splitting parameter (pme.f90:192,198).
 *
 * Reciprocal space: the reference sums it with smooth PME
 * (pme_kernel.f90:251-752); this file instead evaluates the textbook direct
 * structure-factor sum, which has no counterpart in the reference:
 *   E = (2 pi / V) sum_{k != 0} exp(-k^2 / 4 beta^2) / k^2 |S(k)|^2,
 *   S(k) = sum_j q_j exp(i k.r_j),  k = 2 pi Bbox^T n,  |k| <= kmax,
 * summed over the half space n > 0 (lexicographic) and doubled.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc.h"

#define PI_ 3.14159265358979323846
#define EPS_Q 1e-10

int orc_ewald_real(int nat, const double *r, const double *q,
                   const double *cell, const int64_t *offsets, const int *nbr,
                   const int *dc, double cutoff, double beta, double *epot,
                   double *f, double *wpot, double *phi, int nthreads)
{
    const double alpha = beta * beta;
    const double sqrt_alpha_pi = sqrt(alpha / PI_);
    const double cutoff_sq = cutoff * cutoff;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : 1;
#else
    int nt = 1;
    (void)nthreads;
#endif
    double *fbuf = calloc((size_t)nt * nat * 3 + 1, sizeof(double));
    double *pbuf = calloc((size_t)nt * nat + 1, sizeof(double));
    double *wbuf = calloc((size_t)nt * 9, sizeof(double));
    double *ebuf = calloc((size_t)nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *ft = fbuf + (size_t)tid * nat * 3;
        double *pt = pbuf + (size_t)tid * nat;
        double *w = wbuf + 9 * tid;
        double e = 0.0;
#pragma omp for schedule(static)
        for (int i = 0; i < nat; ++i) {
            double dq_i = q[i];
            if (!(fabs(dq_i) > EPS_Q)) continue;
            pt[i] -= 2 * q[i] * sqrt_alpha_pi;
            for (int64_t ni = offsets[i]; ni < offsets[i + 1]; ++ni) {
                int j = nbr[ni];
                if (!(fabs(q[j]) > EPS_Q)) continue;
                if (i > j) continue;
                double d[3];
                for (int x = 0; x < 3; ++x) {
                    double as = 0.0;
                    as = as + cell[3 * 0 + x] * (double)dc[3 * ni + 0];
                    as = as + cell[3 * 1 + x] * (double)dc[3 * ni + 1];
                    as = as + cell[3 * 2 + x] * (double)dc[3 * ni + 2];
                    d[x] = (r[3 * i + x] - r[3 * j + x]) + as;
                }
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (!(r2 < cutoff_sq)) continue;
                double ar = sqrt(r2);
                double ec = erfc(beta * ar);
                double hlp1 = ec / ar;
                double fr = 1.0 / (ar * ar * ar) * (ec + 2 * sqrt_alpha_pi * ar * exp(-alpha * r2));
                if (i < j) {
                    e += dq_i * q[j] * hlp1;
                    pt[i] += q[j] * hlp1;
                    pt[j] += dq_i * hlp1;
                    double hlp[3];
                    for (int x = 0; x < 3; ++x) {
                        hlp[x] = dq_i * q[j] * (d[x] * fr);
                        ft[3 * i + x] += hlp[x];
                        ft[3 * j + x] -= hlp[x];
                    }
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) w[a + 3 * b] -= d[a] * hlp[b];
                } else { /* i == j, periodic image */
                    e += 0.5 * q[j] * q[j] * hlp1;
                    pt[i] += q[j] * hlp1;
                    double hlp[3];
                    for (int x = 0; x < 3; ++x) hlp[x] = 0.5 * dq_i * q[j] * (d[x] * fr);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) w[a + 3 * b] -= d[a] * hlp[b];
                }
            }
        }
        ebuf[tid] = e;
    }
    double e = 0.0, eself = 0.0;
    for (int t = 0; t < nt; ++t) e += ebuf[t];
    for (int i = 0; i < nat; ++i) eself += q[i] * q[i];
    eself = eself * sqrt_alpha_pi;
    *epot += e - eself;
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 3 * nat; ++k) f[k] += fbuf[(size_t)t * nat * 3 + k];
        if (phi)
            for (int i = 0; i < nat; ++i) phi[i] += pbuf[(size_t)t * nat + i];
        for (int a = 0; a < 9; ++a) wpot[a] += wbuf[9 * t + a];
    }
    free(fbuf); free(pbuf); free(wbuf); free(ebuf);
    return ORC_OK;
}

int orc_ewald_kspace(int nat, const double *r, const double *q,
                     const double *cell, double beta, double kmax,
                     double *epot, double *f, double *wpot, double *phi,
                     int nthreads)
{
    double bbox[9];
    int err = orc_inverse_cell(cell, bbox);
    if (err) return err;
    /* V = |det(Abox)| */
    const double *a0 = cell, *a1 = cell + 3, *a2 = cell + 6;
    double V = fabs(a0[0] * (a1[1] * a2[2] - a1[2] * a2[1]) - a0[1] * (a1[0] * a2[2] - a1[2] * a2[0]) +
                    a0[2] * (a1[0] * a2[1] - a1[1] * a2[0]));
    int nmax[3];
    for (int k = 0; k < 3; ++k) {
        double len = sqrt(cell[3 * k] * cell[3 * k] + cell[3 * k + 1] * cell[3 * k + 1] +
                          cell[3 * k + 2] * cell[3 * k + 2]);
        nmax[k] = (int)floor(kmax * len / (2 * PI_)) + 1;
    }
    /* fractional coordinates s = Bbox r */
    double *s = malloc(sizeof(double) * 3 * (nat > 0 ? nat : 1));
    for (int j = 0; j < nat; ++j)
        for (int x = 0; x < 3; ++x)
            s[3 * j + x] = bbox[3 * 0 + x] * r[3 * j] + bbox[3 * 1 + x] * r[3 * j + 1] +
                           bbox[3 * 2 + x] * r[3 * j + 2];
    const double kmax2 = kmax * kmax, inv4b2 = 1.0 / (4 * beta * beta);
    double esum = 0.0, w[9] = {0};
    int nt = nthreads > 0 ? nthreads : 1;
    (void)nt;
    for (int n0 = 0; n0 <= nmax[0]; ++n0)
    for (int n1 = (n0 == 0 ? 0 : -nmax[1]); n1 <= nmax[1]; ++n1)
    for (int n2 = ((n0 == 0 && n1 == 0) ? 1 : -nmax[2]); n2 <= nmax[2]; ++n2) {
        double kv[3];
        for (int x = 0; x < 3; ++x)
            kv[x] = 2 * PI_ * (n0 * bbox[3 * x + 0] + n1 * bbox[3 * x + 1] + n2 * bbox[3 * x + 2]);
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        if (k2 > kmax2) continue;
        double c = exp(-k2 * inv4b2) / k2;
        double Sc = 0.0, Ss = 0.0;
#pragma omp parallel for reduction(+ : Sc, Ss) num_threads(nt)
        for (int j = 0; j < nat; ++j) {
            double th = 2 * PI_ * (n0 * s[3 * j] + n1 * s[3 * j + 1] + n2 * s[3 * j + 2]);
            Sc += q[j] * cos(th);
            Ss += q[j] * sin(th);
        }
        double S2 = Sc * Sc + Ss * Ss;
        esum += c * S2;
        double fac = 2 * (inv4b2 + 1.0 / k2);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                w[a + 3 * b] += c * S2 * (fac * kv[a] * kv[b] - (a == b ? 1.0 : 0.0));
#pragma omp parallel for num_threads(nt)
        for (int j = 0; j < nat; ++j) {
            double th = 2 * PI_ * (n0 * s[3 * j] + n1 * s[3 * j + 1] + n2 * s[3 * j + 2]);
            double cs = cos(th), sn = sin(th);
            double im = sn * Sc - cs * Ss;
            for (int x = 0; x < 3; ++x) f[3 * j + x] += (8 * PI_ / V) * q[j] * c * kv[x] * im;
            if (phi) phi[j] += (8 * PI_ / V) * c * (Sc * cs + Ss * sn);
        }
    }
    *epot += (4 * PI_ / V) * esum;
    for (int a = 0; a < 9; ++a) wpot[a] += (4 * PI_ / V) * w[a];
    free(s);
    return ORC_OK;
}
