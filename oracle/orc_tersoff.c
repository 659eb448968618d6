/*
 * Oracle: Tersoff energy, forces and virial through the generic bond-order
 * kernel.  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates:
 *   - BOP_KERNEL loop1 (internal bond list, trig_off cutoff)
 *       src/potentials/bop/bop_kernel.f90:563-1068
 *       src/support/cutoff.f90:152-196
 *   - BOP_KERNEL loop2 (zeta, b_ij, forces on i, j, k, virial)
 *       src/potentials/bop/bop_kernel.f90:1075-1529
 *   - reduction epot += 1/2 sum(pe), epot_per_at = pe/2
 *       src/potentials/bop/bop_kernel.f90:1613-1628
 *   - VA, VR, g, bo, h    src/potentials/bop/tersoff/tersoff_func.f90:32-211
 *   - fam = 1: g, bo, h of src/potentials/bop/kumagai/kumagai_func.f90:86-221
 *     (VA, VR have the Tersoff form, :32-78)
 *   - fam = 2: VA, VR, g, bo of src/potentials/bop/brenner/brenner_func.f90:27-147
 *     with the constants of BIND_TO_FUNC (brenner_module.f90:267-289); its h
 *     (:153-196) is the Tersoff form
 *   - PAIR_INDEX          src/macros.inc:123
 *   - element map / nebmax src/potentials/bop/default_compute_func.f90:62-73
 * The OpenMP split (per-thread force/pe buffers summed in thread order) is the
 * model of src/support/tls.f90:58-269; with nthreads = 1 it is the serial loop.
 *
 * Virial layout: wpot[a + 3*b] = W(a, b) (Fortran wpot(3,3) image),
 * wpot_per_at[9*i + a + 3*b].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc.h"

#define PI_ 3.14159265358979323846

static int pair_index(int i, int j, int nel)
{
    int a = i + j * nel, b = j + i * nel;
    int c = i * (i + 1) / 2, d = j * (j + 1) / 2;
    return (a < b ? a : b) - (c < d ? c : d);
}

typedef struct {
    int j, typ;
    int dc[3];
    double len, nm[3], fc, dfc;
} bond_t;

static void trig_off(double r1, double r2, double r, double *val, double *dval)
{
    if (r <= r1) { *val = 1.0; *dval = 0.0; }
    else if (r >= r2) { *val = 0.0; *dval = 0.0; }
    else {
        double fac = PI_ / (r2 - r1);
        double x = fac * (r - r1);
        *val = 0.5 * (1.0 + cos(x));
        *dval = -0.5 * fac * sin(x);
    }
}

static void h_fn(const orc_tersoff_params *p, int ikpot, double dr, double *val, double *dval)
{
    if (p->fam == 1) {
        /* kumagai_func.f90:179-221 (the beta = 3 branch has no alpha in the
         * exponent, as written there) */
        double alpha = p->kalpha[ikpot];
        if (alpha == 0.0) { *val = 1.0; *dval = 0.0; return; }
        int beta = p->kbeta[ikpot];
        if (beta == 1) { *val = exp(alpha * dr); *dval = alpha * *val; }
        else if (beta == 3) { *val = exp(dr * dr * dr); *dval = 3 * alpha * dr * dr * *val; }
        else { *val = exp(alpha * pow(dr, beta)); *dval = beta * alpha * pow(dr, beta - 1) * *val; }
        return;
    }
    double mu = p->mubo[ikpot];
    if (mu == 0.0) { *val = 1.0; *dval = 0.0; return; }
    int m = p->m[ikpot];
    if (m == 1) { *val = exp(2 * mu * dr); *dval = 2 * mu * *val; }
    else if (m == 3) {
        double arg = 2 * mu * dr;
        *val = exp(arg * arg * arg);
        *dval = 2 * mu * m * arg * arg * *val;
    } else {
        *val = exp(pow(2 * mu * dr, m));
        *dval = 2 * mu * m * pow(2 * mu * dr, m - 1) * *val;
    }
}

int orc_tersoff(const orc_tersoff_params *p, int nat, const double *r,
                const int *el, const double *cell, const int64_t *offsets,
                const int *nbr, const int *dc, const int *mask,
                double *epot, double *f, double *wpot, double *epot_per_at,
                double *wpot_per_at, int nthreads)
{
    const int nel = p->nel;
    /* loop1: internal bond lists */
    bond_t *bonds = malloc(sizeof(bond_t) * (offsets[nat] > 0 ? offsets[nat] : 1));
    int64_t *bseed = malloc(sizeof(int64_t) * (nat + 1));
    int64_t *blast = malloc(sizeof(int64_t) * (nat + 1));
    int nebmax = 0;
    for (int i = 0; i < nat; ++i) {
        bseed[i] = offsets[i];
        int64_t nb = offsets[i];
        if (el[i] >= 0) {
            for (int64_t jn = offsets[i]; jn < offsets[i + 1]; ++jn) {
                int j = nbr[jn];
                if (el[j] < 0) continue;
                double rij[3];
                for (int x = 0; x < 3; ++x) {
                    double as = 0.0;
                    as = as + cell[3 * 0 + x] * (double)dc[3 * jn + 0];
                    as = as + cell[3 * 1 + x] * (double)dc[3 * jn + 1];
                    as = as + cell[3 * 2 + x] * (double)dc[3 * jn + 2];
                    rij[x] = (r[3 * j + x] - r[3 * i + x]) - as;
                }
                double rl = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
                int t = pair_index(el[i], el[j], nel);
                if (rl < p->r2[t] * p->r2[t]) {
                    bond_t *b = &bonds[nb];
                    b->j = j; b->typ = t;
                    b->dc[0] = dc[3 * jn]; b->dc[1] = dc[3 * jn + 1]; b->dc[2] = dc[3 * jn + 2];
                    b->len = sqrt(rl);
                    for (int x = 0; x < 3; ++x) b->nm[x] = rij[x] / b->len;
                    if (rl < p->r1[t] * p->r1[t]) { b->fc = 1.0; b->dfc = 0.0; }
                    else trig_off(p->r1[t], p->r2[t], b->len, &b->fc, &b->dfc);
                    ++nb;
                }
            }
        }
        blast[i] = nb;
        if (nb - bseed[i] > nebmax) nebmax = (int)(nb - bseed[i]);
    }

#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : 1;
#else
    int nt = 1;
    (void)nthreads;
#endif
    double *fbuf = calloc((size_t)nt * nat * 3 + 1, sizeof(double));
    double *pebuf = calloc((size_t)nt * nat + 1, sizeof(double));
    double *wbuf = calloc((size_t)nt * 9, sizeof(double));
    double *wpa = wpot_per_at;

#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *ft = fbuf + (size_t)tid * nat * 3;
        double *pe = pebuf + (size_t)tid * nat;
        double *w = wbuf + 9 * tid;
        double (*dbidk)[3] = malloc(sizeof(double[3]) * (nebmax + 1));
#pragma omp for schedule(static)
        for (int i = 0; i < nat; ++i) {
            int eli = el[i];
            if (eli < 0) continue;
            double fi[3] = {0, 0, 0};
            for (int64_t ij = bseed[i]; ij < blast[i]; ++ij) {
                const bond_t *bnd_ij = &bonds[ij];
                int j = bnd_ij->j;
                int maskfac = 2;
                if (mask) {
                    if (mask[i] == 0 && mask[j] == 0) maskfac = 0;
                    else if (mask[i] == 0 || mask[j] == 0) maskfac = 1;
                }
                int el2ij = bnd_ij->typ;
                double rlij = bnd_ij->len;
                if (!(maskfac > 0 && rlij < p->r2[el2ij])) continue;
                double fj[3] = {0, 0, 0};
                int elj = el[j];
                (void)elj;
                double rlijr = 1.0 / rlij;
                const double *rnij = bnd_ij->nm;
                double rij[3] = {rlij * rnij[0], rlij * rnij[1], rlij * rnij[2]};
                double fcarij = bnd_ij->fc, dfcarijr = bnd_ij->dfc;
                double VAij, dVAij, VRij, dVRij;
                if (p->fam == 2) {
                    double S = p->bS[el2ij];
                    double expR = p->bbeta[el2ij] * sqrt(2 * S), expA = p->bbeta[el2ij] * sqrt(2 / S);
                    double VR_f = p->bD0[el2ij] / (S - 1), VA_f = S * p->bD0[el2ij] / (S - 1);
                    double ex = exp(-expA * (rlij - p->br0[el2ij]));
                    VAij = -VA_f * ex;
                    dVAij = VA_f * expA * ex;
                    ex = exp(-expR * (rlij - p->br0[el2ij]));
                    VRij = VR_f * ex;
                    dVRij = -VR_f * expR * ex;
                } else {
                    double ex = exp(-p->mu[el2ij] * rlij);
                    VAij = -p->B[el2ij] * ex;
                    dVAij = p->B[el2ij] * p->mu[el2ij] * ex;
                    ex = exp(-p->lambda[el2ij] * rlij);
                    VRij = p->A[el2ij] * ex;
                    dVRij = -p->A[el2ij] * p->lambda[el2ij] * ex;
                }
                VAij = 0.5 * maskfac * VAij;
                dVAij = 0.5 * maskfac * dVAij;
                VRij = 0.5 * maskfac * VRij;
                dVRij = 0.5 * maskfac * dVRij;
                double wij[9] = {0}, wijb[9] = {0};
                double zij = 0.0, dbidi[3] = {0, 0, 0}, dbidj[3] = {0, 0, 0};
                int ikc = 0;
                for (int64_t ik = bseed[i]; ik < blast[i]; ++ik, ++ikc) {
                    const bond_t *bik = &bonds[ik];
                    dbidk[ikc][0] = dbidk[ikc][1] = dbidk[ikc][2] = 0.0;
                    if (ik == ij) continue;
                    int ikpot = bik->typ;
                    double rlik = bik->len;
                    if (!(rlik < p->r2[ikpot])) continue;
                    const double *rnik = bik->nm;
                    double rik[3] = {rlik * rnik[0], rlik * rnik[1], rlik * rnik[2]};
                    double fcik = bik->fc, dfcikr = bik->dfc;
                    double h_Dr, dh_dDr;
                    h_fn(p, ikpot, rlij - rlik, &h_Dr, &dh_dDr);
                    double costh = rnik[0] * rnij[0] + rnik[1] * rnij[1] + rnik[2] * rnij[2];
                    double g_costh, dg_dcosth;
                    if (p->fam == 2) {
                        /* brenner_func.f90:77-101, parameters of the pair ik */
                        double c_sq = p->bc[ikpot] * p->bc[ikpot], d_sq = p->bd[ikpot] * p->bd[ikpot];
                        double hh = d_sq + (p->bh[ikpot] + costh) * (p->bh[ikpot] + costh);
                        g_costh = p->bgam[ikpot] * (1 + c_sq / d_sq - c_sq / hh);
                        dg_dcosth = 2 * p->bgam[ikpot] * c_sq * (p->bh[ikpot] + costh) / (hh * hh);
                    } else if (p->fam == 1) {
                        /* kumagai_func.f90:86-131, parameters of the centre atom */
                        double h_cos = p->h[eli] - costh;
                        double h_cos_sq = h_cos * h_cos;
                        double tmp = h_cos / (p->kc3[eli] + h_cos_sq);
                        double go = p->kc2[eli] * tmp;
                        double ga1 = p->kc4[eli] * exp(-p->kc5[eli] * h_cos_sq);
                        double v = go * (1.0 + ga1);
                        dg_dcosth = -2 * (1.0 - h_cos * tmp) * v + 2 * p->kc5[eli] * h_cos_sq * go * ga1;
                        g_costh = p->kc1[eli] + h_cos * v;
                    } else {
                        /* g(theta): c, d, h of the centre atom, omega of ik */
                        double omega = p->omega[ikpot];
                        double h_c = p->h[eli] - costh;
                        double c_sq = p->c[eli] * p->c[eli];
                        double d_sq = p->d[eli] * p->d[eli];
                        double hh = d_sq + h_c * h_c;
                        g_costh = omega * (1.0 + c_sq / d_sq - c_sq / hh);
                        dg_dcosth = -2 * omega * c_sq * h_c / (hh * hh);
                    }
                    double dkc[3];
                    for (int x = 0; x < 3; ++x) dkc[x] = rnik[x] * rlik - rnij[x] * rlij;
                    double disjk = sqrt(dkc[0] * dkc[0] + dkc[1] * dkc[1] + dkc[2] * dkc[2]);
                    for (int x = 0; x < 3; ++x) dkc[x] = dkc[x] / disjk;
                    double dcsdij = 1.0 / rlik - costh * rlijr;
                    double dcsdik = rlijr - costh / rlik;
                    double dcsdjk = -disjk * rlijr / rlik;
                    double dzfac = fcik * dg_dcosth * h_Dr;
                    zij = zij + fcik * g_costh * h_Dr;
                    double dzdrij = g_costh * fcik * dh_dDr;
                    double dzdrik = g_costh * (dfcikr * h_Dr - fcik * dh_dDr);
                    double df[3];
                    for (int x = 0; x < 3; ++x) {
                        double dcsdi = -dcsdij * rnij[x] - dcsdik * rnik[x];
                        double dcsdj = dcsdij * rnij[x] - dcsdjk * dkc[x];
                        double dcsdk = dcsdik * rnik[x] + dcsdjk * dkc[x];
                        double dgdi = dzfac * dcsdi, dgdj = dzfac * dcsdj, dgdk = dzfac * dcsdk;
                        dbidi[x] = dbidi[x] - dzdrij * rnij[x] - dzdrik * rnik[x] + dgdi;
                        df[x] = dzdrij * rnij[x] + dgdj;
                        dbidj[x] = dbidj[x] + df[x];
                        dbidk[ikc][x] = dzdrik * rnik[x] + dgdk;
                    }
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            wijb[a + 3 * b] = wijb[a + 3 * b] - rij[a] * df[b] - rik[a] * dbidk[ikc][b];
                }
                int numnbi = ikc;
                /* bo: b_ij and dE/dzeta */
                double bij = 1.0, dbij_dzij = 0.0;
                if (p->fam == 2) {
                    /* brenner_func.f90:107-147 */
                    double n = p->bn[el2ij];
                    double bo_exp = -0.5 / n, bo_fac = 0.5 * bo_exp * n, bo_exp1 = bo_exp - 1.0;
                    if (n == 1.0) {
                        double arg = 1.0 + zij;
                        bij = pow(arg, bo_exp);
                        dbij_dzij = bo_fac * fcarij * VAij * pow(arg, bo_exp1);
                    } else if (zij > 0.0) {
                        double arg = 1.0 + pow(zij, n);
                        bij = pow(arg, bo_exp);
                        dbij_dzij = bo_fac * fcarij * VAij * pow(zij, n - 1.0) * pow(arg, bo_exp1);
                    }
                } else if (zij > 0.0 && p->fam == 1) {
                    /* kumagai_func.f90:139-171 */
                    double eta = p->keta[eli], delta = -p->kdelta[eli];
                    double arg = 1.0 + pow(zij, eta);
                    bij = pow(arg, delta);
                    dbij_dzij = 0.5 * fcarij * VAij * eta * pow(zij, eta - 1.0) * delta * pow(arg, delta - 1.0);
                } else if (zij > 0.0) {
                    double nn = p->n[eli];
                    double e = -0.5 / nn;
                    double b = pow(p->beta[eli], nn);
                    double arg = 1.0 + b * pow(zij, nn);
                    bij = p->xi[el2ij] * pow(arg, e);
                    dbij_dzij = -0.25 * fcarij * VAij * p->xi[el2ij] * b *
                                pow(zij, nn - 1.0) * pow(arg, e - 1.0);
                }
                double dffac = 0.5 * fcarij * (VRij + bij * VAij);
                pe[i] += dffac;
                pe[j] += dffac;
                dffac = 0.5 * (dVRij * fcarij + bij * dVAij * fcarij + VRij * dfcarijr +
                               bij * VAij * dfcarijr);
                double df[3];
                for (int x = 0; x < 3; ++x) {
                    df[x] = dffac * rnij[x];
                    fi[x] += df[x];
                    fj[x] -= df[x];
                }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        wij[a + 3 * b] = wij[a + 3 * b] + rij[a] * df[b] - dbij_dzij * wijb[a + 3 * b];
                for (int x = 0; x < 3; ++x) {
                    fi[x] += -dbij_dzij * dbidi[x];
                    fj[x] += -dbij_dzij * dbidj[x];
                }
                ikc = 0;
                for (int64_t ik = bseed[i]; ik < blast[i] && ikc < numnbi; ++ik, ++ikc) {
                    const bond_t *bik = &bonds[ik];
                    int k = bik->j;
                    if (k != j || bik->dc[0] != bnd_ij->dc[0] || bik->dc[1] != bnd_ij->dc[1] ||
                        bik->dc[2] != bnd_ij->dc[2])
                        for (int x = 0; x < 3; ++x) ft[3 * k + x] += -dbij_dzij * dbidk[ikc][x];
                }
                for (int a = 0; a < 9; ++a) w[a] += wij[a];
                if (wpa) {
#pragma omp critical
                    for (int a = 0; a < 9; ++a) {
                        wpa[9 * i + a] += wij[a] / 2;
                        wpa[9 * j + a] += wij[a] / 2;
                    }
                }
                for (int x = 0; x < 3; ++x) ft[3 * j + x] += fj[x];
            }
            for (int x = 0; x < 3; ++x) ft[3 * i + x] += fi[x];
        }
        free(dbidk);
    }
    /* tls_reduce model: thread buffers summed in thread order */
    double esum = 0.0;
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < nat; ++i) esum += pebuf[(size_t)t * nat + i];
    *epot += 0.5 * esum;
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 3 * nat; ++k) f[k] += fbuf[(size_t)t * nat * 3 + k];
        if (epot_per_at)
            for (int i = 0; i < nat; ++i) epot_per_at[i] += 0.5 * pebuf[(size_t)t * nat + i];
        for (int a = 0; a < 9; ++a) wpot[a] += wbuf[9 * t + a];
    }
    free(fbuf); free(pebuf); free(wbuf);
    free(bonds); free(bseed); free(blast);
    return ORC_OK;
}
