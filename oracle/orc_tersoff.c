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
 *   - fam = 3: VA, VR, g, bo, h of src/potentials/bop/juslin/juslin_func.f90:69-313
 *     with the constants of BIND_TO_FUNC (juslin_module.f90:263-330):
 *     ordered pair index PAIR_INDEX_NS, triplet index TRIPLET_INDEX_NS
 *     (src/macros.inc:139-146)
 *   - PAIR_INDEX          src/macros.inc:123
 *   - element map / nebmax src/potentials/bop/default_compute_func.f90:62-73
 *   - scr = 1 (TersoffScr, KumagaiScr, BrennerScr): loop1's screening function
 *     and its cutoff regions (bop_kernel.f90:563-1066, exp_cutoff of
 *     src/support/cutoff.f90:232-293, C_dr_cut and the requested range of
 *     default_bind_to_func.f90:44-150), the forces of the pair-term screening
 *     in loop2 (:1448-1492) and of the bond-order screening in the last loop
 *     (:1531-1609)
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
    double len, nm[3], fc, dfc;     /* fc: the pair (ar) cutoff */
    double fcbo, dfcbo;             /* bond-order cutoff (= fc unscreened) */
    int64_t s0, s1;                 /* screening records [s0, s1) */
    int64_t jn;                     /* list entry of the bond (nbb, bop_kernel.f90:1030) */
} bond_t;

/* one third atom k screening bond ij: its list entry and the derivative
 * factors of the pair (ar) and bond-order (bo) cutoffs along r_ik and r_jk */
typedef struct {
    int64_t kn;
    double arik, arjk, boik, bojk;
} scr_t;

/* exp_cutoff_f (cutoff.f90:245-293): exp(-8 x^3) corrected so that value,
 * first and second derivative vanish at x = 1 */
static void exp_cut(double r1, double r2, double r, double *val, double *dval)
{
    if (r <= r1) { *val = 1.0; *dval = 0.0; return; }
    if (r >= r2) { *val = 0.0; *dval = 0.0; return; }
    const double v1 = exp(-8.0), dv1 = -24 * v1, ddv1 = -48 * v1 - 24 * dv1;
    const double c = (-3 * dv1 + ddv1) / 3, d = (2 * dv1 - ddv1) / 4;
    const double fac1 = 1.0 / (r2 - r1), fac2 = 1.0 / (1 - v1 - c - d), off = v1 + c + d;
    double x = fac1 * (r - r1), x2 = x * x;
    double v = exp(-8 * x * x2);
    double dv = -24 * x2 * v;
    *dval = fac1 * fac2 * (dv + 3 * c * x2 + 4 * d * x * x2);
    *val = fac2 * (v + c * x * x2 + d * x2 * x2 - off);
}

static void entry_vec(const double *r, const double *cell, const int *dc, int i, int j, int64_t jn,
                      double *rij)
{
    for (int x = 0; x < 3; ++x) {
        double as = 0.0;
        as = as + cell[3 * 0 + x] * (double)dc[3 * jn + 0];
        as = as + cell[3 * 1 + x] * (double)dc[3 * jn + 1];
        as = as + cell[3 * 2 + x] * (double)dc[3 * jn + 2];
        rij[x] = (r[3 * j + x] - r[3 * i + x]) - as;
    }
}

/* loop1 of the screened kernels for the list entry jn of atom i (r^2 = rl):
 * returns 1 and fills *b (and appends records) when the bond is kept */
static int screened_bond(const orc_tersoff_params *p, int i, int64_t jn, const double *rij, double rl,
                         int t, const double *r, const double *cell, const int64_t *offsets,
                         const int *nbr, const int *dc, bond_t *b, scr_t **rec, int64_t *nrec,
                         int64_t *cap)
{
    const double cin_l = p->r1[t], cin_h = p->r2[t];
    double mx = cin_h;
    if (p->or2[t] > mx) mx = p->or2[t];
    if (p->bor2[t] > mx) mx = p->bor2[t];
    const double max_cut_sq = mx * mx;
    const double Cmin = p->Cmin[t], Cmax = p->Cmax[t], dC = Cmax - Cmin;
    const double C_dr_cut = Cmax > 2.0 ? Cmax * Cmax / (4 * (Cmax - 1)) : 1.0;
    const double thr = log(1e-6), dot_thr = 1e-10;
    b->s0 = b->s1 = *nrec;
    if (rl < cin_l * cin_l) {
        /* region (a): no screening */
        b->fc = 1.0; b->dfc = 0.0; b->fcbo = 1.0; b->dfcbo = 0.0;
        b->len = sqrt(rl);
        return 1;
    }
    if (rl < max_cut_sq && p->or1[t] < p->or2[t]) {
        int screened = 0, need = 0;
        double sij = 0.0, dsijdrij = 0.0;
        const int64_t ineb = *nrec;
        for (int64_t kn = offsets[i]; !(screened || sij < thr) && kn < offsets[i + 1]; ++kn) {
            const int k = nbr[kn];
            double rik[3];
            entry_vec(r, cell, dc, i, k, kn, rik);
            if (!(rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2] < C_dr_cut * rl)) continue;
            if (k == nbr[jn] && dc[3 * kn] == dc[3 * jn] && dc[3 * kn + 1] == dc[3 * jn + 1] &&
                dc[3 * kn + 2] == dc[3 * jn + 2])
                continue;
            const double dot_ij_ik = rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2];
            const double rlik = rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2];
            double rjk[3];
            for (int x = 0; x < 3; ++x) rjk[x] = -rij[x] + rik[x];
            const double dot_ij_jk = rij[0] * rjk[0] + rij[1] * rjk[1] + rij[2] * rjk[2];
            const double rljk = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
            if (!(dot_ij_ik > dot_thr && dot_ij_jk < -dot_thr)) continue;
            const double xik = rlik / rl, xjk = rljk / rl;
            const double xm = xik - xjk, xp = xik + xjk;
            double fac = 1.0 / (1 - xm * xm);
            const double C = (2 * xp - xm * xm - 1) * fac;
            if (C <= Cmin) {
                screened = 1;
            } else if (C < Cmax) {
                need = 1;
                const double Cmax_C = Cmax - C, C_Cmin = C - Cmin;
                sij = sij - (Cmax_C / C_Cmin) * (Cmax_C / C_Cmin);
                const double dCdrik = 4 * xik * fac * (1 + (C - 1) * xm);
                const double dCdrjk = 4 * xjk * fac * (1 - (C - 1) * xm);
                const double dCdrij = -(dCdrik + dCdrjk);
                fac = 2 * Cmax_C * dC / (C_Cmin * C_Cmin * C_Cmin);
                dsijdrij = dsijdrij + fac * dCdrij;
                if (*nrec >= *cap) {
                    *cap = *cap * 2 + 64;
                    *rec = realloc(*rec, sizeof(scr_t) * (size_t)*cap);
                }
                scr_t *q = &(*rec)[(*nrec)++];
                q->kn = kn;
                q->arik = fac * dCdrik / rlik;
                q->arjk = fac * dCdrjk / rljk;
            }
        }
        if ((screened || sij < thr) && rl > cin_h * cin_h) {
            *nrec = ineb;           /* fully screened: dropped */
            return 0;
        }
        const double len = sqrt(rl);
        b->len = len;
        double fcin, dfcin;
        exp_cut(cin_l, cin_h, len, &fcin, &dfcin);
        if (screened) {
            b->fc = fcin; b->dfc = dfcin; b->fcbo = fcin; b->dfcbo = dfcin;
            *nrec = ineb;
        } else if (need) {
            sij = exp(sij);
            double fcar, dfcar, fcbo, dfcbo;
            exp_cut(p->or1[t], p->or2[t], len, &fcar, &dfcar);
            exp_cut(p->bor1[t], p->bor2[t], len, &fcbo, &dfcbo);
            b->fc = (1.0 - fcin) * sij * fcar + fcin;
            b->dfc = (1.0 - fcin) * sij * (dfcar + fcar * dsijdrij / len) - dfcin * sij * fcar + dfcin;
            b->fcbo = (1.0 - fcin) * sij * fcbo + fcin;
            b->dfcbo = (1.0 - fcin) * sij * (dfcbo + fcbo * dsijdrij / len) - dfcin * sij * fcbo + dfcin;
            for (int64_t q = ineb; q < *nrec; ++q) {
                scr_t *e = &(*rec)[q];
                e->boik = e->arik * sij * fcbo * (1.0 - fcin);
                e->bojk = e->arjk * sij * fcbo * (1.0 - fcin);
                e->arik = e->arik * sij * fcar * (1.0 - fcin);
                e->arjk = e->arjk * sij * fcar * (1.0 - fcin);
            }
        } else {
            double fcar, dfcar, fcbo, dfcbo;
            exp_cut(p->or1[t], p->or2[t], len, &fcar, &dfcar);
            exp_cut(p->bor1[t], p->bor2[t], len, &fcbo, &dfcbo);
            if (len < cin_h) {
                b->fc = (1.0 - fcin) * fcar + fcin;
                b->dfc = (1.0 - fcin) * dfcar - dfcin * fcar + dfcin;
                b->fcbo = (1.0 - fcin) * fcbo + fcin;
                b->dfcbo = (1.0 - fcin) * dfcbo - dfcin * fcbo + dfcin;
            } else {
                b->fc = fcar; b->dfc = dfcar; b->fcbo = fcbo; b->dfcbo = dfcbo;
            }
        }
        b->s0 = ineb;
        b->s1 = *nrec;
        return 1;
    }
    if (rl < cin_h * cin_h) {
        /* beyond the screening range: inner cutoff only */
        b->len = sqrt(rl);
        exp_cut(cin_l, cin_h, b->len, &b->fc, &b->dfc);
        b->fcbo = b->fc;
        b->dfcbo = b->dfc;
        return 1;
    }
    return 0;
}

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

/* element pair of the family: ordered (Juslin) or symmetric */
static int bop_pair(const orc_tersoff_params *p, int i, int j)
{
    return p->fam == 3 ? j + i * p->nel : pair_index(i, j, p->nel);
}

static void h_fn(const orc_tersoff_params *p, int ikpot, double dr, double *val, double *dval, int eli, int elj,
                 int elk)
{
    if (p->fam == 3) {
        /* juslin_func.f90:272-298, h(ktypj, ktypi, ktypk) of centre i */
        const int t = elk + p->nel * (elj + p->nel * eli);
        const double alpha = p->jal[t], omega = p->jom[t];
        const int m = p->jm[t];
        if (m == 1) {
            *val = omega * exp(alpha * dr);
            *dval = alpha * *val;
        } else if (m == 3) {
            const double arg = alpha * dr;
            *val = omega * exp(arg * arg * arg);
            *dval = 3 * alpha * arg * arg * *val;
        } else {
            const double arg = alpha * dr;
            *val = omega * exp(pow(arg, m));
            *dval = m * pow(arg, m - 1) * alpha * *val;
        }
        return;
    }
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
                double *wpot_per_at, double *epot_per_bond, double *f_per_bond,
                double *wpot_per_bond, int nthreads)
{
    const int nel = p->nel;
    /* per-bond arrays are indexed by the bond's list entry; every directed
     * bond belongs to one centre, hence to one thread */
    double *epb = epot_per_bond, *fpb = f_per_bond, *wpb = wpot_per_bond;
    /* loop1: internal bond lists */
    bond_t *bonds = malloc(sizeof(bond_t) * (offsets[nat] > 0 ? offsets[nat] : 1));
    int64_t *bseed = malloc(sizeof(int64_t) * (nat + 1));
    int64_t *blast = malloc(sizeof(int64_t) * (nat + 1));
    int nebmax = 0;
    scr_t *rec = NULL;
    int64_t nrec = 0, rcap = 0;
    for (int i = 0; i < nat; ++i) {
        bseed[i] = offsets[i];
        int64_t nb = offsets[i];
        if (el[i] >= 0) {
            for (int64_t jn = offsets[i]; jn < offsets[i + 1]; ++jn) {
                int j = nbr[jn];
                if (el[j] < 0) continue;
                double rij[3];
                entry_vec(r, cell, dc, i, j, jn, rij);
                double rl = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
                int t = bop_pair(p, el[i], el[j]);
                if (p->scr) {
                    bond_t *b = &bonds[nb];
                    if (screened_bond(p, i, jn, rij, rl, t, r, cell, offsets, nbr, dc, b, &rec, &nrec,
                                      &rcap)) {
                        b->j = j; b->typ = t; b->jn = jn;
                        b->dc[0] = dc[3 * jn]; b->dc[1] = dc[3 * jn + 1]; b->dc[2] = dc[3 * jn + 2];
                        for (int x = 0; x < 3; ++x) b->nm[x] = rij[x] / b->len;
                        ++nb;
                    }
                } else if (rl < p->r2[t] * p->r2[t]) {
                    bond_t *b = &bonds[nb];
                    b->j = j; b->typ = t; b->jn = jn;
                    b->dc[0] = dc[3 * jn]; b->dc[1] = dc[3 * jn + 1]; b->dc[2] = dc[3 * jn + 2];
                    b->len = sqrt(rl);
                    for (int x = 0; x < 3; ++x) b->nm[x] = rij[x] / b->len;
                    if (rl < p->r1[t] * p->r1[t]) { b->fc = 1.0; b->dfc = 0.0; }
                    else trig_off(p->r1[t], p->r2[t], b->len, &b->fc, &b->dfc);
                    b->fcbo = b->fc;
                    b->dfcbo = b->dfc;
                    b->s0 = b->s1 = 0;
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
    /* bond-order screening factor per bond (sfacbo of bop_kernel.f90:1456) */
    double *sfac = calloc((size_t)(offsets[nat] > 0 ? offsets[nat] : 1), sizeof(double));
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
        double *zfac = malloc(sizeof(double) * (nebmax + 1));
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
                /* cut_ar_h is the outer cutoff or2 when screened */
                if (!(maskfac > 0 && rlij < (p->scr ? p->or2[el2ij] : p->r2[el2ij]))) continue;
                double fj[3] = {0, 0, 0};
                int elj = el[j];
                (void)elj;
                double rlijr = 1.0 / rlij;
                const double *rnij = bnd_ij->nm;
                double rij[3] = {rlij * rnij[0], rlij * rnij[1], rlij * rnij[2]};
                double fcarij = bnd_ij->fc, dfcarijr = bnd_ij->dfc;
                double VAij, dVAij, VRij, dVRij;
                if (p->fam >= 2) {
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
                    zfac[ikc] = 0.0;
                    int ikpot = bik->typ;
                    double rlik = bik->len;
                    if (!(rlik < (p->scr ? p->bor2[ikpot] : p->r2[ikpot]))) continue;
                    const double *rnik = bik->nm;
                    double rik[3] = {rlik * rnik[0], rlik * rnik[1], rlik * rnik[2]};
                    double fcik = bik->fcbo, dfcikr = bik->dfcbo;
                    double h_Dr, dh_dDr;
                    h_fn(p, ikpot, rlij - rlik, &h_Dr, &dh_dDr, eli, el[j], el[bik->j]);
                    double costh = rnik[0] * rnij[0] + rnik[1] * rnij[1] + rnik[2] * rnij[2];
                    double g_costh, dg_dcosth;
                    if (p->fam >= 2) {
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
                    zfac[ikc] = g_costh * h_Dr;
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
                if (p->fam >= 2) {
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
                /* bop_kernel.f90:1370-1373 */
                if (epb) epb[bnd_ij->jn] += dffac;
                dffac = 0.5 * (dVRij * fcarij + bij * dVAij * fcarij + VRij * dfcarijr +
                               bij * VAij * dfcarijr);
                double df[3];
                for (int x = 0; x < 3; ++x) {
                    df[x] = dffac * rnij[x];
                    fi[x] += df[x];
                    fj[x] -= df[x];
                }
                /* bop_kernel.f90:1406-1408: the pair term only */
                if (fpb)
                    for (int x = 0; x < 3; ++x) fpb[3 * bnd_ij->jn + x] += df[x];
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
                        bik->dc[2] != bnd_ij->dc[2]) {
                        for (int x = 0; x < 3; ++x) ft[3 * k + x] += -dbij_dzij * dbidk[ikc][x];
                        if (p->scr) sfac[ik] += zfac[ikc] * dbij_dzij;
                    }
                }
                if (p->scr) {
                    /* forces of the pair-term screening (bop_kernel.f90:1469-1490) */
                    double dff = 0.5 * (VRij + bij * VAij);
                    for (int64_t q = bnd_ij->s0; q < bnd_ij->s1; ++q) {
                        const scr_t *e = &rec[q];
                        int k = nbr[e->kn];
                        double rik[3], rjk[3], d1[3], d2[3];
                        entry_vec(r, cell, dc, i, k, e->kn, rik);
                        for (int x = 0; x < 3; ++x) rjk[x] = -rij[x] + rik[x];
                        for (int x = 0; x < 3; ++x) {
                            d1[x] = dff * e->arik * rik[x];
                            fi[x] += d1[x];
                            ft[3 * k + x] += -d1[x];
                        }
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b) wij[a + 3 * b] += rik[a] * d1[b];
                        for (int x = 0; x < 3; ++x) {
                            d2[x] = dff * e->arjk * rjk[x];
                            fj[x] += d2[x];
                            ft[3 * k + x] += -d2[x];
                        }
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b) wij[a + 3 * b] += rjk[a] * d2[b];
                    }
                }
                for (int a = 0; a < 9; ++a) w[a] += wij[a];
                /* bop_kernel.f90:1511-1512 */
                if (wpb)
                    for (int a = 0; a < 9; ++a) wpb[9 * bnd_ij->jn + a] += wij[a];
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
        if (p->scr) {
            /* forces of the bond-order screening (bop_kernel.f90:1531-1609);
             * the omp for above ends with a barrier, so sfac is complete */
#pragma omp for schedule(static)
            for (int i = 0; i < nat; ++i) {
                if (el[i] < 0) continue;
                double fi[3] = {0, 0, 0};
                for (int64_t ij = bseed[i]; ij < blast[i]; ++ij) {
                    const bond_t *b = &bonds[ij];
                    int j = b->j;
                    double fj[3] = {0, 0, 0}, wij[9] = {0};
                    double rij[3] = {b->len * b->nm[0], b->len * b->nm[1], b->len * b->nm[2]};
                    for (int64_t q = b->s0; q < b->s1; ++q) {
                        const scr_t *e = &rec[q];
                        double cik = sfac[ij] * e->boik, cjk = sfac[ij] * e->bojk;
                        int k = nbr[e->kn];
                        double rik[3], rjk[3], d1[3], d2[3];
                        entry_vec(r, cell, dc, i, k, e->kn, rik);
                        for (int x = 0; x < 3; ++x) rjk[x] = -rij[x] + rik[x];
                        for (int x = 0; x < 3; ++x) {
                            d1[x] = cik * rik[x];
                            fi[x] += d1[x];
                            ft[3 * k + x] += -d1[x];
                        }
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) wij[a + 3 * c] += rik[a] * d1[c];
                        for (int x = 0; x < 3; ++x) {
                            d2[x] = cjk * rjk[x];
                            fj[x] += d2[x];
                            ft[3 * k + x] += -d2[x];
                        }
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) wij[a + 3 * c] += rjk[a] * d2[c];
                    }
                    for (int a = 0; a < 9; ++a) w[a] += wij[a];
                    /* bop_kernel.f90:1593-1594 */
                    if (wpb)
                        for (int a = 0; a < 9; ++a) wpb[9 * b->jn + a] += wij[a];
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
        }
        free(dbidk);
        free(zfac);
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
    free(sfac); free(rec);
    return ORC_OK;
}
