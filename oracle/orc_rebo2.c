/*
 * Oracle: second-generation REBO (Brenner et al. 2002) for C and H, without
 * screening; the dihedral (torsion) term is optional as in the reference
 * (with_dihedral, default .false.).  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates:
 *   - parameters and constants      src/potentials/bop/rebo2/rebo2_type.f90:73-200
 *   - bond-order constants, h prefactors, cut-offs, table set-up
 *                                   src/potentials/bop/rebo2/rebo2_db.f90:81-303
 *   - g(cos) quintic splines for C  src/potentials/bop/rebo2/rebo2_db.f90:405-524
 *   - default F, P, T tables        src/potentials/bop/rebo2/rebo2_default_tables.f90
 *   - bicubic / tricubic tables     src/special/table2d.f90, src/special/table3d.f90
 *   - fconj, fCin, VA, VR, g, bo, h src/potentials/bop/rebo2/rebo2_func.f90:31-485
 *   - trig_off cut-off              src/support/cutoff.f90:152-196
 *   - the kernel (bond list, coordination, conjugation, half-pair loop with
 *     bond orders, P, F, forces on i, j, k, l and second neighbours m, n)
 *                                   src/potentials/bop/rebo2/bop_kernel_rebo2.f90:700-2881
 *
 * The tables are set up the way the reference does it: one dense linear
 * system per table (value, first-derivative and cross-derivative rows at the
 * box corners), solved here by Gaussian elimination with partial pivoting.
 *
 * Internal element codes follow the reference: C = 1, H = 3; pair codes are
 * CC = 1, CH = 3, HH = 6 (Z2pair).
 *
 * Virial layout: wpot[a + 3*b] = W(a, b), wpot_per_at[9*i + a + 3*b].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc.h"

#define PI_ 3.14159265358979323846
#define KC 1
#define KH 3
#define PCC 1
#define PCH 3
#define PHH 6

/* default-real (single precision) literals of the reference, widened */
#define SP(x) ((double)(float)(x))

struct orc_rebo2_s {
    /* pair functions */
    double cc_B1, cc_B2, cc_B3, cc_beta1, cc_beta2, cc_beta3, cc_Q, cc_A, cc_alpha;
    double ch_B1, ch_beta1, ch_Q, ch_A, ch_alpha;
    double hh_B1, hh_beta1, hh_Q, hh_A, hh_alpha;
    double cc_g_theta[6], cc_g_g1[6], cc_g_dg1[6], cc_g_d2g1[6], cc_g_g2[6];
    double spgh[3][6];
    int igh[25];
    double hhh_lambda, cc_re, ch_re, hh_re;
    double cc_in_r1, cc_in_r2, ch_r1, ch_r2, hh_r1, hh_r2;
    /* derived */
    double conpe[4], conan[4], conpf[4], conalp, conear[7][7];
    double cut_l[7], cut_h[7], cut_h2[7];
    double g1c[3][6], g2c[3][6];
    /* tables: coefficients c[box][p1][p2][p3] (3d) and c[box][p1][p2] (2d) */
    double Fcc[4 * 4 * 9][64], Fch[4 * 4 * 9][64], Fhh[4 * 4 * 9][64];
    double Pcc[5 * 5][16], Pch[5 * 5][16];
    double Tcc[4 * 4 * 9][64];
    int with_dihedral;
};

/* ------------------------------------------------------------------------ */
/* dense solve with partial pivoting: A (n x n, row-major) X = B (n x m)     */
/* ------------------------------------------------------------------------ */
static int solve_multi(int n, double *A, int m, double *B)
{
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r * n + c]) > fabs(A[piv * n + c])) piv = r;
        if (A[piv * n + c] == 0.0) return ORC_ERR_SINGULAR;
        if (piv != c) {
            for (int k = 0; k < n; ++k) {
                double t = A[c * n + k]; A[c * n + k] = A[piv * n + k]; A[piv * n + k] = t;
            }
            for (int k = 0; k < m; ++k) {
                double t = B[c * m + k]; B[c * m + k] = B[piv * m + k]; B[piv * m + k] = t;
            }
        }
        for (int r = c + 1; r < n; ++r) {
            double fac = A[r * n + c] / A[c * n + c];
            if (fac == 0.0) continue;
            for (int k = c; k < n; ++k) A[r * n + k] -= fac * A[c * n + k];
            for (int k = 0; k < m; ++k) B[r * m + k] -= fac * B[c * m + k];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        for (int k = 0; k < m; ++k) {
            double s = B[c * m + k];
            for (int j = c + 1; j < n; ++j) s -= A[c * n + j] * B[j * m + k];
            B[c * m + k] = s / A[c * n + c];
        }
    }
    return ORC_OK;
}

static double ipow(double x, int p)
{
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= x;
    return r;
}

/* d^a/dx^a of x^p at x (integer corner coordinates 0/1) */
static double dmono(int x, int p, int a)
{
    if (a == 0) return ipow((double)x, p);
    if (p == 0) return 0.0;
    return (double)p * ipow((double)x, p - 1);
}

/* table2d_init (src/special/table2d.f90): values V[(nx+1)*(ny+1)] indexed
 * V[ix*(ny+1)+iy]; no derivative data (the reference passes none for P).
 * Box order ibox = ny*ix + iy. */
static int table2d_init(int nx, int ny, const double *V, double (*coef)[16])
{
    static const int c1[4] = {0, 1, 1, 0}, c2[4] = {0, 0, 1, 1};
    double A[16 * 16];
    int nb = nx * ny;
    double *B = calloc((size_t)16 * nb, sizeof(double));
    if (!B) return ORC_ERR_ARG;
    for (int corner = 0; corner < 4; ++corner)
        for (int p1 = 0; p1 < 4; ++p1)
            for (int p2 = 0; p2 < 4; ++p2) {
                int col = 4 * p1 + p2;
                for (int d = 0; d < 4; ++d) {
                    int a1 = d & 1, a2 = (d >> 1) & 1;
                    A[(corner + 4 * d) * 16 + col] = dmono(c1[corner], p1, a1) * dmono(c2[corner], p2, a2);
                }
            }
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            int box = ny * ix + iy;
            for (int corner = 0; corner < 4; ++corner)
                B[corner * nb + box] = V[(ix + c1[corner]) * (ny + 1) + iy + c2[corner]];
        }
    int err = solve_multi(16, A, nb, B);
    for (int box = 0; box < nb && !err; ++box)
        for (int k = 0; k < 16; ++k) coef[box][k] = B[k * nb + box];
    free(B);
    return err;
}

/* table3d_init (src/special/table3d.f90): arrays indexed
 * V[(ix*(ny+1)+iy)*(nz+1)+iz]; dvdx/dvdy/dvdz may be NULL.
 * Box order ibox = nx*(ny*iz + iy) + ix. */
static int table3d_init(int nx, int ny, int nz, const double *V, const double *Dx,
                        const double *Dy, const double *Dz, double (*coef)[64])
{
    static const int c1[8] = {0, 1, 1, 0, 0, 1, 1, 0}, c2[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                     c3[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    /* derivative rows in the reference's order: f, x, y, z, xy, xz, yz, xyz */
    static const int dd[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                 {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    double *A = malloc(sizeof(double) * 64 * 64);
    int nb = nx * ny * nz;
    double *B = calloc((size_t)64 * nb, sizeof(double));
    if (!A || !B) { free(A); free(B); return ORC_ERR_ARG; }
    for (int corner = 0; corner < 8; ++corner)
        for (int p1 = 0; p1 < 4; ++p1)
            for (int p2 = 0; p2 < 4; ++p2)
                for (int p3 = 0; p3 < 4; ++p3) {
                    int col = 16 * p1 + 4 * p2 + p3;
                    for (int d = 0; d < 8; ++d)
                        A[(corner + 8 * d) * 64 + col] = dmono(c1[corner], p1, dd[d][0]) *
                                                         dmono(c2[corner], p2, dd[d][1]) *
                                                         dmono(c3[corner], p3, dd[d][2]);
                }
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                int box = nx * (ny * iz + iy) + ix;
                for (int corner = 0; corner < 8; ++corner) {
                    int g = ((ix + c1[corner]) * (ny + 1) + iy + c2[corner]) * (nz + 1) + iz + c3[corner];
                    B[corner * nb + box] = V[g];
                    if (Dx) B[(corner + 8) * nb + box] = Dx[g];
                    if (Dy) B[(corner + 16) * nb + box] = Dy[g];
                    if (Dz) B[(corner + 24) * nb + box] = Dz[g];
                }
            }
    int err = solve_multi(64, A, nb, B);
    for (int box = 0; box < nb && !err; ++box)
        for (int k = 0; k < 64; ++k) coef[box][k] = B[k * nb + box];
    free(A);
    free(B);
    return err;
}

static int clampbox(double v, int n)
{
    int b = (int)v;           /* Fortran int(): truncation toward zero */
    if (b < 0) b = 0;
    if (b >= n) b = n - 1;
    return b;
}

/* table2d_eval: nested Horner, x1 outer */
static void table2d_eval(const double (*coef)[16], int nx, int ny, double x, double y,
                         double *val, double *dx, double *dy)
{
    int bx = clampbox(x, nx), by = clampbox(y, ny);
    const double *c = coef[ny * bx + by];
    double x1 = x - bx, x2 = y - by;
    double h = 0.0, dh = 0.0, dc = 0.0;
    for (int i = 3; i >= 0; --i) {
        double s = 0.0, sd = 0.0;
        for (int j = 3; j >= 0; --j) {
            double cij = c[4 * i + j];
            s = s * x2 + cij;
            if (j > 0) sd = sd * x2 + j * cij;
        }
        h = h * x1 + s;
        if (i > 0) dh = dh * x1 + i * s;
        dc = dc * x1 + sd;
    }
    *val = h; *dx = dh; *dy = dc;
}

static void table3d_eval(const double (*coef)[64], int nx, int ny, int nz, double x, double y,
                         double z, double *val, double *dx, double *dy, double *dz)
{
    int bx = clampbox(x, nx), by = clampbox(y, ny), bz = clampbox(z, nz);
    const double *c = coef[nx * (ny * bz + by) + bx];
    double x1 = x - bx, x2 = y - by, x3 = z - bz;
    double f = 0.0, fi = 0.0, fj = 0.0, fk = 0.0;
    for (int i = 3; i >= 0; --i) {
        double s = 0.0, sj = 0.0, sk = 0.0;
        for (int j = 3; j >= 0; --j) {
            double t = 0.0, tk = 0.0;
            for (int k = 3; k >= 0; --k) {
                double cijk = c[16 * i + 4 * j + k];
                t = t * x3 + cijk;
                if (k > 0) tk = tk * x3 + k * cijk;
            }
            s = s * x2 + t;
            if (j > 0) sj = sj * x2 + j * t;
            sk = sk * x2 + tk;
        }
        f = f * x1 + s;
        if (i > 0) fi = fi * x1 + i * s;
        fj = fj * x1 + sj;
        fk = fk * x1 + sk;
    }
    *val = f; *dx = fi; *dy = fj; *dz = fk;
}

/* ------------------------------------------------------------------------ */
/* default tables, src/potentials/bop/rebo2/rebo2_default_tables.f90         */
/* ------------------------------------------------------------------------ */
#define I3(i, j, k) (((i) * 5 + (j)) * 10 + (k))
#define I2(i, j) ((i) * 6 + (j))

static void default_Fcc(double *F, double *Fi, double *Fj, double *Fk)
{
    memset(F, 0, sizeof(double) * 250); memset(Fi, 0, sizeof(double) * 250);
    memset(Fj, 0, sizeof(double) * 250); memset(Fk, 0, sizeof(double) * 250);
    F[I3(1, 1, 0)] = 0.105000;
    F[I3(1, 1, 1)] = -0.0041775;
    for (int k = 2; k <= 8; ++k) F[I3(1, 1, k)] = -0.0160856;
    F[I3(2, 2, 0)] = 0.09444957;
    F[I3(2, 2, 1)] = 0.02200000;
    F[I3(2, 2, 2)] = 0.03970587;
    F[I3(2, 2, 8)] = 0.0;
    F[I3(0, 1, 0)] = 0.04338699;
    for (int k = 1; k <= 8; ++k) F[I3(0, 1, k)] = 0.0099172158;
    F[I3(0, 2, 0)] = 0.0493976637;
    F[I3(0, 2, 1)] = -0.011942669;
    for (int k = 2; k <= 8; ++k) F[I3(0, 2, k)] = F[I3(0, 1, 1)];
    for (int k = 0; k <= 1; ++k) F[I3(0, 3, k)] = -0.119798935;
    for (int k = 2; k <= 8; ++k) F[I3(0, 3, k)] = F[I3(0, 1, 1)];
    F[I3(1, 2, 0)] = 0.0096495698;
    F[I3(1, 2, 1)] = 0.030;
    F[I3(1, 2, 2)] = -0.0200;
    for (int k = 5; k <= 8; ++k) F[I3(1, 2, k)] = -0.030133632;
    for (int k = 1; k <= 8; ++k) F[I3(1, 3, k)] = -0.124836752;
    for (int k = 0; k <= 8; ++k) F[I3(2, 3, k)] = -0.044709383;
    /* linear ramps between the tabulated end points */
    for (int i = 3; i <= 7; ++i)
        F[I3(2, 2, i)] = F[I3(2, 2, 2)] + (i - 2) * (F[I3(2, 2, 8)] - F[I3(2, 2, 2)]) / 6;
    for (int i = 3; i <= 4; ++i)
        F[I3(1, 2, i)] = F[I3(1, 2, 2)] + (i - 2) * (F[I3(1, 2, 5)] - F[I3(1, 2, 2)]) / 3;
    Fi[I3(2, 1, 0)] = -0.052500;
    for (int k = 4; k <= 8; ++k) Fi[I3(2, 1, k)] = -0.054376;
    Fi[I3(2, 3, 0)] = 0.0;
    for (int k = 1; k <= 5; ++k) Fi[I3(2, 3, k)] = 0.062418;
    for (int k = 3; k <= 7; ++k) Fk[I3(2, 2, k)] = -0.006618;
    for (int k = 6; k <= 8; ++k) Fi[I3(2, 3, k)] = 0.062418;
    Fk[I3(1, 1, 1)] = -0.060543;
    Fk[I3(1, 2, 3)] = -0.020044;
    Fk[I3(1, 2, 4)] = -0.020044;
    for (int k = 0; k <= 9; ++k)
        for (int i = 0; i <= 3; ++i)
            for (int j = i + 1; j <= 3; ++j) {
                double x = F[I3(i, j, k)] + F[I3(j, i, k)];
                F[I3(i, j, k)] = x; F[I3(j, i, k)] = x;
                x = Fi[I3(i, j, k)] + Fj[I3(j, i, k)];
                Fi[I3(i, j, k)] = x; Fj[I3(j, i, k)] = x;
                x = Fi[I3(j, i, k)] + Fj[I3(i, j, k)];
                Fi[I3(j, i, k)] = x; Fj[I3(i, j, k)] = x;
                x = Fk[I3(i, j, k)] + Fk[I3(j, i, k)];
                Fk[I3(i, j, k)] = x; Fk[I3(j, i, k)] = x;
            }
}

static void default_Fch(double *F)
{
    memset(F, 0, sizeof(double) * 250);
    for (int k = 4; k <= 8; ++k) F[I3(0, 2, k)] = -0.0090477875161288110;
    for (int k = 0; k <= 8; ++k) {
        F[I3(1, 3, k)] = -0.213;
        F[I3(1, 2, k)] = -0.25;
        F[I3(1, 1, k)] = -0.5;
    }
    for (int k = 0; k <= 9; ++k)
        for (int i = 0; i <= 2; ++i)
            for (int j = i + 1; j <= 3; ++j) {
                double x = F[I3(i, j, k)] + F[I3(j, i, k)];
                F[I3(i, j, k)] = x; F[I3(j, i, k)] = x;
            }
}

static void default_Fhh(double *F)
{
    memset(F, 0, sizeof(double) * 250);
    F[I3(1, 1, 0)] = 0.249831916;
}

static void default_Tcc(double *T)
{
    memset(T, 0, sizeof(double) * 250);
    T[I3(2, 2, 0)] = -0.070280085;
    for (int k = 1; k <= 8; ++k) T[I3(2, 2, k)] = -0.00809675;
}

static void default_Pcc(double *P)
{
    memset(P, 0, sizeof(double) * 36);
    P[I2(1, 1)] = 0.003026697473481;
    P[I2(2, 0)] = 0.007860700254745;
    P[I2(3, 0)] = 0.016125364564267;
    P[I2(1, 2)] = 0.003179530830731;
    P[I2(2, 1)] = 0.006326248241119;
}

static void default_Pch(double *P)
{
    memset(P, 0, sizeof(double) * 36);
    P[I2(1, 0)] = 0.2093367328250380;
    P[I2(2, 0)] = -0.064449615432525;
    P[I2(3, 0)] = -0.303927546346162;
    P[I2(0, 1)] = 0.01;
    P[I2(0, 2)] = -0.1220421462782555;
    P[I2(1, 1)] = -0.1251234006287090;
    P[I2(2, 1)] = -0.298905245783;
    P[I2(0, 3)] = -0.307584705066;
    P[I2(1, 2)] = -0.3005291724067579;
}

/* ------------------------------------------------------------------------ */
/* g(cos) splines for carbon: rebo2_db.f90:405-524                           */
/* ------------------------------------------------------------------------ */
static int make_cc_g_spline(struct orc_rebo2_s *p)
{
    double A[36], As[36], B[6];
    /* third interval: values at theta(3..6), slope and curvature at theta(3) */
    memset(A, 0, sizeof A);
    for (int i = 2; i < 6; ++i)
        for (int j = 0; j < 6; ++j) A[(i - 2) * 6 + j] = ipow(p->cc_g_theta[i], j);
    double z = p->cc_g_theta[2];
    A[4 * 6 + 1] = 1.0;
    A[5 * 6 + 2] = 2.0;
    for (int j = 2; j < 6; ++j) {
        A[4 * 6 + j] = j * ipow(z, j - 1);
        if (j >= 3) A[5 * 6 + j] = (j - 1) * j * ipow(z, j - 2);
    }
    memcpy(As, A, sizeof A);
    for (int i = 0; i < 4; ++i) B[i] = p->cc_g_g1[2 + i];
    B[4] = p->cc_g_dg1[2];
    B[5] = p->cc_g_d2g1[2];
    if (solve_multi(6, A, 1, B)) return ORC_ERR_SINGULAR;
    memcpy(p->g1c[2], B, sizeof B);
    memcpy(A, As, sizeof A);
    for (int i = 0; i < 4; ++i) B[i] = p->cc_g_g2[2 + i];
    B[4] = p->cc_g_dg1[2];
    B[5] = p->cc_g_d2g1[2];
    if (solve_multi(6, A, 1, B)) return ORC_ERR_SINGULAR;
    memcpy(p->g2c[2], B, sizeof B);
    /* first and second interval: value, slope, curvature at both ends */
    for (int k = 0; k < 2; ++k) {
        memset(A, 0, sizeof A);
        for (int e = 0; e < 2; ++e) {
            double ze = e ? p->cc_g_theta[1 + k] : p->cc_g_theta[k];
            A[(3 * e + 0) * 6 + 0] = 1.0;
            A[(3 * e + 1) * 6 + 1] = 1.0;
            A[(3 * e + 2) * 6 + 2] = 2.0;
            for (int j = 1; j < 6; ++j) {
                A[(3 * e + 0) * 6 + j] = ipow(ze, j);
                if (j >= 2) A[(3 * e + 1) * 6 + j] = j * ipow(ze, j - 1);
                if (j >= 3) A[(3 * e + 2) * 6 + j] = (j - 1) * j * ipow(ze, j - 2);
            }
        }
        B[0] = p->cc_g_g1[k];     B[1] = p->cc_g_dg1[k];     B[2] = p->cc_g_d2g1[k];
        B[3] = p->cc_g_g1[k + 1]; B[4] = p->cc_g_dg1[k + 1]; B[5] = p->cc_g_d2g1[k + 1];
        if (solve_multi(6, A, 1, B)) return ORC_ERR_SINGULAR;
        memcpy(p->g1c[k], B, sizeof B);
        memcpy(p->g2c[k], B, sizeof B);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* construction                                                               */
/* ------------------------------------------------------------------------ */
orc_rebo2_t *orc_rebo2_new(int zero_tables, int with_dihedral)
{
    struct orc_rebo2_s *p = calloc(1, sizeof(struct orc_rebo2_s));
    if (!p) return NULL;
    p->cc_B1 = 12388.79197798; p->cc_B2 = 17.56740646509; p->cc_B3 = 30.71493208065;
    p->cc_beta1 = 4.7204523127; p->cc_beta2 = 1.4332132499; p->cc_beta3 = 1.3826912506;
    p->cc_Q = 0.3134602960833; p->cc_A = 10953.544162170; p->cc_alpha = 4.7465390606595;
    static const double th[6] = {-1.0, -1.0 / 2, -1.0 / 3, 0.0, 1.0 / 2, 1.0};
    const double g1[6] = {SP(-0.01), SP(0.05280), SP(0.09733), SP(0.37545), SP(2.0014), SP(8.0)};
    const double dg1[6] = {SP(0.10400), SP(0.17000), SP(0.40000), 0, 0, 0};
    const double d2g1[6] = {SP(0.00000), SP(0.37000), SP(1.98000), 0, 0, 0};
    const double g2[6] = {0.0, 0.0, SP(0.09733), SP(0.271856), SP(0.416335), SP(1.0)};
    memcpy(p->cc_g_theta, th, sizeof th);
    memcpy(p->cc_g_g1, g1, sizeof g1);
    memcpy(p->cc_g_dg1, dg1, sizeof dg1);
    memcpy(p->cc_g_d2g1, d2g1, sizeof d2g1);
    memcpy(p->cc_g_g2, g2, sizeof g2);
    p->ch_B1 = 32.3551866587; p->ch_beta1 = 1.43445805925; p->ch_Q = 0.340775728;
    p->ch_A = 149.94098723; p->ch_alpha = 4.10254983;
    p->hh_B1 = 29.632593; p->hh_beta1 = 1.71589217; p->hh_Q = 0.370471487045;
    p->hh_A = 32.817355747; p->hh_alpha = 3.536298648;
    static const double spgh[3][6] = {
        {270.467795364007301, 1549.701314596994564, 3781.927258631323866,
         4582.337619544424228, 2721.538161662818368, 630.658598136730774},
        {16.956325544514659, -21.059084522755980, -102.394184748124742,
         -210.527926707779059, -229.759473570467513, -94.968528666251945},
        {19.065031149937783, 2.017732531534021, -2.566444502991983,
         3.291353893907436, -2.653536801884563, 0.837650930130006}};
    memcpy(p->spgh, spgh, sizeof spgh);
    for (int i = 0; i < 25; ++i) p->igh[i] = i < 18 ? 3 : (i < 22 ? 2 : 1);
    p->hhh_lambda = 4.0; p->cc_re = 1.4; p->ch_re = 1.09; p->hh_re = 0.7415886997;
    p->cc_in_r1 = 1.70; p->cc_in_r2 = 2.00;
    p->ch_r1 = 1.30; p->ch_r2 = 1.80;
    p->hh_r1 = 1.10; p->hh_r2 = 1.70;
    p->with_dihedral = with_dihedral != 0;

    /* rebo2_db_init_with_parameters */
    p->conpe[KC] = -0.5; p->conpe[KH] = -0.5;
    p->conan[KC] = 0.5 * p->conpe[KC]; p->conan[KH] = 0.5 * p->conpe[KH];
    p->conpf[KC] = p->conpe[KC] - 1.0; p->conpf[KH] = p->conpe[KH] - 1.0;
    p->conalp = p->hhh_lambda;
    p->conear[PCC][PCC] = 1.0;
    p->conear[PCC][PCH] = exp(p->conalp * (p->ch_re - p->cc_re));
    p->conear[PCC][PHH] = exp(p->conalp * (p->hh_re - p->cc_re));
    p->conear[PCH][PCC] = 1.0 / p->conear[PCC][PCH];
    p->conear[PCH][PCH] = 1.0;
    p->conear[PCH][PHH] = exp(p->conalp * (p->hh_re - p->ch_re));
    p->conear[PHH][PCC] = 1.0 / p->conear[PCC][PHH];
    p->conear[PHH][PCH] = 1.0 / p->conear[PCH][PHH];
    p->conear[PHH][PHH] = 1.0;
    p->cut_l[PCC] = p->cc_in_r1; p->cut_l[PCH] = p->ch_r1; p->cut_l[PHH] = p->hh_r1;
    p->cut_h[PCC] = p->cc_in_r2; p->cut_h[PCH] = p->ch_r2; p->cut_h[PHH] = p->hh_r2;
    for (int t = 0; t < 7; ++t) p->cut_h2[t] = p->cut_h[t] * p->cut_h[t];

    int err = make_cc_g_spline(p);
    double *F = malloc(sizeof(double) * 250 * 4), *P = malloc(sizeof(double) * 36);
    if (!F || !P) err = ORC_ERR_ARG;
    if (!err) {
        default_Fcc(F, F + 250, F + 500, F + 750);
        if (zero_tables) memset(F, 0, sizeof(double) * 1000);
        err = table3d_init(4, 4, 9, F, F + 250, F + 500, F + 750, p->Fcc);
    }
    if (!err) {
        default_Fch(F);
        if (zero_tables) memset(F, 0, sizeof(double) * 250);
        err = table3d_init(4, 4, 9, F, NULL, NULL, NULL, p->Fch);
    }
    if (!err) {
        default_Fhh(F);
        if (zero_tables) memset(F, 0, sizeof(double) * 250);
        err = table3d_init(4, 4, 9, F, NULL, NULL, NULL, p->Fhh);
    }
    if (!err) {
        default_Tcc(F);
        if (zero_tables) memset(F, 0, sizeof(double) * 250);
        err = table3d_init(4, 4, 9, F, NULL, NULL, NULL, p->Tcc);
    }
    if (!err) {
        default_Pcc(P);
        if (zero_tables) memset(P, 0, sizeof(double) * 36);
        err = table2d_init(5, 5, P, p->Pcc);
    }
    if (!err) {
        default_Pch(P);
        if (zero_tables) memset(P, 0, sizeof(double) * 36);
        err = table2d_init(5, 5, P, p->Pch);
    }
    free(F);
    free(P);
    if (err) { free(p); return NULL; }
    return p;
}

void orc_rebo2_free(orc_rebo2_t *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* functional forms, rebo2_func.f90                                           */
/* ------------------------------------------------------------------------ */
static int z2pair(int ki, int kj)
{
    if (ki == KC) return kj;
    if (kj == KC) return ki;
    return ki + kj;
}

static void fconj(double x, double *fx, double *dfx)
{
    if (x <= 2.0) { *fx = 1.0; *dfx = 0.0; }
    else if (x >= 3.0) { *fx = 0.0; *dfx = 0.0; }
    else {
        double arg = PI_ * (x - 2.0);
        *fx = 0.5 * (1.0 + cos(arg));
        *dfx = -0.5 * PI_ * sin(arg);
    }
}

static void fcin(const struct orc_rebo2_s *p, int t, double r, double *v, double *dv)
{
    double l = p->cut_l[t], h = p->cut_h[t];
    if (r > h) { *v = 0.0; *dv = 0.0; }
    else if (r < l) { *v = 1.0; *dv = 0.0; }
    else if (r <= l) { *v = 1.0; *dv = 0.0; }
    else if (r >= h) { *v = 0.0; *dv = 0.0; }
    else {
        double fac = PI_ / (h - l);
        double x = fac * (r - l);
        *v = 0.5 * (1.0 + cos(x));
        *dv = -0.5 * fac * sin(x);
    }
}

static void va_fn(const struct orc_rebo2_s *p, int t, double r, double *v, double *dv)
{
    if (t == PCC) {
        double e1 = p->cc_B1 * exp(-p->cc_beta1 * r);
        double e2 = p->cc_B2 * exp(-p->cc_beta2 * r);
        double e3 = p->cc_B3 * exp(-p->cc_beta3 * r);
        *v = -(e1 + e2 + e3);
        *dv = -(-p->cc_beta1 * e1 - p->cc_beta2 * e2 - p->cc_beta3 * e3);
    } else if (t == PCH) {
        double e1 = p->ch_B1 * exp(-p->ch_beta1 * r);
        *v = -e1;
        *dv = p->ch_beta1 * e1;
    } else {
        double e1 = p->hh_B1 * exp(-p->hh_beta1 * r);
        *v = -e1;
        *dv = p->hh_beta1 * e1;
    }
}

static void vr_fn(const struct orc_rebo2_s *p, int t, double r, double *v, double *dv)
{
    double A, Q, al;
    if (t == PCC) { A = p->cc_A; Q = p->cc_Q; al = p->cc_alpha; }
    else if (t == PCH) { A = p->ch_A; Q = p->ch_Q; al = p->ch_alpha; }
    else { A = p->hh_A; Q = p->hh_Q; al = p->hh_alpha; }
    double e1 = A * exp(-al * r);
    double hl = 1 + Q / r;
    *v = hl * e1;
    *dv = (-Q / (r * r) - hl * al) * e1;
}

static void poly5(const double *c, double x, double *v, double *dv)
{
    double h = c[0] + c[1] * x, dh = c[1];
    for (int i = 2; i < 6; ++i) {
        h = h + c[i] * ipow(x, i);
        dh = dh + i * c[i] * ipow(x, i - 1);
    }
    *v = h; *dv = dh;
}

static void cc_g_from_spline(const struct orc_rebo2_s *p, const double (*c)[6], double x,
                             double *v, double *dv)
{
    int j = x < p->cc_g_theta[1] ? 0 : (x < p->cc_g_theta[2] ? 1 : 2);
    poly5(c[j], x, v, dv);
}

static void g_fn(const struct orc_rebo2_s *p, int ktyp, double costh, double n, double *val,
                 double *dcos, double *dn)
{
    if (ktyp == KC) {
        if (n < 3.2) {
            cc_g_from_spline(p, p->g2c, costh, val, dcos);
            *dn = 0.0;
        } else if (n > SP(3.7)) {
            cc_g_from_spline(p, p->g1c, costh, val, dcos);
            *dn = 0.0;
        } else {
            double v1, v2, d1, d2;
            cc_g_from_spline(p, p->g1c, costh, &v1, &d1);
            cc_g_from_spline(p, p->g2c, costh, &v2, &d2);
            double arg = 2 * PI_ * (n - 3.2);
            double s = (1 + cos(arg)) / 2;
            double ds = -PI_ * sin(arg);
            *val = v1 * (1 - s) + v2 * s;
            *dcos = d1 * (1 - s) + d2 * s;
            *dn = (v2 - v1) * ds;
        }
    } else {
        int ig = p->igh[(int)(-costh * 12.0) + 12];
        poly5(p->spgh[ig - 1], costh, val, dcos);
        *dn = 0.0;
    }
}

static void bo_fn(const struct orc_rebo2_s *p, int ktypi, double z, double fc, double fa,
                  double *b, double *db)
{
    double arg = 1.0 + z;
    *b = pow(arg, p->conpe[ktypi]);
    *db = p->conan[ktypi] * fc * fa * pow(arg, p->conpf[ktypi]);
}

static void h_fn(const struct orc_rebo2_s *p, int ijpot, int ikpot, double dr, double *v,
                 double *dv)
{
    if (ijpot + ikpot <= 4) { *v = 1.0; *dv = 0.0; }
    else {
        *v = p->conear[ijpot][ikpot] * exp(p->conalp * dr);
        *dv = p->conalp * *v;
    }
}

/* entry points for the unit tests of the building blocks */
int orc_rebo2_eval_g(const orc_rebo2_t *p, int ktyp, double costh, double n, double *out3)
{
    g_fn(p, ktyp, costh, n, &out3[0], &out3[1], &out3[2]);
    return ORC_OK;
}

int orc_rebo2_eval_table(const orc_rebo2_t *p, int which, double x, double y, double z,
                         double *out4)
{
    out4[3] = 0.0;
    switch (which) {
    case 0: table3d_eval(p->Fcc, 4, 4, 9, x, y, z, &out4[0], &out4[1], &out4[2], &out4[3]); break;
    case 1: table3d_eval(p->Fch, 4, 4, 9, x, y, z, &out4[0], &out4[1], &out4[2], &out4[3]); break;
    case 2: table3d_eval(p->Fhh, 4, 4, 9, x, y, z, &out4[0], &out4[1], &out4[2], &out4[3]); break;
    case 3: table2d_eval(p->Pcc, 5, 5, x, y, &out4[0], &out4[1], &out4[2]); break;
    case 4: table2d_eval(p->Pch, 5, 5, x, y, &out4[0], &out4[1], &out4[2]); break;
    case 5: table3d_eval(p->Tcc, 4, 4, 9, x, y, z, &out4[0], &out4[1], &out4[2], &out4[3]); break;
    default: return ORC_ERR_ARG;
    }
    return ORC_OK;
}

int orc_rebo2_default_table(int which, double *out)
{
    double F[1000];
    switch (which) {
    case 0: default_Fcc(F, F + 250, F + 500, F + 750); memcpy(out, F, sizeof F); break;
    case 1: default_Fch(out); break;
    case 2: default_Fhh(out); break;
    case 3: default_Pcc(out); break;
    case 4: default_Pch(out); break;
    case 5: default_Tcc(out); break;
    default: return ORC_ERR_ARG;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* kernel                                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int j, typ;
    int dc[3];
    double len, nm[3], fc, dfc;
    double fx, dfx;     /* conjugation F(x_ik) and dF/dx for carbon k, else 0 */
    int64_t jn;          /* list entry of the bond (nbb, bop_kernel_rebo2.f90) */
} rbond;

static int dc_is_pos(const int *dc)
{
    /* sign of DCELL_INDEX: the first non-zero component decides */
    if (dc[0]) return dc[0] > 0;
    if (dc[1]) return dc[1] > 0;
    return dc[2] > 0;
}

static int ktyp_of(int Z) { return Z == 6 ? KC : (Z == 1 ? KH : 0); }

static inline void outer_add(double *w, double s, const double *a, const double *b)
{
    /* w(x, y) += s * a(x) * b(y) with w[x + 3*y] */
    for (int y = 0; y < 3; ++y)
        for (int x = 0; x < 3; ++x) w[x + 3 * y] = w[x + 3 * y] + s * (a[x] * b[y]);
}

int orc_rebo2(const orc_rebo2_t *p, int nat, const double *r, const int *Z,
              const double *cell, const int64_t *offsets, const int *nbr, const int *dc,
              double *epot, double *f, double *wpot, double *epot_per_at,
              double *wpot_per_at, double *epot_per_bond, double *f_per_bond,
              double *wpot_per_bond, int nthreads)
{
    int *ktyp = malloc(sizeof(int) * (nat > 0 ? nat : 1));
    int64_t *bseed = malloc(sizeof(int64_t) * (nat + 1));
    int64_t cap = offsets[nat] > 0 ? offsets[nat] : 1;
    rbond *bd = malloc(sizeof(rbond) * cap);
    double *nn = calloc((size_t)4 * (nat > 0 ? nat : 1), sizeof(double));
    double *nconjit = calloc((size_t)(nat > 0 ? nat : 1), sizeof(double));
    if (!ktyp || !bseed || !bd || !nn || !nconjit) {
        free(ktyp); free(bseed); free(bd); free(nn); free(nconjit);
        return ORC_ERR_ARG;
    }
    for (int i = 0; i < nat; ++i) ktyp[i] = ktyp_of(Z[i]);

    /* loop1: internal bond list (bop_kernel_rebo2.f90:700-1181) */
    int64_t nb = 0;
    for (int i = 0; i < nat; ++i) {
        bseed[i] = nb;
        int ki = ktyp[i];
        if (ki <= 0) continue;
        for (int64_t s = offsets[i]; s < offsets[i + 1]; ++s) {
            int j = nbr[s];
            int kj = ktyp[j];
            if (kj <= 0) continue;
            const int *d = dc + 3 * s;
            double rij[3];
            for (int x = 0; x < 3; ++x) {
                double as = 0.0;
                as = as + cell[0 + x] * (double)d[0];
                as = as + cell[3 + x] * (double)d[1];
                as = as + cell[6 + x] * (double)d[2];
                rij[x] = -((r[3 * i + x] - r[3 * j + x]) + as);
            }
            double rl = 0.0;
            rl = rl + rij[0] * rij[0];
            rl = rl + rij[1] * rij[1];
            rl = rl + rij[2] * rij[2];
            int t = z2pair(ki, kj);
            rbond b;
            if (rl < p->cut_l[t] * p->cut_l[t]) {
                b.fc = 1.0; b.dfc = 0.0;
                b.len = sqrt(rl);
            } else if (rl < p->cut_h2[t]) {
                b.len = sqrt(rl);
                fcin(p, t, b.len, &b.fc, &b.dfc);
            } else {
                continue;
            }
            for (int x = 0; x < 3; ++x) b.nm[x] = rij[x] / b.len;
            b.j = j; b.typ = t; b.jn = s;
            b.dc[0] = d[0]; b.dc[1] = d[1]; b.dc[2] = d[2];
            b.fx = 0.0; b.dfx = 0.0;
            bd[nb++] = b;
        }
    }
    bseed[nat] = nb;
    /* coordination numbers (bop_kernel_rebo2.f90:1190-1205) */
    for (int i = 0; i < nat; ++i)
        for (int64_t b = bseed[i]; b < bseed[i + 1]; ++b)
            nn[4 * i + ktyp[bd[b].j]] += bd[b].fc;
    /* conjugation per bond (ik_loop1, bop_kernel_rebo2.f90:1231-1325) */
    for (int i = 0; i < nat; ++i) {
        double s = 0.0;
        for (int64_t b = bseed[i]; b < bseed[i + 1]; ++b) {
            int k = bd[b].j;
            if (ktyp[k] == KC) {
                double xik = nn[4 * k + KC] + nn[4 * k + KH] - bd[b].fc;
                fconj(xik, &bd[b].fx, &bd[b].dfx);
                s = s + bd[b].fc * bd[b].fx;
            }
        }
        nconjit[i] = s;
    }

    int nth = nthreads > 0 ? nthreads : 1;
#ifndef _OPENMP
    nth = 1;
#endif
    double *fbuf = calloc((size_t)nth * 3 * (nat > 0 ? nat : 1), sizeof(double));
    double *pebuf = calloc((size_t)nth * (nat > 0 ? nat : 1), sizeof(double));
    double *wbuf = calloc((size_t)nth * 9, sizeof(double));
    double *wabuf = wpot_per_at ? calloc((size_t)nth * 9 * (nat > 0 ? nat : 1), sizeof(double)) : NULL;
    int err = ORC_OK;
    if (!fbuf || !pebuf || !wbuf || (wpot_per_at && !wabuf)) err = ORC_ERR_ARG;

    if (!err) {
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *fl = fbuf + (size_t)tid * 3 * nat;
        double *pe = pebuf + (size_t)tid * nat;
        double *wl = wbuf + 9 * tid;
        double *wal = wabuf ? wabuf + (size_t)tid * 9 * nat : NULL;
        int maxb = 0;
        for (int i = 0; i < nat; ++i) {
            int c = (int)(bseed[i + 1] - bseed[i]);
            if (c > maxb) maxb = c;
        }
        double (*dbidk)[3] = malloc(sizeof(double[3]) * (maxb + 1));
        double (*dbjdl)[3] = malloc(sizeof(double[3]) * (maxb + 1));

#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int i = 0; i < nat; ++i) {
            int ki = ktyp[i];
            if (ki <= 0) continue;
            double fi[3] = {0, 0, 0};
            int64_t is = bseed[i], ie = bseed[i + 1];
            for (int64_t ij = is; ij < ie; ++ij) {
                const rbond *bij_ = &bd[ij];
                int j = bij_->j;
                if (!((bij_->dc[0] == 0 && bij_->dc[1] == 0 && bij_->dc[2] == 0 && j > i) ||
                      dc_is_pos(bij_->dc)))
                    continue;
                int ijpot = bij_->typ;
                double rlij = bij_->len;
                if (!(rlij < p->cut_h[ijpot])) continue;
                double fj[3] = {0, 0, 0};
                int kj = ktyp[j];
                double rlijr = 1.0 / rlij;
                const double *rnij = bij_->nm;
                double rij[3] = {rlij * rnij[0], rlij * rnij[1], rlij * rnij[2]};
                double fcarij = bij_->fc, dfcarijr = bij_->dfc;
                double ni[4] = {0, nn[4 * i + 1], 0, nn[4 * i + 3]};
                double nj[4] = {0, nn[4 * j + 1], 0, nn[4 * j + 3]};
                ni[kj] = ni[kj] - bij_->fc;
                nj[ki] = nj[ki] - bij_->fc;
                if (ni[KC] > 4.0) ni[KC] = 4.0;
                if (ni[KH] > 4.0) ni[KH] = 4.0;
                double nti = ni[KC] + ni[KH];
                if (nj[KC] > 4.0) nj[KC] = 4.0;
                if (nj[KH] > 4.0) nj[KH] = 4.0;
                double ntj = nj[KC] + nj[KH];
                double faij, dfaijr, frij, dfrijr;
                va_fn(p, ijpot, rlij, &faij, &dfaijr);
                vr_fn(p, ijpot, rlij, &frij, &dfrijr);
                double wij[9] = {0}, wijb[9] = {0}, wjib[9] = {0};
                double zij = 0.0, dbidi[3] = {0, 0, 0}, dbidj[3] = {0, 0, 0}, dzdni = 0.0;
                double nconji = 0.0;
                /* ik_loop2: zeta_ij and its derivatives */
                for (int64_t ik = is; ik < ie; ++ik) {
                    int ikc = (int)(ik - is);
                    const rbond *bk = &bd[ik];
                    double fcik = bk->fc;
                    if (ik != ij) {
                        int ikpot = bk->typ;
                        double rlik = bk->len;
                        if (rlik < p->cut_h[ikpot]) {
                            const double *rnik = bk->nm;
                            double rik[3] = {rlik * rnik[0], rlik * rnik[1], rlik * rnik[2]};
                            double dfcikr = bk->dfc;
                            double qfacan, qfadan, gfacan, gddan, dgdn;
                            h_fn(p, ijpot, ikpot, rlij - rlik, &qfacan, &qfadan);
                            double costh = 0.0;
                            costh = costh + rnik[0] * rnij[0];
                            costh = costh + rnik[1] * rnij[1];
                            costh = costh + rnik[2] * rnij[2];
                            g_fn(p, ki, costh, nti, &gfacan, &gddan, &dgdn);
                            double dkc[3], dd = 0.0;
                            for (int x = 0; x < 3; ++x) dkc[x] = rnik[x] * rlik - rnij[x] * rlij;
                            dd = dd + dkc[0] * dkc[0];
                            dd = dd + dkc[1] * dkc[1];
                            dd = dd + dkc[2] * dkc[2];
                            double disjk = sqrt(dd);
                            for (int x = 0; x < 3; ++x) dkc[x] = dkc[x] / disjk;
                            double dcsdij = 1.0 / rlik - costh * rlijr;
                            double dcsdik = rlijr - costh / rlik;
                            double dcsdjk = -disjk * rlijr / rlik;
                            double dcsdi[3], dcsdj[3], dcsdk[3];
                            for (int x = 0; x < 3; ++x) {
                                dcsdi[x] = -dcsdij * rnij[x] - dcsdik * rnik[x];
                                dcsdj[x] = dcsdij * rnij[x] - dcsdjk * dkc[x];
                                dcsdk[x] = dcsdik * rnik[x] + dcsdjk * dkc[x];
                            }
                            dzdni = dzdni + fcik * dgdn * qfacan;
                            double dzfac = fcik * gddan * qfacan;
                            zij = zij + fcik * gfacan * qfacan;
                            double dzdrij = gfacan * fcik * qfadan;
                            double dzdrik = gfacan * (dfcikr * qfacan - fcik * qfadan);
                            double df[3];
                            for (int x = 0; x < 3; ++x) {
                                dbidi[x] = dbidi[x] - dzdrij * rnij[x] - dzdrik * rnik[x] + dzfac * dcsdi[x];
                                df[x] = dzdrij * rnij[x] + dzfac * dcsdj[x];
                                dbidj[x] = dbidj[x] + df[x];
                                dbidk[ikc][x] = dzdrik * rnik[x] + dzfac * dcsdk[x];
                            }
                            for (int y = 0; y < 3; ++y)
                                for (int x = 0; x < 3; ++x)
                                    wijb[x + 3 * y] = wijb[x + 3 * y] - rij[x] * df[y] - rik[x] * dbidk[ikc][y];
                        } else {
                            dbidk[ikc][0] = dbidk[ikc][1] = dbidk[ikc][2] = 0.0;
                        }
                    } else {
                        nconji = nconjit[i] - fcik * bk->fx;
                    }
                }
                double pij = 0.0, dpdnci = 0.0, dpdnhi = 0.0;
                if (ki == KC) {
                    table2d_eval(ijpot == PCC ? p->Pcc : p->Pch, 5, 5, ni[KH], ni[KC], &pij,
                                 &dpdnhi, &dpdnci);
                    zij = zij + pij;
                    dpdnci = dpdnci + dzdni;
                    dpdnhi = dpdnhi + dzdni;
                }
                double bij, dfbij;
                bo_fn(p, ki, zij, fcarij, faij, &bij, &dfbij);

                /* jl_loop: zeta_ji, conjugation of j */
                double zji = 0.0, dbjdi[3] = {0, 0, 0}, dbjdj[3] = {0, 0, 0}, dzdnj = 0.0;
                double nconjj = 0.0;
                int64_t js = bseed[j], je = bseed[j + 1];
                for (int64_t jl = js; jl < je; ++jl) {
                    int jlc = (int)(jl - js);
                    const rbond *bl = &bd[jl];
                    int l = bl->j;
                    int ldc[3] = {bij_->dc[0] + bl->dc[0], bij_->dc[1] + bl->dc[1], bij_->dc[2] + bl->dc[2]};
                    if (l == i && ldc[0] == 0 && ldc[1] == 0 && ldc[2] == 0) {
                        dbjdl[jlc][0] = dbjdl[jlc][1] = dbjdl[jlc][2] = 0.0;
                        continue;
                    }
                    int jlpot = bl->typ;
                    double rljl = bl->len;
                    const double *rnjl = bl->nm;
                    if (ktyp[l] == KC) nconjj = nconjj + bl->fc * bl->fx;
                    if (rljl < p->cut_h[jlpot]) {
                        double rjl[3] = {rljl * rnjl[0], rljl * rnjl[1], rljl * rnjl[2]};
                        double fcjl = bl->fc, dfcjlr = bl->dfc;
                        double qfacan, qfadan, gfacan, gddan, dgdn;
                        h_fn(p, ijpot, jlpot, rlij - rljl, &qfacan, &qfadan);
                        double costh = 0.0;
                        costh = costh + rnjl[0] * rnij[0];
                        costh = costh + rnjl[1] * rnij[1];
                        costh = costh + rnjl[2] * rnij[2];
                        costh = -costh;
                        g_fn(p, kj, costh, ntj, &gfacan, &gddan, &dgdn);
                        double dlc[3], dd = 0.0;
                        for (int x = 0; x < 3; ++x) dlc[x] = rnjl[x] * rljl + rnij[x] * rlij;
                        dd = dd + dlc[0] * dlc[0];
                        dd = dd + dlc[1] * dlc[1];
                        dd = dd + dlc[2] * dlc[2];
                        double disil = sqrt(dd);
                        for (int x = 0; x < 3; ++x) dlc[x] = dlc[x] / disil;
                        double dcsdji = 1.0 / rljl - costh * rlijr;
                        double dcsdjl = rlijr - costh / rljl;
                        double dcsdil = -disil * rlijr / rljl;
                        double dcsdj[3], dcsdi[3], dcsdl[3];
                        for (int x = 0; x < 3; ++x) {
                            dcsdj[x] = dcsdji * rnij[x] - dcsdjl * rnjl[x];
                            dcsdi[x] = -dcsdji * rnij[x] - dcsdil * dlc[x];
                            dcsdl[x] = dcsdjl * rnjl[x] + dcsdil * dlc[x];
                        }
                        dzdnj = dzdnj + fcjl * dgdn * qfacan;
                        double dzfac = fcjl * gddan * qfacan;
                        zji = zji + fcjl * gfacan * qfacan;
                        double dzdrji = gfacan * fcjl * qfadan;
                        double dzdrjl = gfacan * (dfcjlr * qfacan - fcjl * qfadan);
                        double df[3];
                        for (int x = 0; x < 3; ++x) {
                            dbjdj[x] = dbjdj[x] + dzdrji * rnij[x] - dzdrjl * rnjl[x] + dzfac * dcsdj[x];
                            df[x] = -dzdrji * rnij[x] + dzfac * dcsdi[x];
                            dbjdi[x] = dbjdi[x] + df[x];
                            dbjdl[jlc][x] = dzdrjl * rnjl[x] + dzfac * dcsdl[x];
                        }
                        for (int y = 0; y < 3; ++y)
                            for (int x = 0; x < 3; ++x)
                                wjib[x + 3 * y] = wjib[x + 3 * y] + rij[x] * df[y] - rjl[x] * dbjdl[jlc][y];
                    } else {
                        dbjdl[jlc][0] = dbjdl[jlc][1] = dbjdl[jlc][2] = 0.0;
                    }
                }
                double pji = 0.0, dpdncj = 0.0, dpdnhj = 0.0;
                if (kj == KC) {
                    table2d_eval(ijpot == PCC ? p->Pcc : p->Pch, 5, 5, nj[KH], nj[KC], &pji,
                                 &dpdnhj, &dpdncj);
                    zji = zji + pji;
                    dpdncj = dpdncj + dzdnj;
                    dpdnhj = dpdnhj + dzdnj;
                }
                double bji, dfbji;
                bo_fn(p, kj, zji, fcarij, faij, &bji, &dfbji);

                double nconj = nconji * nconji + nconjj * nconjj;
                if (nconj > 8.0) nconj = 8.0;
                if (nti > 3.0) nti = 3.0;
                if (ntj > 3.0) ntj = 3.0;
                /* torsion (bop_kernel_rebo2.f90:1939-2120, non-ALT_DIHEDRAL form) */
                double bdh = 0.0, tij = 0.0, dtdni = 0.0, dtdnj = 0.0, dtdncn = 0.0;
                if (p->with_dihedral && ijpot == PCC) {
                    table3d_eval(p->Tcc, 4, 4, 9, nti, ntj, nconj, &tij, &dtdni, &dtdnj, &dtdncn);
                    double tije = tij * faij * fcarij;
                    if (tij != 0) {
                        for (int64_t ik = is; ik < ie; ++ik) {
                            if (ik == ij) continue;
                            const rbond *bk = &bd[ik];
                            int k = bk->j;
                            double rlik = bk->len;
                            const double *rnik = bk->nm;
                            double fcik = bk->fc, dfcikr = bk->dfc;
                            double dot_ij_ik = 0.0;
                            dot_ij_ik = dot_ij_ik + rnij[0] * rnik[0];
                            dot_ij_ik = dot_ij_ik + rnij[1] * rnik[1];
                            dot_ij_ik = dot_ij_ik + rnij[2] * rnik[2];
                            double dcik = 1.0 - dot_ij_ik * dot_ij_ik;
                            for (int64_t jl = js; jl < je; ++jl) {
                                const rbond *bl = &bd[jl];
                                int l = bl->j;
                                int ldc[3] = {bij_->dc[0] + bl->dc[0], bij_->dc[1] + bl->dc[1],
                                              bij_->dc[2] + bl->dc[2]};
                                if (l == i && ldc[0] == 0 && ldc[1] == 0 && ldc[2] == 0) continue;
                                if (l == k && ldc[0] == bk->dc[0] && ldc[1] == bk->dc[1] &&
                                    ldc[2] == bk->dc[2])
                                    continue;
                                double rljl = bl->len;
                                const double *rnjl = bl->nm;
                                double fcjl = bl->fc, dfcjlr = bl->dfc;
                                double dot_ij_jl = 0.0, dot_ik_jl = 0.0;
                                dot_ij_jl = dot_ij_jl + rnij[0] * rnjl[0];
                                dot_ij_jl = dot_ij_jl + rnij[1] * rnjl[1];
                                dot_ij_jl = dot_ij_jl + rnij[2] * rnjl[2];
                                dot_ik_jl = dot_ik_jl + rnik[0] * rnjl[0];
                                dot_ik_jl = dot_ik_jl + rnik[1] * rnjl[1];
                                dot_ik_jl = dot_ik_jl + rnik[2] * rnjl[2];
                                double dcjl = 1.0 - dot_ij_jl * dot_ij_jl;
                                double abs_dc = sqrt(dcik * dcjl);
                                double cost = (dot_ij_ik * dot_ij_jl - dot_ik_jl) / abs_dc;
                                double bdhij = 1 - cost * cost;
                                bdh = bdh + bdhij * fcik * fcjl;
                                bdhij = bdhij * tij * faij * fcarij / 2;
                                double dbdhij = -2 * cost * tije * fcik * fcjl / 2;
                                double df[3], v[3];
                                for (int x = 0; x < 3; ++x)
                                    df[x] = dbdhij *
                                            ((dot_ij_jl / abs_dc + cost * dot_ij_ik / dcik) * rnik[x] +
                                             (dot_ij_ik / abs_dc + cost * dot_ij_jl / dcjl) * rnjl[x] -
                                             (2 * dot_ik_jl / abs_dc + cost * (1.0 / dcik + 1.0 / dcjl)) * rnij[x]) /
                                            rlij;
                                for (int x = 0; x < 3; ++x) {
                                    fi[x] += df[x];
                                    fj[x] -= df[x];
                                    v[x] = rlij * rnij[x];
                                }
                                outer_add(wij, 1.0, v, df);
                                for (int x = 0; x < 3; ++x)
                                    df[x] = dbdhij *
                                                (-1.0 / dcik * cost * rnik[x] - 1.0 / abs_dc * rnjl[x] +
                                                 (dot_ij_jl / abs_dc + cost * dot_ij_ik / dcik) * rnij[x]) /
                                                rlik +
                                            bdhij * dfcikr * fcjl * rnik[x];
                                for (int x = 0; x < 3; ++x) {
                                    fi[x] += df[x];
                                    fl[3 * k + x] += -df[x];
                                    v[x] = rlik * rnik[x];
                                }
                                outer_add(wij, 1.0, v, df);
                                for (int x = 0; x < 3; ++x)
                                    df[x] = dbdhij *
                                                (-1.0 / dcjl * cost * rnjl[x] - 1.0 / abs_dc * rnik[x] +
                                                 (dot_ij_ik / abs_dc + cost * dot_ij_jl / dcjl) * rnij[x]) /
                                                rljl +
                                            bdhij * fcik * dfcjlr * rnjl[x];
                                for (int x = 0; x < 3; ++x) {
                                    fj[x] += df[x];
                                    fl[3 * l + x] += -df[x];
                                    v[x] = rljl * rnjl[x];
                                }
                                outer_add(wij, 1.0, v, df);
                            }
                        }
                    }
                }
                double fij = 0.0, dfdni = 0.0, dfdnj = 0.0, dfdncn = 0.0;
                if (ijpot == PCC)
                    table3d_eval(p->Fcc, 4, 4, 9, nti, ntj, nconj, &fij, &dfdni, &dfdnj, &dfdncn);
                else if (ijpot == PHH)
                    table3d_eval(p->Fhh, 4, 4, 9, nti, ntj, nconj, &fij, &dfdni, &dfdnj, &dfdncn);
                else if (ki == KC)
                    table3d_eval(p->Fch, 4, 4, 9, ntj, nti, nconj, &fij, &dfdnj, &dfdni, &dfdncn);
                else if (kj == KC)
                    table3d_eval(p->Fch, 4, 4, 9, nti, ntj, nconj, &fij, &dfdni, &dfdnj, &dfdncn);
                dfdni = dfdni + dtdni * bdh;
                dfdnj = dfdnj + dtdnj * bdh;
                dfdncn = dfdncn + dtdncn * bdh;
                dfdni = 0.5 * fcarij * faij * dfdni;
                dfdnj = 0.5 * fcarij * faij * dfdnj;
                dfdncn = 0.5 * fcarij * faij * dfdncn;
                double dfdncni = 2 * dfdncn * nconji;
                double dfdncnj = 2 * dfdncn * nconjj;

                /* forces on neighbours of i from N_i, N_conj (and their neighbours) */
                for (int64_t ik = is; ik < ie; ++ik) {
                    const rbond *bk = &bd[ik];
                    int k = bk->j;
                    if (ik == ij) continue;
                    double dn[3], dncn[3], dri[3], df[3];
                    int kc = ktyp[k] == KC;
                    for (int x = 0; x < 3; ++x) {
                        dn[x] = bk->nm[x] * bk->dfc;
                        dncn[x] = kc ? (bk->fx * bk->dfc) * bk->nm[x] : 0.0;
                        dri[x] = bk->len * bk->nm[x];
                    }
                    double dpn = ktyp[k] == KC ? dpdnci : dpdnhi;
                    for (int x = 0; x < 3; ++x)
                        df[x] = -(dfdni * (dn[x] + 0.0) + dfdncni * dncn[x]) - dfbij * (dpn * dn[x] + 0.0);
                    for (int x = 0; x < 3; ++x) {
                        fl[3 * k + x] += df[x];
                        fi[x] -= df[x];
                    }
                    outer_add(wij, -1.0, dri, df);
                    if (kc) {
                        double nconjdx = bk->fc * bk->dfx;
                        for (int64_t km = bseed[k]; km < bseed[k + 1]; ++km) {
                            const rbond *bm = &bd[km];
                            int m = bm->j;
                            int mdc[3] = {bk->dc[0] + bm->dc[0], bk->dc[1] + bm->dc[1], bk->dc[2] + bm->dc[2]};
                            if (m == i && mdc[0] == 0 && mdc[1] == 0 && mdc[2] == 0) continue;
                            double drk[3], d2[3];
                            for (int x = 0; x < 3; ++x) {
                                drk[x] = bm->len * bm->nm[x];
                                d2[x] = -dfdncni * (nconjdx * (bm->dfc * bm->nm[x]));
                            }
                            for (int x = 0; x < 3; ++x) {
                                fl[3 * m + x] += d2[x];
                                fl[3 * k + x] += -d2[x];
                            }
                            outer_add(wij, -1.0, drk, d2);
                        }
                    }
                }
                /* forces on neighbours of j */
                for (int64_t jl = js; jl < je; ++jl) {
                    const rbond *bl = &bd[jl];
                    int l = bl->j;
                    int ldc[3] = {bij_->dc[0] + bl->dc[0], bij_->dc[1] + bl->dc[1], bij_->dc[2] + bl->dc[2]};
                    if (l == i && ldc[0] == 0 && ldc[1] == 0 && ldc[2] == 0) continue;
                    double dn[3], dncn[3], drj[3], df[3];
                    int lc = ktyp[l] == KC;
                    for (int x = 0; x < 3; ++x) {
                        dn[x] = bl->nm[x] * bl->dfc;
                        dncn[x] = lc ? (bl->fx * bl->dfc) * bl->nm[x] : 0.0;
                        drj[x] = bl->len * bl->nm[x];
                    }
                    double dpn = lc ? dpdncj : dpdnhj;
                    for (int x = 0; x < 3; ++x)
                        df[x] = -(dfdnj * (dn[x] + 0.0) + dfdncnj * dncn[x]) - dfbji * (dpn * dn[x] + 0.0);
                    for (int x = 0; x < 3; ++x) {
                        fl[3 * l + x] += df[x];
                        fj[x] -= df[x];
                    }
                    outer_add(wij, -1.0, drj, df);
                    if (lc) {
                        double nconjdx = bl->fc * bl->dfx;
                        for (int64_t ln = bseed[l]; ln < bseed[l + 1]; ++ln) {
                            const rbond *bn = &bd[ln];
                            int n = bn->j;
                            int ndc[3] = {ldc[0] + bn->dc[0], ldc[1] + bn->dc[1], ldc[2] + bn->dc[2]};
                            if (n == j && ndc[0] == bij_->dc[0] && ndc[1] == bij_->dc[1] &&
                                ndc[2] == bij_->dc[2])
                                continue;
                            double drl[3], d2[3];
                            for (int x = 0; x < 3; ++x) {
                                drl[x] = bn->len * bn->nm[x];
                                d2[x] = -dfdncnj * (nconjdx * (bn->dfc * bn->nm[x]));
                            }
                            for (int x = 0; x < 3; ++x) {
                                fl[3 * n + x] += d2[x];
                                fl[3 * l + x] += -d2[x];
                            }
                            outer_add(wij, -1.0, drl, d2);
                        }
                    }
                }

                /* pair energy and the remaining forces */
                double baveij = 0.5 * (bij + bji + fij + tij * bdh);
                double hlfvij = fcarij * (frij + baveij * faij) / 2;
                pe[i] += hlfvij;
                pe[j] += hlfvij;
                /* per-bond properties on the half pair's entry i->j (one
                 * thread per centre i, so no two threads share an entry):
                 * bop_kernel_rebo2.f90:2546-2547 */
                if (epot_per_bond) epot_per_bond[bij_->jn] += 2 * hlfvij;
                double dffac = dfrijr * fcarij + baveij * dfaijr * fcarij + frij * dfcarijr +
                               baveij * faij * dfcarijr;
                double df[3];
                for (int x = 0; x < 3; ++x) {
                    df[x] = dffac * rnij[x];
                    fi[x] += df[x];
                    fj[x] -= df[x];
                }
                /* :2571-2572, the pair term only */
                if (f_per_bond)
                    for (int x = 0; x < 3; ++x) f_per_bond[3 * bij_->jn + x] += df[x];
                for (int y = 0; y < 3; ++y)
                    for (int x = 0; x < 3; ++x)
                        wij[x + 3 * y] = wij[x + 3 * y] + rij[x] * df[y] - dfbij * wijb[x + 3 * y] -
                                         dfbji * wjib[x + 3 * y];
                for (int x = 0; x < 3; ++x) {
                    fi[x] += -(dfbij * dbidi[x] + dfbji * dbjdi[x]);
                    fj[x] += -(dfbij * dbidj[x] + dfbji * dbjdj[x]);
                }
                for (int64_t ik = is; ik < ie; ++ik) {
                    if (ik == ij) continue;
                    int ikc = (int)(ik - is);
                    int k = bd[ik].j;
                    for (int x = 0; x < 3; ++x) fl[3 * k + x] += -dfbij * dbidk[ikc][x];
                }
                for (int64_t jl = js; jl < je; ++jl) {
                    const rbond *bl = &bd[jl];
                    int jlc = (int)(jl - js);
                    int l = bl->j;
                    if (l == i && bij_->dc[0] + bl->dc[0] == 0 && bij_->dc[1] + bl->dc[1] == 0 &&
                        bij_->dc[2] + bl->dc[2] == 0)
                        continue;
                    for (int x = 0; x < 3; ++x) fl[3 * l + x] += -dfbji * dbjdl[jlc][x];
                }
                for (int k = 0; k < 9; ++k) wl[k] += wij[k];
                /* :2762-2763 */
                if (wpot_per_bond)
                    for (int k = 0; k < 9; ++k) wpot_per_bond[9 * bij_->jn + k] += wij[k];
                if (wal) {
                    for (int k = 0; k < 9; ++k) {
                        wal[9 * i + k] += wij[k] / 2;
                        wal[9 * j + k] += wij[k] / 2;
                    }
                }
                for (int x = 0; x < 3; ++x) fl[3 * j + x] += fj[x];
            }
            for (int x = 0; x < 3; ++x) fl[3 * i + x] += fi[x];
        }
        free(dbidk);
        free(dbjdl);
    }
    }

    if (!err) {
        double e = 0.0;
        for (int t = 0; t < nth; ++t) {
            for (int i = 0; i < nat; ++i) {
                for (int x = 0; x < 3; ++x) f[3 * i + x] += fbuf[(size_t)t * 3 * nat + 3 * i + x];
                if (epot_per_at) epot_per_at[i] += pebuf[(size_t)t * nat + i];
                if (wpot_per_at)
                    for (int k = 0; k < 9; ++k) wpot_per_at[9 * i + k] += wabuf[(size_t)t * 9 * nat + 9 * i + k];
            }
            for (int k = 0; k < 9; ++k) wpot[k] += wbuf[9 * t + k];
        }
        for (int i = 0; i < nat; ++i)
            for (int t = 0; t < nth; ++t) e += pebuf[(size_t)t * nat + i];
        *epot += e;
    }
    free(fbuf); free(pebuf); free(wbuf); free(wabuf);
    free(ktyp); free(bseed); free(bd); free(nn); free(nconjit);
    return err;
}
