/*
 * CPU oracle for the potential-evaluation path: a plain-C FP64 restatement
 * of the reference algorithms.  TEST INFRASTRUCTURE ONLY -- the product path
 * (atomistica_amd) never links or calls anything here; tests, the smoke check
 * and the bench's cpu_baseline leg use it as the checker.
 *
 * Conventions shared by every routine:
 *   - cell[9] is row-major with cell[3*k + x] = component x of lattice vector
 *     k (ASE order).  This is the memory image of the reference's Fortran
 *     Abox(3,3) with Abox(:,k) = lattice vector k
 *     (src/python/f90/particles_wrap.f90:147-166).
 *   - r[3*i + x] are unwrapped positions (the reference's r_non_cyc).
 *   - neighbour lists are CSR: offsets[nat+1], nbr[j], dc[3*slot]; the image
 *     convention follows src/macros.inc:76: dr_ij = r_i - r_j + Abox.dc.
 *
 * Compiled with -ffp-contract=off so that sums and products round exactly as
 * written, in the order the reference's Fortran writes them.
 */
#ifndef ORC_H
#define ORC_H
#include <stdint.h>

#define ORC_OK 0
#define ORC_ERR_OVERFLOW 1
#define ORC_ERR_DOMAIN 2
#define ORC_ERR_SPLINE 3
#define ORC_ERR_ARG 4
#define ORC_ERR_SINGULAR 5
#define ORC_ERR_NEB 6

/* ---- cell / neighbour list ---- */
int orc_inverse_cell(const double *cell, double *bbox);
int orc_nl_geometry(const double *cell, const int *pbc, double cutoff,
                    double bin_size, int *n_cells, int *n_stencil);
int orc_nl_build(int nat, const double *r, const double *cell, const int *pbc,
                 double cutoff, double bin_size, int64_t cap,
                 int64_t *offsets, int *nbr, int *dc);
int orc_nl_build_mt(int nat, const double *r, const double *cell, const int *pbc,
                    double cutoff, double bin_size, int64_t cap,
                    int64_t *offsets, int *nbr, int *dc, int nthreads);
int orc_nl_brute(int nat, const double *r, const double *cell, const int *pbc,
                 double cutoff, int64_t cap, int64_t *offsets, int *nbr,
                 int *dc);

/* ---- Tersoff ---- */
typedef struct {
    int nel;
    int Z[3];
    double A[6], B[6], xi[6], lambda[6], mu[6], omega[6], mubo[6];
    int m[6];
    double beta[3], n[3], c[3], d[3], h[3];
    double r1[9], r2[9];
    /* fam = 1: the Kumagai forms (kumagai_func.f90) with A, B, lambda
     * (= lambda1), mu (= lambda2), h, r1, r2 from above */
    int fam;
    double kc1[3], kc2[3], kc3[3], kc4[3], kc5[3], keta[3], kdelta[3];
    double kalpha[6];
    int kbeta[6];
    /* fam = 2: the Brenner forms (brenner_func.f90), every parameter per
     * element pair; the h-term takes mu, m from mubo, m above */
    double bD0[9], br0[9], bS[9], bbeta[9], bgam[9], bc[9], bd[9], bh[9], bn[9];
    /* scr = 1: the screened variants (TersoffScr, KumagaiScr, BrennerScr):
     * exp_cutoff on [r1, r2] (inner), [or1, or2] (pair), [bor1, bor2] (bond
     * order) and the screening of bonds by third atoms with Cmin, Cmax */
    int scr;
    double or1[9], or2[9], bor1[9], bor2[9], Cmin[9], Cmax[9];   /* 9: ordered pairs (Juslin) */
    /* fam = 3: the Juslin forms (juslin_func.f90): the Brenner V_A, V_R, g,
     * b above over ordered element pairs j + i nel (PAIR_INDEX_NS, 9 pairs,
     * r1 r2 too), and h = omega exp((alpha dr)^m) per ordered triplet
     * k + nel (j + nel i) (TRIPLET_INDEX_NS) of centre i */
    double jal[27], jom[27];
    int jm[27];
} orc_tersoff_params;

int orc_tersoff(const orc_tersoff_params *p, int nat, const double *r,
                const int *el, const double *cell, const int64_t *offsets,
                const int *nbr, const int *dc, const int *mask,
                double *epot, double *f, double *wpot, double *epot_per_at,
                double *wpot_per_at, double *epot_per_bond, double *f_per_bond,
                double *wpot_per_bond, int nthreads);

/* ---- splines / EAM ---- */
typedef struct {
    int n;
    double x0, dx, cut;
    double *y, *d2y, *c1, *c2, *c3, *dc1, *dc2, *dc3;
} orc_spline;

int orc_spline_init(int n, double x0, double dx, const double *y,
                    double *coef /* 8*n doubles: y,d2y,c1,c2,c3,dc1,dc2,dc3 */);
int orc_eam(int nel, int nat, const double *r, const int *el,
            const double *cell, const int64_t *offsets, const int *nbr,
            const int *dc, const int *mask, double cutoff,
            const int *n_F, const double *dF, const double *const *coef_F,
            const int *n_rho, const double *drho,
            const double *const *coef_rho, const int *n_phi,
            const double *dphi, const double *const *coef_phi,
            double *epot, double *f, double *wpot, double *epot_per_at,
            double *wpot_per_at, int funcfl, int nthreads);

/* ---- Ewald ---- */
int orc_ewald_real(int nat, const double *r, const double *q,
                   const double *cell, const int64_t *offsets, const int *nbr,
                   const int *dc, double cutoff, double beta, double *epot,
                   double *f, double *wpot, double *phi, int nthreads);
int orc_ewald_kspace(int nat, const double *r, const double *q,
                     const double *cell, double beta, double kmax,
                     double *epot, double *f, double *wpot, double *phi,
                     int nthreads);

/* ---- REBO2 (C, H; no screening; optional dihedral term) ---- */
typedef struct orc_rebo2_s orc_rebo2_t;
orc_rebo2_t *orc_rebo2_new(int zero_tables, int with_dihedral);
void orc_rebo2_free(orc_rebo2_t *p);
int orc_rebo2(const orc_rebo2_t *p, int nat, const double *r, const int *Z,
              const double *cell, const int64_t *offsets, const int *nbr, const int *dc,
              double *epot, double *f, double *wpot, double *epot_per_at,
              double *wpot_per_at, double *epot_per_bond, double *f_per_bond,
              double *wpot_per_bond, int nthreads);
int orc_rebo2_eval_g(const orc_rebo2_t *p, int ktyp, double costh, double n, double *out3);
int orc_rebo2_eval_table(const orc_rebo2_t *p, int which, double x, double y, double z,
                         double *out4);
int orc_rebo2_default_table(int which, double *out);

#endif
