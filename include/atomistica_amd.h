/*
 * atomistica_amd -- C-ABI of the MI355X (gfx950) potential-evaluation path.
 *
 * The ABI mirrors the per-class function table that the reference generates
 * for its Python and LAMMPS front ends, so a host that drives Atomistica
 * through that table can drive this library the same way:
 *
 *   reference potential_class_t        src/python/c/factory.template.h:82-98
 *   reference LAMMPS class table       src/lammps/factory.template.h:37-49
 *   reference Coulomb class table      src/python/gen_factory.py:570-615
 *
 * Conventions (identical to the reference's Fortran memory images):
 *   - positions r[3*i + x], forces f[3*i + x]                    (r(3,nat), f(3,nat))
 *   - cell[3*k + x] = component x of lattice vector k            (Abox(3,3))
 *   - wpot[a + 3*b] = W(a,b); wpot_per_at[9*i + a + 3*b]         (wpot(3,3), wpot_per_at(3,3,nat))
 *   - neighbour images: dr_ij = r_i - r_j + Abox.dc              (src/macros.inc:76)
 *   - atom and neighbour indices are 0-based in every array here.
 *   - f, wpot and per-atom outputs are accumulated (+=); *epot is set to
 *     the energy of the call (the generated wrapper zeroes it first,
 *     src/python/gen_factory.py:283).
 *   - every call returns AGPU_OK (0) or an error code; the message of the
 *     last error on the calling thread is returned by agpu_get_error().
 *   - array arguments are device (HBM) pointers unless the call's flags
 *     contain AGPU_HOST; then they are host pointers staged through HBM.
 *   - calls are synchronous on the library stream (agpu_set_stream) and not
 *     re-entrant on one handle, like the reference (src/support/tls.f90:45-49).
 */
#ifndef ATOMISTICA_AMD_H
#define ATOMISTICA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes (reference error cases: error.inc:58-69) ---- */
#define AGPU_OK 0
#define AGPU_ERR_ARG 1            /* invalid argument / shape */
#define AGPU_ERR_HIP 2            /* HIP runtime failure */
#define AGPU_ERR_NL_OVERFLOW 3    /* python_neighbors.f90:713-718 */
#define AGPU_ERR_DOMAIN 4         /* python_neighbors.f90:944-948; image shift out of range */
#define AGPU_ERR_SPLINE 5         /* simple_spline.f90:426-429 */
#define AGPU_ERR_UNSUPPORTED 6    /* optional argument not supported: gen_factory.py:303-304 */
#define AGPU_ERR_ELEMENT 7        /* default_bind_to_func.f90:83 */
#define AGPU_ERR_PBC 8            /* pme.f90:185-187 */
#define AGPU_ERR_IO 9             /* table file unreadable */
#define AGPU_ERR_STATE 10         /* handle not initialised / not bound */
#define AGPU_ERR_NEB 11           /* bond buffer exhausted: bop_kernel.f90:672-678 */

/* ---- flags ---- */
#define AGPU_HOST 1               /* array arguments are host pointers */

/* ---- precision modes ---- */
#define AGPU_PREC_FP64 0          /* FP64 arithmetic and accumulation */
#define AGPU_PREC_MIXED 1         /* FP32 pair/triplet arithmetic, FP64 accumulation */

/* ---- potential kinds ---- */
#define AGPU_KIND_TERSOFF 1
#define AGPU_KIND_EAM_ALLOY 2
#define AGPU_KIND_EWALD 3
#define AGPU_KIND_REBO2 4
#define AGPU_KIND_KUMAGAI 5
#define AGPU_KIND_BRENNER 6
#define AGPU_KIND_TERSOFF_SCR 7   /* tersoff_scr.f90, classname TersoffScr */
#define AGPU_KIND_KUMAGAI_SCR 8   /* kumagai_scr.f90, classname KumagaiScr */
#define AGPU_KIND_BRENNER_SCR 9   /* brenner_scr.f90, classname BrennerScr */
#define AGPU_KIND_PME 10          /* pme.f90, classname PME */
#define AGPU_KIND_JUSLIN 11       /* juslin.f90, classname Juslin */
#define AGPU_KIND_JUSLIN_SCR 12   /* juslin_scr.f90, classname JuslinScr */
#define AGPU_KIND_EAM 13          /* tabulated_eam.f90, classname TabulatedEAM */

/* agpu_rebo2_create flags */
#define AGPU_REBO2_ZERO_TABLES 1  /* rebo2 zero_tables = .true. (rebo2_module.f90:30-42) */
#define AGPU_REBO2_DIHEDRAL 2     /* with_dihedral = .true. (rebo2_type.f90, default .false.) */

typedef struct agpu_particles agpu_particles;
typedef struct agpu_neighbors agpu_neighbors;
typedef struct agpu_potential agpu_potential;
typedef struct agpu_md agpu_md;

/* Internal argument pack (not used across the ABI). */
typedef struct agpu_eval_args {
    const double *q;
    double *epot;
    double *f;
    double *wpot;
    const int *mask;
    double *epot_per_at;
    double *wpot_per_at;
    double *epot_per_bond;       /* one value per list entry (CSR order) */
    double *f_per_bond;          /* 3 per list entry */
    double *wpot_per_bond;       /* 3x3 per list entry, W(a, b) at a + 3 b */
    int flags;
} agpu_eval_args;

/* ======================================================================
 * library / device
 * ====================================================================== */
const char *agpu_version(void);
/* Copies the last error message of this thread into buf (NUL-terminated).
 * Replaces get_full_error_string (src/lammps/lammps_particles.f90:404-425)
 * and error_to_py (src/python/c/py_f.c:36). */
int agpu_get_error(char *buf, int len);
int agpu_device_count(int *n);
int agpu_set_device(int device);
/* Use an existing hipStream_t (passed as void*) for every launch; NULL
 * selects the library's own stream. */
int agpu_set_stream(void *stream);
int agpu_get_stream(void **stream);
int agpu_synchronize(void);
int agpu_malloc(void **ptr, size_t bytes);
int agpu_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int agpu_memcpy(void *dst, const void *src, size_t bytes, int kind);
/* Per-kernel HIP-event timing on the launch stream (0 = off). */
int agpu_profile_enable(int on);
int agpu_profile_reset(void);
/* Sum of event-measured durations (ms) and launch count of one kernel. */
int agpu_profile_query(const char *kernel, double *total_ms, int *count);
/* Comma-separated names of every kernel timed since the last reset. */
int agpu_profile_list(char *buf, int len);

/* ======================================================================
 * particles -- python_particles.f90:84-170 (particles_t),
 *              lammps_particles.f90:273-298 (set_pointers)
 * ====================================================================== */
int agpu_particles_create(agpu_particles **p);
int agpu_particles_destroy(agpu_particles *p);
/* particles_allocate: device storage for nat atoms */
int agpu_particles_allocate(agpu_particles *p, int nat);
/* particles_set_cell (python_particles.f90:286-346): cell rows = lattice
 * vectors, pbc[3] nonzero = periodic.  Bbox = cell^-1 by partial-pivot LU. */
int agpu_particles_set_cell(agpu_particles *p, const double *cell, const int *pbc);
/* Z[nat] atomic numbers (host or device per flags) -- update_elements
 * (python_particles.f90:617-658) */
int agpu_particles_set_atomic_numbers(agpu_particles *p, const int *Z, int flags);
/* copies r(3,nat) and marks positions changed (I_changed_positions) */
int agpu_particles_set_positions(agpu_particles *p, const double *r, int flags);
/* aliases caller-owned device positions r(3,nat) (LAMMPS-style
 * particles_set_pointers); the caller calls agpu_particles_changed_positions
 * after modifying them in place */
int agpu_particles_set_position_pointer(agpu_particles *p, int nat, double *r_device);
int agpu_particles_changed_positions(agpu_particles *p);
int agpu_particles_get_nat(agpu_particles *p, int *nat);

/* ======================================================================
 * neighbours -- python_neighbors.f90:51-139 (neighbors_t)
 * ====================================================================== */
/* Neighbors(avgn): capacity nat*avgn entries (python_neighbors.f90:499) */
int agpu_neighbors_create(agpu_neighbors **nl, int avgn);
int agpu_neighbors_destroy(agpu_neighbors *nl);
/* request_interaction_range (python_neighbors.f90:381-403): keeps the max */
int agpu_neighbors_request_interaction_range(agpu_neighbors *nl, double cutoff);
/* bin_size > 0 overrides the bin edge (default: the cutoff, :518) */
int agpu_neighbors_set_bin_size(agpu_neighbors *nl, double bin_size);
/* neighbors_update (python_neighbors.f90:430-456): rebuild when positions
 * or cell changed since the last build (verlet shell 0) */
int agpu_neighbors_update(agpu_neighbors *nl, agpu_particles *p);
int agpu_neighbors_get_size(agpu_neighbors *nl, int64_t *total, int *maxrow);
/* CSR export: offsets[nat+1] (int64), neighbours[total] (0-based j),
 * dc[3*total] image shifts; row order is the reference's fill order
 * (stencil x-major, then ascending j within a bin). */
int agpu_neighbors_get(agpu_neighbors *nl, int64_t *offsets, int *neighbors, int *dc,
                       int flags);

/* ======================================================================
 * potentials -- init / bind_to / energy_and_forces
 * ====================================================================== */

/* Tersoff parameter database (BOP_DB_TYPE, tersoff_params.f90:32-78).
 * Pair-indexed arrays use the symmetric pair index of macros.inc:123
 * (0-based): for nel elements, pair(i,j) = min(i+j*nel, j+i*nel)
 * - min(i(i+1)/2, j(j+1)/2). */
typedef struct agpu_tersoff_params {
    int nel;
    int Z[3];                    /* atomic numbers of the database elements */
    double A[6], B[6], xi[6], lambda[6], mu[6], omega[6], mubo[6];
    int m[6];
    double beta[3], n[3], c[3], d[3], h[3];
    double r1[6], r2[6];
} agpu_tersoff_params;

/* Tersoff: tersoff_init (tersoff_module.f90:30-78).  params == NULL selects
 * Tersoff_PRB_39_5566_SiC (tersoff_params.f90:99-138). */
int agpu_tersoff_create(agpu_potential **pot, const agpu_tersoff_params *params);
int agpu_tersoff_set_params(agpu_potential *pot, const agpu_tersoff_params *params);

/* Kumagai-Izumi-Hara-Sakai bond-order potential (one element), the
 * parameters of kumagai_params.f90:34-94 with the pair/element arrays reduced
 * to one entry.  INIT_FUNC (kumagai_module.f90:30-78) over the generic bond-order
 * kernel (bop_kernel.f90:563-1529) with the forms of kumagai_func.f90:32-221:
 * V_R = A exp(-lambda1 r), V_A = -B exp(-lambda2 r),
 * g = c1 + t c2 t/(c3 + t^2) (1 + c4 exp(-c5 t^2)), t = h - cos,
 * b = (1 + zeta^eta)^(-delta), h(dr) = exp(alpha dr^beta), trig_off cut-off
 * on [r1, r2].  params == NULL selects Kumagai_CompMaterSci_39_457_Si
 * (kumagai_params.f90:97-132). */
typedef struct agpu_kumagai_params {
    int Z;
    double A, B, lambda1, lambda2, eta, delta, alpha;
    int beta;
    double c1, c2, c3, c4, c5, h, r1, r2;
} agpu_kumagai_params;
int agpu_kumagai_create(agpu_potential **pot, const agpu_kumagai_params *params);

/* Brenner-form (Albe/Erhart) bond-order potential, up to three elements,
 * every parameter per element pair (PAIR_INDEX order as Tersoff):
 * brenner_params.f90:34-71 without screening.  INIT_FUNC / BIND_TO_FUNC
 * (brenner_module.f90:27-152, 267-296) over the generic bond-order kernel
 * with the forms of brenner_func.f90:27-196:
 * V_R = D0/(S-1) exp(-beta sqrt(2S) (r - r0)),
 * V_A = -S D0/(S-1) exp(-beta sqrt(2/S) (r - r0)),
 * g = gamma (1 + c^2/d^2 - c^2/(d^2 + (h + cos)^2)), b = (1 + zeta^n)^(-1/2n),
 * h-term exp((2 mu dr)^m), trig_off cut-off on [r1, r2].
 * params == NULL selects Erhart_PRB_71_035211_SiC. */
typedef struct agpu_brenner_params {
    int nel;
    int Z[3];
    double D0[6], r0[6], S[6], beta[6], gamma[6], c[6], d[6], h[6], mu[6], n[6];
    int m[6];
    double r1[6], r2[6];
} agpu_brenner_params;
int agpu_brenner_create(agpu_potential **pot, const agpu_brenner_params *params);

/* Juslin (juslin.f90, juslin_func.f90): the Brenner forms of V_R, V_A, g and
 * b over ORDERED element pairs q = j + i nel (PAIR_INDEX_NS,
 * macros.inc:139; 9 entries for 3 elements), and the h-term
 * omega exp((alpha dr)^m) per ordered triplet t = k + nel (j + nel i) of the
 * centre i (TRIPLET_INDEX_NS, macros.inc:146).  A pair with r0 < 0 takes the
 * reversed pair's parameters (juslin_module.f90:273-305).  NULL selects
 * Juslin_J_Appl_Phys_98_123520_WCH (juslin_params.f90:74-112). */
typedef struct {
    int nel;
    int Z[3];
    double D0[9], r0[9], S[9], beta[9], gamma[9], c[9], d[9], h[9], n[9];
    double alpha[27], omega[27];
    int m[27];
    double r1[9], r2[9];
} agpu_juslin_params;
int agpu_juslin_create(agpu_potential **pot, const agpu_juslin_params *params);

/* Screened variants: the same bond-order kernels with SCREENING defined
 * (bop_kernel.f90:563-1066 loop1 screening, :1448-1492 and :1531-1609 its
 * forces).  r1, r2 of the family's parameter block become the inner cutoff;
 * every cutoff is the exp_cutoff form (cutoff.f90:232-293).  Per element
 * pair (PAIR_INDEX order): outer (pair-term) cutoff [or1, or2], bond-order
 * cutoff [bor1, bor2], screening parameters Cmin < Cmax
 * (tersoff_params.f90:82-89).  The neighbour range requested is
 * sqrt(C_dr_cut) * max(r2, or2, bor2), C_dr_cut = Cmax^2 / (4 (Cmax - 1))
 * for Cmax > 2 (default_bind_to_func.f90:44-150).  NULL parameter blocks
 * select the reference's screened defaults: Tersoff_PRB_39_5566_SiC,
 * Kumagai_CompMaterSci_39_457_Si and Erhart_PRB_71_035211_SiC with
 * SCREENING (tersoff_params.f90:99-138, kumagai_params.f90:97-132,
 * brenner_params.f90:78-113).  Rows of more than 256 neighbours within the
 * screening range return AGPU_ERR_NEB. */
/* Indexed by the family's element pair: the symmetric pair index for
 * Tersoff, Kumagai and Brenner (6 entries used), the ordered pair j + i*nel
 * for Juslin (9). */
typedef struct {
    double or1[9], or2[9], bor1[9], bor2[9], Cmin[9], Cmax[9];
} agpu_screening_params;
int agpu_tersoff_scr_create(agpu_potential **pot, const agpu_tersoff_params *params,
                            const agpu_screening_params *scr);
int agpu_kumagai_scr_create(agpu_potential **pot, const agpu_kumagai_params *params,
                            const agpu_screening_params *scr);
int agpu_brenner_scr_create(agpu_potential **pot, const agpu_brenner_params *params,
                            const agpu_screening_params *scr);
/* JuslinScr: juslin_scr_init (juslin_scr.f90, the shared bop_kernel.f90 with
 * SCREENING); params/scr NULL = the screened Juslin W-C-H set of
 * juslin_params.f90. */
int agpu_juslin_scr_create(agpu_potential **pot, const agpu_juslin_params *params,
                           const agpu_screening_params *scr);

/* TabulatedAlloyEAM: tabulated_alloy_eam_init (tabulated_alloy_eam.f90:147-259)
 * reading a setfl file; elements = filter string ("*" or NULL = all). */
int agpu_eam_alloy_create(agpu_potential **pot, const char *fn, const char *elements);

/* TabulatedEAM: tabulated_eam_init (tabulated_eam.f90:141-214) reading a
 * single-element funcfl file (F(rho), effective charge Z(r), rho(r)); the
 * kernel is tabulated_eam.f90:334-511: pair energy Z(r)^2 / r per directed
 * pair with Z scaled by sqrt(Hartree Bohr / 2) (Units.f90 values).  Like the
 * reference class (features: per_at) it takes no mask and no per-bond
 * arrays.  Arguments beyond a table extend its last interval (the
 * reference's inline spline macros do not check their range there). */
int agpu_eam_create(agpu_potential **pot, const char *fn, const char *elements);

/* Ewald Coulomb: real-space erfc sum and self term as pme.f90:332-484, with
 * the reciprocal space as a direct FP64 structure-factor sum over
 * |k| <= kmax.  beta <= 0 selects the reference rule
 * beta^2 = ln(10)*12/cutoff^2 (pme.f90:192); kmax <= 0 selects
 * exp(-kmax^2/(4 beta^2)) = 1e-13. */
int agpu_ewald_create(agpu_potential **pot, double cutoff, double beta, double kmax);

/* PME: pme_init(cutoff, order, grid) (pme.f90:129-147) -- the same real space
 * and self term with beta^2 = ln(10)*12/cutoff^2, and the smooth PME
 * reciprocal sum of pme_kernel.f90:251-752 (cardinal B-splines of the given
 * order, FFT grid K1 x K2 x K3).  cutoff <= 0, order <= 0 and grid NULL
 * select the defaults 10, 8, 64^3.  order 3 .. 12, every K >= order.  As in
 * the reference, scaled fractional coordinates are taken with the columns of
 * Bbox, exact for orthorhombic cells. */
int agpu_pme_create(agpu_potential **pot, double cutoff, int order, const int *grid);

/* REBO2 (Brenner et al. 2002) for C and H without screening: rebo2_init and
 * rebo2_db_init_with_parameters (rebo2_module.f90:30-42, rebo2_db.f90:81-303)
 * with the default tables of rebo2_default_tables.f90; energy_and_forces is
 * rebo2_kernel (bop_kernel_rebo2.f90), including the torsion term when
 * AGPU_REBO2_DIHEDRAL is set.  No mask (the reference kernel takes none).
 * AGPU_PREC_MIXED keeps the bond records, bond orders, radial pair terms and
 * the torsion in FP64 and runs the per-neighbour angular and conjugation
 * loops in FP32 (force RMS within 1e-5 eV/A of FP64 on a-C:H). */
int agpu_rebo2_create(agpu_potential **pot, int flags);

int agpu_potential_destroy(agpu_potential *pot);
int agpu_potential_kind(agpu_potential *pot, int *kind);
int agpu_potential_set_precision(agpu_potential *pot, int precision);
int agpu_potential_get_cutoff(agpu_potential *pot, double *cutoff);

/* bind_to (default_bind_to_func.f90:25-146, tabulated_alloy_eam.f90:297-350,
 * pme.f90:174-207): element map, interaction range request on nl. */
int agpu_potential_bind_to(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl);

/* energy_and_forces: the argument list of potential_class_t.energy_and_forces
 * (factory.template.h:92-94).  q is used by Coulomb classes only.
 *
 * Per-bond arrays (the bond-order classes and REBO2, whose reference kernels
 * fill them: bop_kernel.f90:1370-1373,1406-1408,1511-1512,1593-1594 and
 * bop_kernel_rebo2.f90:2546-2547,2571-2572,2762-2763,2855-2856; the other
 * classes return AGPU_ERR_UNSUPPORTED for a non-NULL one, as
 * gen_factory.py:303-304 does) hold one entry per neighbour-list entry, in
 * the CSR order agpu_neighbors_get exports (row i = offsets[i] ..
 * offsets[i+1]-1).  That is the packed order in which the reference's Python
 * layer returns epot_per_bond and wpot_per_bond (potential.c:587-611,
 * neighbors_wrap.f90:498-547); f_per_bond is packed the same way here.  The
 * arrays need room for the list total after the update
 * (agpu_neighbors_get_size; the reference sizes them nat*avgn,
 * default_compute_func.f90:37-44, which bounds it) and are accumulated (+=):
 *   epot_per_bond[e]      += 1/2 fc (V_R + b V_A) of bond e (bond-order
 *                            classes) or the whole bond energy fc (V_R +
 *                            bbar V_A) on the half-pair entry (REBO2);
 *   f_per_bond[3e + x]    += the pair-term force dE/dr r_hat (no bond-order
 *                            part), as the reference;
 *   wpot_per_bond[9e + q] += the virial of every term evaluated for that
 *                            bond (pair, bond order, screening, REBO2
 *                            coordination, conjugation and torsion). */
int agpu_potential_energy_and_forces(agpu_potential *pot, agpu_particles *p,
                                     agpu_neighbors *nl, const double *q, double *epot,
                                     double *f, double *wpot, const int *mask,
                                     double *epot_per_at, double *epot_per_bond,
                                     double *f_per_bond, double *wpot_per_at,
                                     double *wpot_per_bond, int flags);

/* Allocation-free energy_and_forces.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * unscreened bond-order classes (Tersoff, Kumagai, Brenner, Juslin) only;
 * device pointers only (epot set, f and wpot accumulated, mask nullable).
 * Nothing is allocated: the potential must be bound to p and nl with the
 * current atomic numbers, and the list current for the positions
 * (agpu_neighbors_update called beforehand), else AGPU_ERR_STATE.  The
 * per-call scratch is a fixed 64 MiB device array per GPU (24 bytes per list
 * entry plus about 30 per atom; larger systems return AGPU_ERR_ARG), held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU race.  Rows of more than 64 neighbours return
 * AGPU_ERR_UNSUPPORTED.  Results equal those of
 * agpu_potential_energy_and_forces on its row path (set_launch variant 0,
 * the default). */
int agpu_potential_energy_and_forces_noalloc(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                                             double *epot, double *f, double *wpot, const int *mask);

/* Launch-shape autotuner (no counterpart in the reference, whose kernels are
 * CPU loops): times the interchangeable kernel variants on the bound
 * configuration with HIP events, reps calls each, and keeps the fastest for
 * later calls on pot and nl -- the neighbour-list pair search (thread per
 * atom or wavefront per bin block; identical rows), for single-element
 * tabulated EAM, lanes per atom x row entries per lane, for the
 * unscreened bond-order classes, the row path (a lane per row entry) or the
 * bond-compacting path, and for REBO2, the lanes per half pair of the
 * torsion pass (4, 8 or 16).  A text report of the timings and choices is
 * written to report (len bytes, may be NULL). */
int agpu_potential_autotune(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int reps,
                            char *report, int len);

/* Pins the launch variants instead of timing them (to repeat an earlier
 * autotuner choice): pair_mode 0 = from the bin occupancy, 1 = thread per
 * atom, 2 = wavefront per bin block; variant -1 keeps the current kernel
 * variant, else selects 0 <= variant < *nvariants, the number of variants
 * this binding offers (single-element EAM: the lanes x entries shapes;
 * unscreened bond-order classes: 0 = row path, 1 = bond-compacting path;
 * REBO2: 0, 1, 2 = 4, 8, 16 lanes per half pair in the torsion pass;
 * 0 for the other classes, variant 0 is then the only accepted value). */
int agpu_potential_set_launch(agpu_potential *pot, agpu_neighbors *nl, int pair_mode, int variant,
                              int *nvariants);

/* Coulomb potential phi_i = dE/dq_i (pme_potential, pme.f90:215-321);
 * phi is accumulated (+=). */
int agpu_coulomb_potential(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                           const double *q, double *phi, int flags);

/* ======================================================================
 * batched Tersoff for parameter fitting (no counterpart in the reference,
 * which evaluates one structure per call): many small fully periodic
 * structures of one element evaluated together, with the gradient of
 * energies and forces with respect to the 11 parameters
 * theta = (A, B, lambda, mu, beta, n, c, d, h, R, S) of the single-element
 * Tersoff form of bop_kernel.f90 / tersoff_func.f90 (xi = omega = 1, no
 * h-term).  The batch builds its own neighbour list (every periodic image
 * within S, all structures in one pass).
 * ====================================================================== */
typedef struct agpu_batch agpu_batch;
int agpu_batch_create(agpu_batch **b);
int agpu_batch_destroy(agpu_batch *b);
/* atom_offsets[nstruct+1] (host, atom_offsets[0] = 0), cells[9*nstruct]
 * (host, cell[9*s + 3*k + x] = component x of lattice vector k) */
int agpu_batch_set_structures(agpu_batch *b, int nstruct, const int *atom_offsets,
                              const double *cells);
/* r[3*ntot] concatenated positions (device, or host with AGPU_HOST) */
int agpu_batch_set_positions(agpu_batch *b, const double *r, int flags);
int agpu_batch_get_size(agpu_batch *b, int *nstruct, int *ntot, int64_t *npairs);
/* e_struct[nstruct] is set, f[3*ntot] accumulated; theta[11] on the host;
 * precision AGPU_PREC_FP64 or AGPU_PREC_MIXED (FP32 pair/triplet
 * arithmetic, FP64 displacements and accumulation) */
int agpu_tersoff_batch_forward(agpu_batch *b, const double *theta, int precision,
                               double *e_struct, double *f, int flags);
/* grad_theta[11] = sum_s grad_e[s] dE_s/dtheta + sum_i grad_f[i] . dF_i/dtheta
 * (set, FP64); grad_e or grad_f may be NULL (zero) */
int agpu_tersoff_batch_backward(agpu_batch *b, const double *theta, const double *grad_e,
                                const double *grad_f, double *grad_theta, int flags);

/* Batched Ewald Coulomb sum over the structures of b (e^2/Angstrom): real
 * space + self term as pme_energy_and_forces / pme_potential
 * (src/potentials/coulomb/pme.f90:215-321, 386-437, 467-472), reciprocal space
 * as the direct S(k) sum over the half space |k| <= kmax in FP64 (the
 * reference uses smooth PME, pme_kernel.f90:251-752).  Per-ion charges q[ntot]
 * (|q| <= 1e-10 skipped as in the reference).  e_struct[nstruct] and
 * wpot_struct[9*nstruct] (optional) are set; f[3*ntot] and phi[ntot] (both
 * optional; phi = dE/dq) are accumulated.  precision AGPU_PREC_MIXED runs the
 * real-space erfc/exp in FP32 with FP64 distances and accumulation.  No
 * counterpart in the reference, which evaluates one structure per call. */
int agpu_ewald_batch_compute(agpu_batch *b, const double *q, double cutoff, double beta, double kmax,
                             int precision, double *e_struct, double *f, double *wpot_struct,
                             double *phi, int flags);

/* ======================================================================
 * NVE molecular dynamics around one potential, with the steps replayed as
 * a HIP graph -- the velocity-Verlet loop of the reference's MDCORE driver
 * (src/standalone/main.f90:429-531) with a list kept inside a Verlet shell
 * (the reference's Python path never rebuilds with a shell,
 * python_particles.f90:125-128, so the shell logic is this library's own).
 * Units are the potential's: with eV, Angstrom and amu the time unit is
 * Angstrom sqrt(amu/eV) = 10.18 fs.  No counterpart in the reference's
 * potential interface.
 * ====================================================================== */
/* masses[nat] and v0[3*nat] (optional, zero if NULL) on the host; the
 * potential must be a Tersoff-family or EAM class bound to p and nl; skin is
 * the Verlet shell added to the list cut-off (0: rebuild every step). */
int agpu_md_create(agpu_md **md, agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                   const double *masses, const double *v0, double dt, double skin);
int agpu_md_destroy(agpu_md *md);
/* nsteps steps.  graph_steps > 0 replays captured graphs of up to that many
 * steps (32, 16, 8 or 4, the longest that the displacement so far and the
 * largest speed predict to keep the list valid; a replay that still goes
 * stale is rolled back and redone step by step), giving the same trajectory
 * as graph_steps = 0, bitwise.  *epot (host, optional) = potential energy
 * after the last step. */
int agpu_md_run(agpu_md *md, int nsteps, int graph_steps, double *epot);
/* r, v, f (3*nat), wpot (9) copies, host or device per flags; counters[4] =
 * steps, list builds, graph replays, replays redone step by step. */
int agpu_md_get(agpu_md *md, double *r, double *v, double *f, double *wpot, int64_t *counters, int flags);

/* LAMMPS pair-style step (PairAtomistica::FAtomistica,
 * src/lammps/pair_style/pair_atomistica.cpp:499-561): x[3*nall] holds the
 * nlocal local atoms followed by the ghosts (host memory, LAMMPS order), Z
 * their atomic numbers.  The atoms are placed in an open box and the local
 * atoms are masked in, so only terms centred on local atoms carry energy;
 * f[3*nall] is accumulated on local atoms and ghosts (LAMMPS' reverse
 * communication returns the ghost forces), *eng += energy, virial[6] -=
 * wpot in LAMMPS Voigt order (xx yy zz xy xz yz) as the reference does.
 * eatom (nall) is optional.  The bond-order classes and EAM; REBO2 needs the
 * atom tags (agpu_lammps_compute_tagged); Ewald and PME return
 * AGPU_ERR_UNSUPPORTED. */
int agpu_lammps_compute(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                        const double *x, const int *Z, double *f, double *eng, double *virial,
                        double *eatom);

/* The same LAMMPS step over LAMMPS' own full neighbour list with ghosts
 * instead of a list built here: PairAtomistica::Atomistica_neigh
 * (src/lammps/pair_style/pair_atomistica.cpp:385-470) hands the reference
 * seed/last offsets into the LAMMPS neighbour memory; here row i (of all nall
 * local and ghost atoms) is neigh[seed[i] .. seed[i] + numneigh[i]), entries
 * are indices into x with LAMMPS' special-bond bits (the top two) masked
 * off, and no periodic images.  The list is imported into nl and kept until
 * the next call on nl. */
int agpu_lammps_compute_list(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                             const double *x, const int *Z, const int64_t *seed, const int *numneigh,
                             const int *neigh, double *f, double *eng, double *virial, double *eatom);

/* The LAMMPS step with the atom tags (atom->tag of local atoms and ghosts),
 * which REBO2 needs: it has no mask (bop_kernel_rebo2.f90 takes none) and
 * evaluates each bond once, from the local atom i for tag(j) >= tag(i)
 * (bop_kernel_rebo2.f90:1328-1329).  seed == NULL builds the list here (as
 * agpu_lammps_compute), else imports LAMMPS' list (as
 * agpu_lammps_compute_list); tag may be NULL for the other classes.  vatom
 * (6 nall, optional) receives LAMMPS' per-atom virial: -= (W_xx, W_yy, W_zz,
 * W_yx, W_zx, W_zy) of each atom's half of its terms, as SUM_VIRIAL in
 * LAMMPS mode (macros.inc:202, pair_atomistica.cpp:523-524). */
int agpu_lammps_compute_tagged(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                               const double *x, const int *Z, const int *tag, const int64_t *seed,
                               const int *numneigh, const int *neigh, double *f, double *eng, double *virial,
                               double *eatom, double *vatom);

/* Synthetic-workload utility (no counterpart in the reference): n random
 * points in the periodic cell (cell[3k+x] = component x of lattice vector k)
 * by random sequential addition with minimum distance dmin over every
 * periodic image; dmin is relaxed by 3 % after max_tries rejections in a row.
 * Deterministic for a given seed (splitmix64).  pos[3*n], host memory. */
int agpu_random_positions(int n, const double *cell, double dmin, uint64_t seed, int max_tries,
                          double *pos);

#ifdef __cplusplus
}
#endif
#endif
