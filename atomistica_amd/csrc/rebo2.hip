// Second-generation REBO (Brenner et al. 2002) for carbon and hydrogen: the
// non-screened rebo2 class of the reference (src/potentials/bop/rebo2/),
// with the optional dihedral (torsion) term of with_dihedral = .true.
//
// The kernels work on slots, the bin order of the neighbour-list builder
// (slot p holds atom order[p]): bond records, coordinations, accumulators
// and the half-pair list are indexed by slot, so the atoms a half pair
// touches lie close together in memory whatever the caller's atom order.
// Forces and per-atom properties go back to the caller's order at the end.
//
// Kernels (one stream, no host round trip inside a call):
//   k_rb_slots    elements (and LAMMPS tags) in slot order; clears the
//                 fixed-point accumulators, flags and energy.  The bond
//                 slots of slot p are its list row's entries (the list's
//                 slot-order offsets)
//   k_rb_bonds    thread per slot: compact the neighbour row to its bonds
//                 (r^2 < r2^2 of the pair type), bond record = unit vector,
//                 length, cut-off and derivative; coordination N_C, N_H
//                 (loop1 and the nn sum of bop_kernel_rebo2.f90:700-1205);
//                 the centre's width class (4, 8, 16 bonds), listed per
//                 class by k_rb_class_list after a scan
//   k_rb_conj     thread per atom: F(x_ik) of every carbon neighbour k and
//                 the sum N^conj_i (ik_loop1, bop_kernel_rebo2.f90:1231-1325)
//   k_rb_bo       thread per directed bond: zeta with P, b and dV_A b', the
//                 reverse bond's slot, the half-pair flag
//   k_rb_list     ordered list of half pairs (exclusive scan of the flags)
//   k_rb_pair     thread per half pair i->j ((dc = 0 and j > i) or dc > 0):
//                 F_ij(N_i, N_j, N_conj) and the torsion factor
//                 T_ij(N_i, N_j, N_conj), the pair energy and radial force,
//                 the neighbour-pass coefficients of both directed bonds
//                 (ij_loop, bop_kernel_rebo2.f90:1326-2779)
//   k_rb_torsion  group per half pair with T != 0: the torsion sum over the
//                 bonds k of i and l of j with its forces
//   neighbour pass, per centre (k_rb_centre, k_rb_conj_force per width
//                 class, then the owner gather k_rb_gather), or, when
//                 per-atom or per-bond virials are asked for, a thread per
//                 directed bond (k_rb_nbrs) and k_rb_finish.
// Forces on an atom from its own bonds' passes are kept per directed bond
// and summed by the owner; forces that many half pairs put on one atom
// (torsion k and l, the per-bond-thread pass) go to 64-bit fixed-point
// accumulators with integer atomics (exact and order independent).  Every
// sum therefore has a fixed order, and results are bitwise reproducible;
// total energy and virial are reduced per block, then in a fixed order.
//
// Tables (F_CC, F_CH, F_HH, T_CC tricubic, P_CC, P_CH bicubic) are built on the
// host from their node data in closed form: the tensor product of cubic
// Hermite polynomials with the tabulated values and first derivatives and
// zero cross derivatives, which is the unique solution of the corner system
// the reference solves with gaussn (src/special/table3d.f90:40-196).
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

// pair codes of the reference (Z2pair): C = 1, H = 3; CC = 1, CH = 3, HH = 6
enum { RB_C = 1, RB_H = 3, RB_CC = 1, RB_CH = 3, RB_HH = 6 };

template <typename R> struct RbDevT {
    R ccB[3], ccBeta[3];        // VA of C-C: three exponentials
    R B1[7], beta1[7];          // VA of C-H and H-H
    R A[7], Q[7], alpha[7];     // VR
    R cut_l[7], cut_h[7], cut_h2[7], cut_fac[7];
    R conear[7][7];
    R conalp, conpe, conan, conpf;
    R th1, th2, n_hi;           // g interval boundaries; upper N switch
    R g1c[3][6], g2c[3][6];     // carbon g(cos): high / low coordination
    R spgh[3][6];               // hydrogen g(cos) polynomials
    int igh[25];
    int with_dihedral;
};
using RbDev = RbDevT<double>;

// the same constants rounded to FP32 (the mixed form)
static RbDevT<float> rb_dev_float(const RbDev &D)
{
    RbDevT<float> F;
    for (int k = 0; k < 3; ++k) { F.ccB[k] = (float)D.ccB[k]; F.ccBeta[k] = (float)D.ccBeta[k]; }
    for (int k = 0; k < 7; ++k) {
        F.B1[k] = (float)D.B1[k]; F.beta1[k] = (float)D.beta1[k];
        F.A[k] = (float)D.A[k]; F.Q[k] = (float)D.Q[k]; F.alpha[k] = (float)D.alpha[k];
        F.cut_l[k] = (float)D.cut_l[k]; F.cut_h[k] = (float)D.cut_h[k];
        F.cut_h2[k] = (float)D.cut_h2[k]; F.cut_fac[k] = (float)D.cut_fac[k];
        for (int q = 0; q < 7; ++q) F.conear[k][q] = (float)D.conear[k][q];
    }
    F.conalp = (float)D.conalp; F.conpe = (float)D.conpe; F.conan = (float)D.conan; F.conpf = (float)D.conpf;
    F.th1 = (float)D.th1; F.th2 = (float)D.th2; F.n_hi = (float)D.n_hi;
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 6; ++q) {
            F.g1c[a][q] = (float)D.g1c[a][q];
            F.g2c[a][q] = (float)D.g2c[a][q];
            F.spgh[a][q] = (float)D.spgh[a][q];
        }
    for (int q = 0; q < 25; ++q) F.igh[q] = D.igh[q];
    F.with_dihedral = D.with_dihedral;
    return F;
}

// table layout in one device block: 3 tricubic tables of 4*4*9 boxes x 64
// coefficients, then 2 bicubic tables of 5*5 boxes x 16 coefficients
constexpr int RB_T3_BOXES = 4 * 4 * 9;
constexpr int RB_T3 = RB_T3_BOXES * 64;
constexpr int RB_T2 = 25 * 16;
constexpr int RB_TAB_FCC = 0, RB_TAB_FCH = RB_T3, RB_TAB_FHH = 2 * RB_T3, RB_TAB_PCC = 3 * RB_T3,
              RB_TAB_PCH = 3 * RB_T3 + RB_T2, RB_TAB_TCC = 3 * RB_T3 + 2 * RB_T2,
              RB_TAB_N = 4 * RB_T3 + 2 * RB_T2;

// fixed point: 2^-40 eV/A (or eV) per unit
constexpr double RB_FX = 1099511627776.0;
constexpr double RB_FX_INV = 1.0 / 1099511627776.0;
constexpr double RB_FX_MAX = 2097152.0;   // 2^21: contributions beyond flag an error

// bond record: unit vector r_ij / r_ij, length, cut-off and derivative, the
// conjugation F(x_ik) and dF/dx (64 bytes in FP64, 32 in FP32)
template <typename R> struct RbBondT {
    R n[3];
    R len;
    R fc, dfc, fx, dfx;
};
using RbBond = RbBondT<double>;

// per directed bond: b, dV_A b', dP/dN_C + dzeta/dN, dP/dN_H + dzeta/dN
template <typename R> struct alignas(4 * sizeof(R)) Rb4 {
    R x, y, z, w;
};

struct Rebo2State {
    int zero_tables = 0;
    int with_dihedral = 0;
    RbDev hdev{};
    std::vector<double> htab;
    DevBuf dev, tab, bond, bij, bown, bcnt, nn, nconj, facc, eacc, wacc, partial, stats, bo, rev, half, hpos,
        list, scan_tmp, bsrc, el_s, tag_s, dev32, tab32, bond32, coef, tor, wdir, fnb, fcj,
        bconj, own, tflag, tpos, tlist, cls, fpr;
    int z2db[128];
    uint64_t Z_rev = ~0ull;
    double cutoff = 0.0;
    int tor_g = 4;      // lanes per half pair of the torsion pass (4, 8 or 16)
    // centres per width class on the last per-centre pass (-1: none yet)
    int chint[3] = {-1, -1, -1};
    bool chint_pending = false;
};

// ---------------------------------------------------------------------------
// host set-up
// ---------------------------------------------------------------------------
static int solve6(double A[6][6], double b[6])
{
    for (int c = 0; c < 6; ++c) {
        int p = c;
        for (int r = c + 1; r < 6; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (A[p][c] == 0.0) return -1;
        if (p != c) {
            for (int k = 0; k < 6; ++k) std::swap(A[c][k], A[p][k]);
            std::swap(b[c], b[p]);
        }
        for (int r = c + 1; r < 6; ++r) {
            double m = A[r][c] / A[c][c];
            for (int k = c; k < 6; ++k) A[r][k] -= m * A[c][k];
            b[r] -= m * b[c];
        }
    }
    for (int c = 5; c >= 0; --c) {
        double s = b[c];
        for (int k = c + 1; k < 6; ++k) s -= A[c][k] * b[k];
        b[c] = s / A[c][c];
    }
    return 0;
}

static double hpow(double x, int p)
{
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= x;
    return r;
}

// Quintic pieces of g_C on [-1,-1/2], [-1/2,-1/3] (value, slope, curvature
// at both ends) and [-1/3, 1] (values at -1/3, 0, 1/2, 1 plus slope and
// curvature at -1/3), rebo2_db.f90:405-524.
static int g_spline(const double th[6], const double *g, const double *dg, const double *d2g,
                    const double *gtail, double c[3][6])
{
    double A[6][6], b[6];
    std::memset(A, 0, sizeof A);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 6; ++j) A[i][j] = hpow(th[2 + i], j);
    for (int j = 1; j < 6; ++j) A[4][j] = j * hpow(th[2], j - 1);
    for (int j = 2; j < 6; ++j) A[5][j] = j * (j - 1) * hpow(th[2], j - 2);
    for (int i = 0; i < 4; ++i) b[i] = gtail[2 + i];
    b[4] = dg[2];
    b[5] = d2g[2];
    if (solve6(A, b)) return -1;
    std::memcpy(c[2], b, sizeof b);
    for (int k = 0; k < 2; ++k) {
        std::memset(A, 0, sizeof A);
        for (int e = 0; e < 2; ++e) {
            double z = th[k + e];
            for (int j = 0; j < 6; ++j) A[3 * e][j] = hpow(z, j);
            for (int j = 1; j < 6; ++j) A[3 * e + 1][j] = j * hpow(z, j - 1);
            for (int j = 2; j < 6; ++j) A[3 * e + 2][j] = j * (j - 1) * hpow(z, j - 2);
        }
        b[0] = g[k]; b[1] = dg[k]; b[2] = d2g[k];
        b[3] = g[k + 1]; b[4] = dg[k + 1]; b[5] = d2g[k + 1];
        if (solve6(A, b)) return -1;
        std::memcpy(c[k], b, sizeof b);
    }
    return 0;
}

// monomial coefficients of the cubic Hermite basis on [0,1]:
// value at 0, value at 1, slope at 0, slope at 1
static const double HV[2][4] = {{1, 0, -3, 2}, {0, 0, 3, -2}};
static const double HD[2][4] = {{0, 1, -2, 1}, {0, 0, -1, 1}};

// V, Dx, Dy, Dz indexed [(ix*(ny+1)+iy)*(nz+1)+iz]; box (bx,by,bz) stored at
// ((bx*ny)+by)*nz+bz with coefficient [16 p1 + 4 p2 + p3]
static void hermite3(int nx, int ny, int nz, const double *V, const double *Dx, const double *Dy,
                     const double *Dz, double *out)
{
    for (int bx = 0; bx < nx; ++bx)
        for (int by = 0; by < ny; ++by)
            for (int bz = 0; bz < nz; ++bz) {
                double *c = out + (size_t)(((bx * ny) + by) * nz + bz) * 64;
                for (int q = 0; q < 64; ++q) c[q] = 0.0;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int e = 0; e < 2; ++e) {
                            int g = ((bx + a) * (ny + 1) + by + b) * (nz + 1) + bz + e;
                            double f = V[g], fx = Dx ? Dx[g] : 0.0, fy = Dy ? Dy[g] : 0.0,
                                   fz = Dz ? Dz[g] : 0.0;
                            for (int p1 = 0; p1 < 4; ++p1)
                                for (int p2 = 0; p2 < 4; ++p2)
                                    for (int p3 = 0; p3 < 4; ++p3) {
                                        double v = f * HV[a][p1] * HV[b][p2] * HV[e][p3] +
                                                   fx * HD[a][p1] * HV[b][p2] * HV[e][p3] +
                                                   fy * HV[a][p1] * HD[b][p2] * HV[e][p3] +
                                                   fz * HV[a][p1] * HV[b][p2] * HD[e][p3];
                                        c[16 * p1 + 4 * p2 + p3] += v;
                                    }
                        }
            }
}

static void hermite2(int nx, int ny, const double *V, double *out)
{
    for (int bx = 0; bx < nx; ++bx)
        for (int by = 0; by < ny; ++by) {
            double *c = out + (size_t)(bx * ny + by) * 16;
            for (int q = 0; q < 16; ++q) c[q] = 0.0;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    double f = V[(bx + a) * (ny + 1) + by + b];
                    for (int p1 = 0; p1 < 4; ++p1)
                        for (int p2 = 0; p2 < 4; ++p2) c[4 * p1 + p2] += f * HV[a][p1] * HV[b][p2];
                }
        }
}

// Node data of rebo2_default_tables.f90 (Brenner et al. 2002, Tables 4-9):
// F(0:4, 0:4, 0:9) at [(i*5+j)*10+k], P(0:5, 0:5) at [i*6+j].
#define T3(i, j, k) (((i) * 5 + (j)) * 10 + (k))
static void default_tables(double *Fcc, double *Fi, double *Fj, double *Fk, double *Fch,
                           double *Fhh, double *Tcc, double *Pcc, double *Pch)
{
    for (int q = 0; q < 250; ++q) Fcc[q] = Fi[q] = Fj[q] = Fk[q] = Fch[q] = Fhh[q] = Tcc[q] = 0.0;
    for (int q = 0; q < 36; ++q) Pcc[q] = Pch[q] = 0.0;
    // F_CC values (Table 4) with the linear ramps in N_conj of the reference
    Fcc[T3(1, 1, 0)] = 0.105000;
    Fcc[T3(1, 1, 1)] = -0.0041775;
    for (int k = 2; k <= 8; ++k) Fcc[T3(1, 1, k)] = -0.0160856;
    Fcc[T3(2, 2, 0)] = 0.09444957;
    Fcc[T3(2, 2, 1)] = 0.02200000;
    Fcc[T3(2, 2, 2)] = 0.03970587;
    Fcc[T3(0, 1, 0)] = 0.04338699;
    for (int k = 1; k <= 8; ++k) Fcc[T3(0, 1, k)] = 0.0099172158;
    Fcc[T3(0, 2, 0)] = 0.0493976637;
    Fcc[T3(0, 2, 1)] = -0.011942669;
    for (int k = 2; k <= 8; ++k) Fcc[T3(0, 2, k)] = Fcc[T3(0, 1, 1)];
    Fcc[T3(0, 3, 0)] = Fcc[T3(0, 3, 1)] = -0.119798935;
    for (int k = 2; k <= 8; ++k) Fcc[T3(0, 3, k)] = Fcc[T3(0, 1, 1)];
    Fcc[T3(1, 2, 0)] = 0.0096495698;
    Fcc[T3(1, 2, 1)] = 0.030;
    Fcc[T3(1, 2, 2)] = -0.0200;
    for (int k = 5; k <= 8; ++k) Fcc[T3(1, 2, k)] = -0.030133632;
    for (int k = 1; k <= 8; ++k) Fcc[T3(1, 3, k)] = -0.124836752;
    for (int k = 0; k <= 8; ++k) Fcc[T3(2, 3, k)] = -0.044709383;
    for (int k = 3; k <= 7; ++k)
        Fcc[T3(2, 2, k)] = Fcc[T3(2, 2, 2)] + (k - 2) * (Fcc[T3(2, 2, 8)] - Fcc[T3(2, 2, 2)]) / 6;
    for (int k = 3; k <= 4; ++k)
        Fcc[T3(1, 2, k)] = Fcc[T3(1, 2, 2)] + (k - 2) * (Fcc[T3(1, 2, 5)] - Fcc[T3(1, 2, 2)]) / 3;
    Fi[T3(2, 1, 0)] = -0.052500;
    for (int k = 4; k <= 8; ++k) Fi[T3(2, 1, k)] = -0.054376;
    for (int k = 1; k <= 8; ++k) Fi[T3(2, 3, k)] = 0.062418;
    for (int k = 3; k <= 7; ++k) Fk[T3(2, 2, k)] = -0.006618;
    Fk[T3(1, 1, 1)] = -0.060543;
    Fk[T3(1, 2, 3)] = Fk[T3(1, 2, 4)] = -0.020044;
    // symmetrisation F(i,j) = F(j,i), dF/dNi(i,j) = dF/dNj(j,i)
    for (int k = 0; k <= 9; ++k)
        for (int i = 0; i <= 3; ++i)
            for (int j = i + 1; j <= 3; ++j) {
                double x = Fcc[T3(i, j, k)] + Fcc[T3(j, i, k)];
                Fcc[T3(i, j, k)] = Fcc[T3(j, i, k)] = x;
                x = Fi[T3(i, j, k)] + Fj[T3(j, i, k)];
                Fi[T3(i, j, k)] = Fj[T3(j, i, k)] = x;
                x = Fi[T3(j, i, k)] + Fj[T3(i, j, k)];
                Fi[T3(j, i, k)] = Fj[T3(i, j, k)] = x;
                x = Fk[T3(i, j, k)] + Fk[T3(j, i, k)];
                Fk[T3(i, j, k)] = Fk[T3(j, i, k)] = x;
            }
    // F_CH (Table 9)
    for (int k = 4; k <= 8; ++k) Fch[T3(0, 2, k)] = -0.0090477875161288110;
    for (int k = 0; k <= 8; ++k) {
        Fch[T3(1, 3, k)] = -0.213;
        Fch[T3(1, 2, k)] = -0.25;
        Fch[T3(1, 1, k)] = -0.5;
    }
    for (int k = 0; k <= 9; ++k)
        for (int i = 0; i <= 2; ++i)
            for (int j = i + 1; j <= 3; ++j) {
                double x = Fch[T3(i, j, k)] + Fch[T3(j, i, k)];
                Fch[T3(i, j, k)] = Fch[T3(j, i, k)] = x;
            }
    // F_HH (Table 6)
    Fhh[T3(1, 1, 0)] = 0.249831916;
    // T_CC (Table 5)
    Tcc[T3(2, 2, 0)] = -0.070280085;
    for (int k = 1; k <= 8; ++k) Tcc[T3(2, 2, k)] = -0.00809675;
    // P_CC, P_CH (Table 8), indexed (N_H, N_C)
    Pcc[1 * 6 + 1] = 0.003026697473481;
    Pcc[2 * 6 + 0] = 0.007860700254745;
    Pcc[3 * 6 + 0] = 0.016125364564267;
    Pcc[1 * 6 + 2] = 0.003179530830731;
    Pcc[2 * 6 + 1] = 0.006326248241119;
    Pch[1 * 6 + 0] = 0.2093367328250380;
    Pch[2 * 6 + 0] = -0.064449615432525;
    Pch[3 * 6 + 0] = -0.303927546346162;
    Pch[0 * 6 + 1] = 0.01;
    Pch[0 * 6 + 2] = -0.1220421462782555;
    Pch[1 * 6 + 1] = -0.1251234006287090;
    Pch[2 * 6 + 1] = -0.298905245783;
    Pch[0 * 6 + 3] = -0.307584705066;
    Pch[1 * 6 + 2] = -0.3005291724067579;
}
#undef T3

static int rebo2_setup(Rebo2State *st)
{
    RbDev &D = st->hdev;
    std::memset(&D, 0, sizeof D);
    // rebo2_type.f90 defaults (Brenner et al. 2002, Tables 2, 6, 7)
    D.ccB[0] = 12388.79197798; D.ccB[1] = 17.56740646509; D.ccB[2] = 30.71493208065;
    D.ccBeta[0] = 4.7204523127; D.ccBeta[1] = 1.4332132499; D.ccBeta[2] = 1.3826912506;
    D.Q[RB_CC] = 0.3134602960833; D.A[RB_CC] = 10953.544162170; D.alpha[RB_CC] = 4.7465390606595;
    D.B1[RB_CH] = 32.3551866587; D.beta1[RB_CH] = 1.43445805925;
    D.Q[RB_CH] = 0.340775728; D.A[RB_CH] = 149.94098723; D.alpha[RB_CH] = 4.10254983;
    D.B1[RB_HH] = 29.632593; D.beta1[RB_HH] = 1.71589217;
    D.Q[RB_HH] = 0.370471487045; D.A[RB_HH] = 32.817355747; D.alpha[RB_HH] = 3.536298648;
    // the g knots and node values are single-precision literals there
    const double th[6] = {-1.0, -1.0 / 2, -1.0 / 3, 0.0, 1.0 / 2, 1.0};
    const double g1[6] = {(double)-0.01f, (double)0.05280f, (double)0.09733f,
                          (double)0.37545f, (double)2.0014f, (double)8.0f};
    const double dg1[6] = {(double)0.10400f, (double)0.17000f, (double)0.40000f, 0, 0, 0};
    const double d2g1[6] = {0.0, (double)0.37000f, (double)1.98000f, 0, 0, 0};
    const double g2[6] = {0.0, 0.0, (double)0.09733f, (double)0.271856f, (double)0.416335f, 1.0};
    if (g_spline(th, g1, dg1, d2g1, g1, D.g1c) || g_spline(th, g1, dg1, d2g1, g2, D.g2c))
        return set_error(AGPU_ERR_ARG, "REBO2: singular g spline system");
    D.th1 = th[1];
    D.th2 = th[2];
    D.n_hi = (double)3.7f;
    static const double spgh[3][6] = {
        {270.467795364007301, 1549.701314596994564, 3781.927258631323866,
         4582.337619544424228, 2721.538161662818368, 630.658598136730774},
        {16.956325544514659, -21.059084522755980, -102.394184748124742,
         -210.527926707779059, -229.759473570467513, -94.968528666251945},
        {19.065031149937783, 2.017732531534021, -2.566444502991983,
         3.291353893907436, -2.653536801884563, 0.837650930130006}};
    std::memcpy(D.spgh, spgh, sizeof spgh);
    for (int i = 0; i < 25; ++i) D.igh[i] = i < 18 ? 3 : (i < 22 ? 2 : 1);
    // rebo2_db.f90:150-200
    const double lam = 4.0, cc_re = 1.4, ch_re = 1.09, hh_re = 0.7415886997;
    D.conpe = -0.5;
    D.conan = 0.5 * D.conpe;
    D.conpf = D.conpe - 1.0;
    D.conalp = lam;
    D.conear[RB_CC][RB_CC] = 1.0;
    D.conear[RB_CC][RB_CH] = std::exp(lam * (ch_re - cc_re));
    D.conear[RB_CC][RB_HH] = std::exp(lam * (hh_re - cc_re));
    D.conear[RB_CH][RB_CC] = 1.0 / D.conear[RB_CC][RB_CH];
    D.conear[RB_CH][RB_CH] = 1.0;
    D.conear[RB_CH][RB_HH] = std::exp(lam * (hh_re - ch_re));
    D.conear[RB_HH][RB_CC] = 1.0 / D.conear[RB_CC][RB_HH];
    D.conear[RB_HH][RB_CH] = 1.0 / D.conear[RB_CH][RB_HH];
    D.conear[RB_HH][RB_HH] = 1.0;
    const int t3[3] = {RB_CC, RB_CH, RB_HH};
    const double r1[3] = {1.70, 1.30, 1.10}, r2[3] = {2.00, 1.80, 1.70};
    for (int q = 0; q < 3; ++q) {
        D.cut_l[t3[q]] = r1[q];
        D.cut_h[t3[q]] = r2[q];
        D.cut_h2[t3[q]] = r2[q] * r2[q];
        D.cut_fac[t3[q]] = M_PI / (r2[q] - r1[q]);
    }
    // tables
    D.with_dihedral = st->with_dihedral;
    std::vector<double> F(250 * 7), P(72);
    default_tables(&F[0], &F[250], &F[500], &F[750], &F[1000], &F[1250], &F[1500], &P[0], &P[36]);
    if (st->zero_tables) {
        std::fill(F.begin(), F.end(), 0.0);
        std::fill(P.begin(), P.end(), 0.0);
    }
    st->htab.assign(RB_TAB_N, 0.0);
    hermite3(4, 4, 9, &F[0], &F[250], &F[500], &F[750], &st->htab[RB_TAB_FCC]);
    hermite3(4, 4, 9, &F[1000], nullptr, nullptr, nullptr, &st->htab[RB_TAB_FCH]);
    hermite3(4, 4, 9, &F[1250], nullptr, nullptr, nullptr, &st->htab[RB_TAB_FHH]);
    hermite3(4, 4, 9, &F[1500], nullptr, nullptr, nullptr, &st->htab[RB_TAB_TCC]);
    hermite2(5, 5, &P[0], &st->htab[RB_TAB_PCC]);
    hermite2(5, 5, &P[36], &st->htab[RB_TAB_PCH]);
    AGPU_CHECK(st->dev.reserve(sizeof(RbDev)));
    AGPU_CHECK(st->tab.reserve(sizeof(double) * RB_TAB_N));
    AGPU_HIP(hipMemcpy(st->dev.ptr, &D, sizeof(RbDev), hipMemcpyHostToDevice));
    AGPU_HIP(hipMemcpy(st->tab.ptr, st->htab.data(), sizeof(double) * RB_TAB_N, hipMemcpyHostToDevice));
    // FP32 copies for the mixed form
    const RbDevT<float> D32 = rb_dev_float(D);
    const std::vector<float> tab32(st->htab.begin(), st->htab.end());
    AGPU_CHECK(st->dev32.reserve(sizeof(RbDevT<float>)));
    AGPU_CHECK(st->tab32.reserve(sizeof(float) * RB_TAB_N));
    AGPU_HIP(hipMemcpy(st->dev32.ptr, &D32, sizeof(RbDevT<float>), hipMemcpyHostToDevice));
    AGPU_HIP(hipMemcpy(st->tab32.ptr, tab32.data(), sizeof(float) * RB_TAB_N, hipMemcpyHostToDevice));
    return AGPU_OK;
}

// ---------------------------------------------------------------------------
// device functions (rebo2_func.f90)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rb_pair(int ki, int kj)
{
    return ki == RB_C ? kj : (kj == RB_C ? ki : ki + kj);
}

__device__ __forceinline__ void rb_sincos(double x, double &s, double &c) { sincos(x, &s, &c); }
__device__ __forceinline__ void rb_sincos(float x, float &s, float &c) { sincosf(x, &s, &c); }

template <typename R>
__device__ __forceinline__ void rb_fconj(R x, R &fx, R &dfx)
{
    if (x <= R(2.0)) { fx = R(1.0); dfx = R(0.0); }
    else if (x >= R(3.0)) { fx = R(0.0); dfx = R(0.0); }
    else {
        R s, c;
        rb_sincos(R(M_PI) * (x - R(2.0)), s, c);
        fx = R(0.5) * (R(1.0) + c);
        dfx = -R(0.5) * R(M_PI) * s;
    }
}

template <typename R>
__device__ __forceinline__ void rb_fcin(const RbDevT<R> &D, int t, R r, R &v, R &dv)
{
    const R l = D.cut_l[t], h = D.cut_h[t];
    if (r >= h) { v = R(0.0); dv = R(0.0); }
    else if (r <= l) { v = R(1.0); dv = R(0.0); }
    else {
        const R fac = D.cut_fac[t];
        R s, c;
        rb_sincos(fac * (r - l), s, c);
        v = R(0.5) * (R(1.0) + c);
        dv = -R(0.5) * fac * s;
    }
}

template <typename R>
__device__ __forceinline__ void rb_va(const RbDevT<R> &D, int t, R r, R &v, R &dv)
{
    if (t == RB_CC) {
        const R e1 = D.ccB[0] * exp(-D.ccBeta[0] * r);
        const R e2 = D.ccB[1] * exp(-D.ccBeta[1] * r);
        const R e3 = D.ccB[2] * exp(-D.ccBeta[2] * r);
        v = -(e1 + e2 + e3);
        dv = -(-D.ccBeta[0] * e1 - D.ccBeta[1] * e2 - D.ccBeta[2] * e3);
    } else {
        const R e1 = D.B1[t] * exp(-D.beta1[t] * r);
        v = -e1;
        dv = D.beta1[t] * e1;
    }
}

template <typename R>
__device__ __forceinline__ void rb_vr(const RbDevT<R> &D, int t, R r, R &v, R &dv)
{
    const R e1 = D.A[t] * exp(-D.alpha[t] * r);
    const R hl = 1 + D.Q[t] / r;
    v = hl * e1;
    dv = (-D.Q[t] / (r * r) - hl * D.alpha[t]) * e1;
}

template <typename R>
__device__ __forceinline__ void rb_poly5(const R *c, R x, R &v, R &dv)
{
    const R x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x;
    v = c[0] + c[1] * x + c[2] * x2 + c[3] * x3 + c[4] * x4 + c[5] * x5;
    dv = c[1] + 2 * c[2] * x + 3 * c[3] * x2 + 4 * c[4] * x3 + 5 * c[5] * x4;
}

template <typename R>
__device__ __forceinline__ void rb_g(const RbDevT<R> &D, int ktyp, R costh, R n, R &val,
                                     R &dcos, R &dn)
{
    // one polynomial per lane from a per-lane coefficient set (carbon: g_2
    // below N = 3.2, else g_1; hydrogen: its interval's set), so lanes of
    // different branches share the loads; carbon between 3.2 and n_hi also
    // takes g_2 and the switch between them
    const bool kc = ktyp == RB_C;
    const int j = costh < D.th1 ? 0 : (costh < D.th2 ? 1 : 2);
    int q = (int)(-costh * R(12.0)) + 12;
    q = q < 0 ? 0 : (q > 24 ? 24 : q);
    const R *c1 = kc ? (n < R(3.2) ? D.g2c[j] : D.g1c[j]) : D.spgh[D.igh[q] - 1];
    R v1, d1;
    rb_poly5(c1, costh, v1, d1);
    if (kc && !(n < R(3.2)) && !(n > D.n_hi)) {
        R v2, d2, s, c;
        rb_poly5(D.g2c[j], costh, v2, d2);
        rb_sincos(2 * R(M_PI) * (n - R(3.2)), s, c);
        const R sw = (1 + c) / 2;
        val = v1 * (1 - sw) + v2 * sw;
        dcos = d1 * (1 - sw) + d2 * sw;
        dn = (v2 - v1) * (-R(M_PI) * s);
    } else {
        val = v1;
        dcos = d1;
        dn = R(0.0);
    }
}

template <typename R>
__device__ __forceinline__ void rb_h(const RbDevT<R> &D, int ijpot, int ikpot, R dr, R &v,
                                     R &dv)
{
    if (ijpot + ikpot <= 4) { v = R(1.0); dv = R(0.0); }
    else {
        v = D.conear[ijpot][ikpot] * exp(D.conalp * dr);
        dv = D.conalp * v;
    }
}

template <typename R>
__device__ __forceinline__ int rb_box(R v, int n)
{
    int b = (int)v;
    return b < 0 ? 0 : (b >= n ? n - 1 : b);
}

template <typename R>
__device__ void rb_table3(const R *__restrict__ T, R x, R y, R z, R &f,
                          R &fx, R &fy, R &fz)
{
    const int bx = rb_box(x, 4), by = rb_box(y, 4), bz = rb_box(z, 9);
    const R *c = T + (size_t)((bx * 4 + by) * 9 + bz) * 64;
    const R x1 = x - bx, x2 = y - by, x3 = z - bz;
    R F = R(0.0), Fi = R(0.0), Fj = R(0.0), Fk = R(0.0);
    for (int i = 3; i >= 0; --i) {
        R s = R(0.0), sj = R(0.0), sk = R(0.0);
        for (int j = 3; j >= 0; --j) {
            R t = R(0.0), tk = R(0.0);
            for (int k = 3; k >= 0; --k) {
                const R cijk = c[16 * i + 4 * j + k];
                t = t * x3 + cijk;
                if (k > 0) tk = tk * x3 + k * cijk;
            }
            s = s * x2 + t;
            if (j > 0) sj = sj * x2 + j * t;
            sk = sk * x2 + tk;
        }
        F = F * x1 + s;
        if (i > 0) Fi = Fi * x1 + i * s;
        Fj = Fj * x1 + sj;
        Fk = Fk * x1 + sk;
    }
    f = F; fx = Fi; fy = Fj; fz = Fk;
}

template <typename R>
__device__ void rb_table2(const R *__restrict__ T, R x, R y, R &f, R &fx,
                          R &fy)
{
    const int bx = rb_box(x, 5), by = rb_box(y, 5);
    const R *c = T + (size_t)(bx * 5 + by) * 16;
    const R x1 = x - bx, x2 = y - by;
    R h = R(0.0), dh = R(0.0), dc = R(0.0);
    for (int i = 3; i >= 0; --i) {
        R s = R(0.0), sd = R(0.0);
        for (int j = 3; j >= 0; --j) {
            const R cij = c[4 * i + j];
            s = s * x2 + cij;
            if (j > 0) sd = sd * x2 + j * cij;
        }
        h = h * x1 + s;
        if (i > 0) dh = dh * x1 + i * s;
        dc = dc * x1 + sd;
    }
    f = h; fx = dh; fy = dc;
}

// fixed-point accumulation
__device__ __forceinline__ void fx_add(unsigned long long *a, double v, int &bad)
{
    if (!(fabs(v) < RB_FX_MAX)) bad |= 1;
    const long long q = __double2ll_rn(v * RB_FX);
    atomicAdd(a, (unsigned long long)q);
}

template <typename T>
__device__ __forceinline__ void fx_add3(unsigned long long *a, int i, const T *v, int &bad)
{
    fx_add(a + 3 * (size_t)i + 0, v[0], bad);
    fx_add(a + 3 * (size_t)i + 1, v[1], bad);
    fx_add(a + 3 * (size_t)i + 2, v[2], bad);
}

__device__ __forceinline__ bool dc_zero_sum(int c1, int c2)
{
    int a1, b1, e1, a2, b2, e2;
    dc_unpack(c1, a1, b1, e1);
    dc_unpack(c2, a2, b2, e2);
    return a1 + a2 == 0 && b1 + b2 == 0 && e1 + e2 == 0;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Width classes of the per-centre neighbour pass: up to 4, 8 and 16 bonds
// (groups of 4, 8 and 16 lanes).  Class counts are packed into one 64-bit
// word, RB_CLS_BITS bits each, for a single scan (nat < 2^RB_CLS_BITS).
constexpr int RB_CLS_BITS = 21;
__device__ __forceinline__ int rb_class(int nb) { return nb <= 4 ? 0 : (nb <= 8 ? 1 : 2); }

// Elements and LAMMPS tags in slot order (the bond slots of slot p are its
// row's entries: the list's slot-order offsets); also clears the call's
// fixed-point accumulators (eacc, wacc optional), flags and energy
__global__ void k_rb_slots(int nat, const int *__restrict__ order, const int *__restrict__ el,
                           const int *__restrict__ tag, int *__restrict__ el_s, int *__restrict__ tag_s,
                           unsigned long long *__restrict__ facc, unsigned long long *__restrict__ eacc,
                           unsigned long long *__restrict__ wacc, int *__restrict__ flags, double *__restrict__ epot)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < 4) flags[p] = 0;
    if (p == 0) *epot = 0.0;
    if (p >= nat) return;
    const int i = order[p];
    el_s[p] = el[i];
    if (tag) tag_s[p] = tag[i];
    for (int x = 0; x < 3; ++x) facc[3 * (size_t)p + x] = 0ull;
    if (eacc) eacc[p] = 0ull;
    if (wacc)
        for (int q = 0; q < 9; ++q) wacc[9 * (size_t)p + q] = 0ull;
}

// thread per slot p (atom order[p]): its bonds at offs_s[p] .., neighbours
// by slot (the list's rows name slots, el_s and the slot records hold their
// elements and positions); bsrc (optional) = the bond's entry in the
// caller-order list.  bond32 (optional, the mixed form): the same records
// rounded to FP32
__global__ void __launch_bounds__(256)
k_rb_bonds(int nat, const int *__restrict__ order, const int *__restrict__ el_s, CellArg C,
           const RbDev *__restrict__ Dp, NlRows L, const int *__restrict__ offs_s, const int *__restrict__ offsets,
           RbBond *__restrict__ bond, RbBondT<float> *__restrict__ bond32, int2 *__restrict__ bij,
           int *__restrict__ bown, int *__restrict__ bcnt, double2 *__restrict__ nn, int *__restrict__ bsrc,
           unsigned long long *__restrict__ cflag)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    const RbDev &D = *Dp;
    const int ki = el_s[p];
    const int beg = offs_s[p], end = offs_s[p + 1];
    int nrow;
    const int2 *row = nl_row(L, p, nrow);
    int nb = 0;
    double nC = 0.0, nH = 0.0;
    if (ki > 0) {
        const SortedAtom A = L.atoms[p];
        const double ri[3] = {A.x, A.y, A.z};
        for (int k = 0; k < nrow; ++k) {
            const int2 e = row[k];
            const int kj = el_s[e.x];
            if (kj <= 0) continue;
            int a, b, c;
            dc_unpack(e.y, a, b, c);
            const SortedAtom Aj = L.atoms[e.x];
            const double rj[3] = {Aj.x, Aj.y, Aj.z};
            double d[3];
            delta_r(C, ri, rj, a, b, c, d);
            const double rl = dot3_nc(d);
            const int t = rb_pair(ki, kj);
            double len, fc, dfc;
            if (rl < D.cut_l[t] * D.cut_l[t]) {
                len = sqrt(rl);
                fc = 1.0;
                dfc = 0.0;
            } else if (rl < D.cut_h2[t]) {
                len = sqrt(rl);
                rb_fcin<double>(D, t, len, fc, dfc);
            } else {
                continue;
            }
            // r_ij = r_j - r_i - A.dc = -d, normalised by division
            RbBond B;
            B.len = len;
            B.fc = fc;
            B.dfc = dfc;
            B.n[0] = -d[0] / len;
            B.n[1] = -d[1] / len;
            B.n[2] = -d[2] / len;
            B.fx = 0.0;
            B.dfx = 0.0;
            if (kj == RB_C) nC += fc;
            else nH += fc;
            bond[beg + nb] = B;
            if (bond32)
                bond32[beg + nb] = RbBondT<float>{{(float)B.n[0], (float)B.n[1], (float)B.n[2]}, (float)len,
                                                  (float)fc, (float)dfc, 0.0f, 0.0f};
            bij[beg + nb] = make_int2(e.x, e.y);
            bown[beg + nb] = p;
            if (bsrc) bsrc[beg + nb] = offsets[order[p]] + k;   // list entry of the bond (nbb of the reference)
            ++nb;
        }
    }
    for (int s = beg + nb; s < end; ++s) bown[s] = -1;
    bcnt[p] = nb;
    nn[p] = make_double2(nC, nH);
    if (cflag) {
        cflag[p] = 1ull << (RB_CLS_BITS * rb_class(nb));
        if (p == nat - 1) cflag[nat] = 0ull;
    }
}

// The centres of each width class (rb_class) in ascending slot order, from
// the exclusive scan of the packed class flags: list k at clist + k*nat,
// its length in ccount[k]
__global__ void k_rb_class_list(int nat, const int *__restrict__ bcnt, const unsigned long long *__restrict__ cpos,
                                int *__restrict__ clist, int *__restrict__ ccount)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long m = (1ull << RB_CLS_BITS) - 1;
    if (p == nat)
        for (int k = 0; k < 3; ++k) ccount[k] = (int)((cpos[nat] >> (RB_CLS_BITS * k)) & m);
    if (p >= nat) return;
    const int k = rb_class(bcnt[p]);
    clist[(size_t)k * nat + (int)((cpos[p] >> (RB_CLS_BITS * k)) & m)] = p;
}

__global__ void __launch_bounds__(256)
k_rb_conj(int nat, const int *__restrict__ el, const int *__restrict__ offsets,
          const int *__restrict__ bcnt, const int2 *__restrict__ bij, const double2 *__restrict__ nn,
          RbBond *__restrict__ bond, RbBondT<float> *__restrict__ bond32, double *__restrict__ nconjit)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nat) return;
    const int beg = offsets[i], nb = bcnt[i];
    double s = 0.0;
    for (int b = beg; b < beg + nb; ++b) {
        const int k = bij[b].x;
        if (el[k] != RB_C) continue;
        const double2 nk = nn[k];
        const double fc = bond[b].fc;
        const double x = nk.x + nk.y - fc;
        double fx, dfx;
        rb_fconj(x, fx, dfx);
        bond[b].fx = fx;
        bond[b].dfx = dfx;
        if (bond32) {
            bond32[b].fx = (float)fx;
            bond32[b].dfx = (float)dfx;
        }
        s = s + fc * fx;
    }
    nconjit[i] = s;
}

// virial and force sums are held in FP64 in both forms (products in R)
template <typename R>
__device__ __forceinline__ void outer_sub(double *w, const R *a, const R *b)
{
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) w[x + 3 * y] = w[x + 3 * y] - (double)(a[x] * b[y]);
}

// zeta of the bond c->o over the other bonds of centre c, and dzeta/dN
// (ik_loop2 of bop_kernel_rebo2.f90:1407-1594 without the derivatives)
template <typename R>
__device__ __forceinline__ void rb_zeta(const RbDevT<R> &D, int kc, int ijpot, R rlij, const R *rn,
                                        int beg, int nb, int skip, const RbBondT<R> *__restrict__ bond,
                                        const int2 *__restrict__ bij, const int *__restrict__ el, R nt,
                                        R &z, R &dzdn)
{
    z = R(0.0);
    dzdn = R(0.0);
    for (int b = beg; b < beg + nb; ++b) {
        if (b == skip) continue;
        const RbBondT<R> B = bond[b];
        const int ikpot = rb_pair(kc, el[bij[b].x]);
        const R rlik = B.len;
        if (!(rlik < D.cut_h[ikpot])) continue;
        R qf, qd, gf, gd, dgdn;
        rb_h(D, ijpot, ikpot, rlij - rlik, qf, qd);
        R costh = R(0.0);
        costh = costh + B.n[0] * rn[0];
        costh = costh + B.n[1] * rn[1];
        costh = costh + B.n[2] * rn[2];
        rb_g(D, kc, costh, nt, gf, gd, dgdn);
        dzdn = dzdn + B.fc * dgdn * qf;
        z = z + B.fc * gf * qf;
    }
}

// Forces on the neighbours k of a centre c (bond c->o): the coordination
// and conjugation terms (dF/dN, dF/dN_conj, P) with their second neighbours
// m, and the bond-order terms of b_co, one accumulator update per k
// (bop_kernel_rebo2.f90:2405-2520 and the dbidk/dbjdl parts of 1407-1920).
// sgn = +1 for the centre i (bond towards j), -1 for the centre j; the
// bond-order forces on the centre and on the other atom of the bond go to
// fcen and foth, the virial to w, already scaled by -dV_A b'.
template <typename R>
__device__ __forceinline__ void rb_neighbours(const RbDevT<R> &D, int centre, int kc, int ijpot, R rlij,
                                              R rlijr, const R *rnij, R sgn, int beg, int nb,
                                              int skip, const RbBondT<R> *__restrict__ bond,
                                              const int2 *__restrict__ bij, const int *__restrict__ offsets,
                                              const int *__restrict__ bcnt, const int *__restrict__ el, R nt,
                                              R dfdn, R dfdncn, R dfb, R dpdnc, R dpdnh,
                                              double *fcen, double *foth, double *w,
                                              unsigned long long *__restrict__ facc, int &bad)
{
    const R rco[3] = {sgn * rlij * rnij[0], sgn * rlij * rnij[1], sgn * rlij * rnij[2]};
    for (int b = beg; b < beg + nb; ++b) {
        if (b == skip) continue;
        const RbBondT<R> B = bond[b];
        const int2 e = bij[b];
        const int k = e.x;
        const int ko = el[k];
        const bool kC = ko == RB_C;
        R fk[3], dr[3];
        // coordination / conjugation
        const R dpn = kC ? dpdnc : dpdnh;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const R dn = B.n[x] * B.dfc;
            const R dncn = kC ? (B.fx * B.dfc) * B.n[x] : R(0.0);
            fk[x] = -(dfdn * dn + dfdncn * dncn) - dfb * (dpn * dn);
            fcen[x] -= fk[x];
            dr[x] = B.len * B.n[x];
        }
        outer_sub(w, dr, fk);
        if (kC && dfdncn != R(0.0)) {
            const R nconjdx = B.fc * B.dfx;
            const int kb = offsets[k], kn = bcnt[k];
            for (int m = kb; m < kb + kn; ++m) {
                const int2 em = bij[m];
                // the centre itself (m == centre at the centre's image) is skipped
                if (em.x == centre && dc_zero_sum(e.y, em.y)) continue;
                const RbBondT<R> M = bond[m];
                R d2[3], drm[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    d2[x] = -dfdncn * (nconjdx * (M.dfc * M.n[x]));
                    fk[x] -= d2[x];
                    drm[x] = M.len * M.n[x];
                }
                outer_sub(w, drm, d2);
                fx_add3(facc, em.x, d2, bad);
            }
        }
        // bond order: angular terms of zeta_co with the final dV_A b'
        const int ikpot = rb_pair(kc, ko);
        const R rlik = B.len;
        if (rlik < D.cut_h[ikpot]) {
            R qf, qd, gf, gd, dgdn;
            rb_h(D, ijpot, ikpot, rlij - rlik, qf, qd);
            R costh = R(0.0);
            costh = costh + B.n[0] * rnij[0];
            costh = costh + B.n[1] * rnij[1];
            costh = costh + B.n[2] * rnij[2];
            if (sgn < 0) costh = -costh;
            rb_g(D, kc, costh, nt, gf, gd, dgdn);
            const R fcik = B.fc;
            R dk[3], dd = R(0.0);
#pragma unroll
            for (int x = 0; x < 3; ++x) dk[x] = B.n[x] * rlik - sgn * rnij[x] * rlij;
            dd = dd + dk[0] * dk[0];
            dd = dd + dk[1] * dk[1];
            dd = dd + dk[2] * dk[2];
            const R dis = sqrt(dd);
            const R rdis = R(1.0) / dis, rlikr = R(1.0) / rlik;
#pragma unroll
            for (int x = 0; x < 3; ++x) dk[x] = dk[x] * rdis;
            const R dcs_io = rlikr - costh * rlijr;
            const R dcs_ik = rlijr - costh * rlikr;
            const R dcs_ok = -dis * rlijr * rlikr;
            const R dzfac = fcik * gd * qf;
            const R dzdro = gf * fcik * qd;
            const R dzdrk = gf * (B.dfc * qf - fcik * qd);
            R dfo[3], dbk[3];
            const R rik[3] = {rlik * B.n[0], rlik * B.n[1], rlik * B.n[2]};
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const R u = sgn * rnij[x];      // unit vector centre -> other
                const R dcs_c = -dcs_io * u - dcs_ik * B.n[x];
                const R dcs_o = dcs_io * u - dcs_ok * dk[x];
                const R dcs_k = dcs_ik * B.n[x] + dcs_ok * dk[x];
                fcen[x] += -dfb * (-dzdro * u - dzdrk * B.n[x] + dzfac * dcs_c);
                dfo[x] = dzdro * u + dzfac * dcs_o;
                foth[x] += -dfb * dfo[x];
                dbk[x] = dzdrk * B.n[x] + dzfac * dcs_k;
                fk[x] += -dfb * dbk[x];
            }
            // virial -dV_A b' (-(r_co x dfo) - (r_ck x dbk))
#pragma unroll
            for (int y = 0; y < 3; ++y)
#pragma unroll
                for (int x = 0; x < 3; ++x)
                    w[x + 3 * y] = w[x + 3 * y] + dfb * (rco[x] * dfo[y] + rik[x] * dbk[y]);
        }
        fx_add3(facc, k, fk, bad);
    }
}

template <typename R>
__device__ __forceinline__ void outer_add(double *w, const R *a, const R *b)
{
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) w[x + 3 * y] = w[x + 3 * y] + (double)(a[x] * b[y]);
}

// Bond order of every directed bond c->o (thread per bond slot): zeta with
// P(N_H, N_C) for a carbon centre, b = (1 + zeta)^-1/2 and dV_A b'/dzeta,
// dP/dN + dzeta/dN; the slot of the reverse bond o->c; and whether the slot
// carries a half pair ((dc = 0 and o > c) or dc > 0; for a LAMMPS step, o of
// tag >= tag(c) with c local).
template <typename R>
__global__ void __launch_bounds__(256)
k_rb_bo(int ntot, const RbDevT<R> *__restrict__ Dp, const R *__restrict__ tab, const int *__restrict__ el,
        const int *__restrict__ offsets, const int *__restrict__ bcnt, const int *__restrict__ bown,
        const RbBondT<R> *__restrict__ bond, const int2 *__restrict__ bij, const double2 *__restrict__ nn,
        Rb4<R> *__restrict__ bo, int *__restrict__ rev, int *__restrict__ half, int natloc,
        const int *__restrict__ tag, const int *__restrict__ order, int *__restrict__ flags,
        double4 *__restrict__ coef, double *__restrict__ fpr)
{
    const RbDevT<R> &D = *Dp;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntot) return;
    // neighbour-pass coefficients: zero (.w = 0) until a half pair sets them;
    // the pair forces by directed bond likewise
    coef[s] = make_double4(0.0, 0.0, 0.0, 0.0);
    if (fpr)
        for (int x = 0; x < 3; ++x) fpr[3 * (size_t)s + x] = 0.0;
    const int i = bown[s];
    if (i < 0) {
        half[s] = 0;
        return;
    }
    const int2 e = bij[s];
    const int j = e.x;
    const int ki = el[i], kj = el[j];
    const RbBondT<R> B = bond[s];
    const int ijpot = rb_pair(ki, kj);
    int a, b, c;
    dc_unpack(e.y, a, b, c);
    const bool pos = a != 0 ? a > 0 : (b != 0 ? b > 0 : c > 0);
    const bool zero = a == 0 && b == 0 && c == 0;
    const bool act = B.len < D.cut_h[ijpot];
    // i and j are slots; the rules compare caller indices order[.].  LAMMPS:
    // local centres only, tag(j) >= tag(i) (bop_kernel_rebo2.f90:1328-1329)
    if (tag)
        half[s] = (act && order[i] < natloc && tag[j] >= tag[i]) ? 1 : 0;
    else
        half[s] = (act && ((zero && order[j] > order[i]) || pos)) ? 1 : 0;
    int sr = -1;
    const int jb = offsets[j], jn = bcnt[j];
    for (int q = jb; q < jb + jn; ++q) {
        const int2 eq = bij[q];
        if (eq.x == i && dc_zero_sum(eq.y, e.y)) { sr = q; break; }
    }
    rev[s] = sr;
    if (sr < 0) atomicOr(flags, DERR_BOND_OVERFLOW);
    if (!act) return;
    const double2 nni = nn[i];
    R niC = nni.x, niH = nni.y;
    if (kj == RB_C) niC = niC - B.fc; else niH = niH - B.fc;
    niC = niC > R(4.0) ? R(4.0) : niC;
    niH = niH > R(4.0) ? R(4.0) : niH;
    R z, dzdn;
    rb_zeta(D, ki, ijpot, B.len, B.n, offsets[i], bcnt[i], s, bond, bij, el, niC + niH, z, dzdn);
    R dpc = R(0.0), dph = R(0.0);
    if (ki == RB_C) {
        R pij;
        rb_table2(tab + (ijpot == RB_CC ? RB_TAB_PCC : RB_TAB_PCH), niH, niC, pij, dph, dpc);
        z = z + pij;
        dpc = dpc + dzdn;
        dph = dph + dzdn;
    }
    R fa, dfa;
    rb_va(D, ijpot, B.len, fa, dfa);
    const R bb = pow(R(1.0) + z, D.conpe);
    const R dfb = D.conan * B.fc * fa * pow(R(1.0) + z, D.conpf);
    bo[s] = Rb4<R>{bb, dfb, dpc, dph};
}

// ordered list of the half-pair slots (exclusive scan of the flags)
__global__ void k_rb_list(int ntot, const int *__restrict__ half, const int *__restrict__ pos,
                          int *__restrict__ list, int *__restrict__ count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntot) return;
    if (half[s]) list[pos[s]] = s;
    if (s == ntot - 1) *count = pos[s] + half[s];
}

// [e, W(9)] of a thread reduced over the block into partial[blk]
__device__ __forceinline__ void rb_block_record(double e, const double *w, double *__restrict__ partial, int blk)
{
    __shared__ double s_red[4][10];
    double red[10];
    red[0] = e;
#pragma unroll
    for (int q = 0; q < 9; ++q) red[1 + q] = w[q];
#pragma unroll
    for (int q = 0; q < 10; ++q) red[q] = group_sum<64>(red[q]);
    const int tid = threadIdx.x, wave = tid / 64;
    if ((tid & 63) == 0)
        for (int q = 0; q < 10; ++q) s_red[wave][q] = red[q];
    __syncthreads();
    if (tid < 10)
        partial[(size_t)blk * 10 + tid] = ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
}

// the FP64 radial pair terms of a half pair for a bond-order average bave:
// energy 2 hlf, force df on i (-df on j), virial r_ij (x) df
__device__ __forceinline__ void rb_radial(const RbBond &B, double fa, double dfa, double fr, double dfr,
                                          double bave, double &hlf, double *df)
{
    hlf = B.fc * (fr + bave * fa) / 2;
    const double dffac = dfr * B.fc + bave * dfa * B.fc + fr * B.dfc + bave * fa * B.dfc;
#pragma unroll
    for (int x = 0; x < 3; ++x) df[x] = dffac * B.n[x];
}

// Pass 1, thread per half pair i->j (ij_loop of bop_kernel_rebo2.f90:1326-
// 2779 without the neighbour loops): N^conj of j, F(N_i, N_j, N^conj) and
// T(N_i, N_j, N^conj); the pair energy and radial force with
// b = (b_ij + b_ji + F) / 2 (the torsion part T sum follows in pass 2); the
// neighbour-loop coefficients of both directed bonds into coef[slot] =
// {dE/dN_c, dE/dN^conj_c, N^conj_c}; the torsion table values into tor[t].
// The mixed form takes R = float for the tables and coefficients and the
// FP64 records (bond64) and constants (D64) for the radial terms.
template <typename R>
__global__ void __launch_bounds__(256)
k_rb_pair(const int *__restrict__ list, const int *__restrict__ count, const RbDevT<R> *__restrict__ Dp,
          const R *__restrict__ tab, const int *__restrict__ el, const int *__restrict__ offsets,
          const int *__restrict__ bcnt, const int *__restrict__ bown, const RbBondT<R> *__restrict__ bond,
          const int2 *__restrict__ bij, const double2 *__restrict__ nn, const double *__restrict__ nconjit,
          const Rb4<double> *__restrict__ bo, const int *__restrict__ rev, double4 *__restrict__ coef,
          Rb4<R> *__restrict__ tor, int *__restrict__ tflag, int ntot, unsigned long long *__restrict__ facc,
          unsigned long long *__restrict__ eacc,
          unsigned long long *__restrict__ wacc, double *__restrict__ partial, int *__restrict__ flags,
          const int *__restrict__ bsrc, double *__restrict__ epb, double *__restrict__ fpb,
          double *__restrict__ wpb, const RbDev *__restrict__ D64p, const RbBond *__restrict__ bond64,
          double *__restrict__ fpr)
{
    const RbDevT<R> &D = *Dp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    double epair = 0.0;
    double wij[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) wij[q] = 0.0;
    const int nhalf = *count;
    if (tflag && t >= nhalf && t < ntot) tflag[t] = 0;
    if (t < nhalf) {
        const int s = list[t];
        const int i = bown[s];
        const int j = bij[s].x;
        const int sji = rev[s];
        const int ki = el[i], kj = el[j];
        const RbBondT<R> Bij = bond[s];
        const int ijpot = rb_pair(ki, kj);
        const R fcar = Bij.fc;
        const double2 nni = nn[i], nnj = nn[j];
        R niC = nni.x, niH = nni.y, njC = nnj.x, njH = nnj.y;
        if (kj == RB_C) niC = niC - Bij.fc; else niH = niH - Bij.fc;
        if (ki == RB_C) njC = njC - Bij.fc; else njH = njH - Bij.fc;
        niC = niC > R(4.0) ? R(4.0) : niC;
        niH = niH > R(4.0) ? R(4.0) : niH;
        njC = njC > R(4.0) ? R(4.0) : njC;
        njH = njH > R(4.0) ? R(4.0) : njH;
        // F sees N_i, N_j clamped to 3
        R nti = niC + niH, ntj = njC + njH;
        R fa, dfa;
        rb_va(D, ijpot, Bij.len, fa, dfa);
        const int jb = offsets[j], jnb = bcnt[j];
        const double bi = bo[s].x, bj = bo[sji].x;
        const R nconji = nconjit[i] - Bij.fc * Bij.fx;
        R nconjj = R(0.0);
        for (int b = jb; b < jb + jnb; ++b)
            if (b != sji && el[bij[b].x] == RB_C) nconjj = nconjj + bond[b].fc * bond[b].fx;
        R nconj = nconji * nconji + nconjj * nconjj;
        nconj = nconj > R(8.0) ? R(8.0) : nconj;
        nti = nti > R(3.0) ? R(3.0) : nti;
        ntj = ntj > R(3.0) ? R(3.0) : ntj;
        Rb4<R> tv{R(0.0), R(0.0), R(0.0), R(0.0)};
        if (D.with_dihedral && ijpot == RB_CC) rb_table3(tab + RB_TAB_TCC, nti, ntj, nconj, tv.x, tv.y, tv.z, tv.w);
        if (tor) {
            tor[t] = tv;
            tflag[t] = tv.x != R(0.0) ? 1 : 0;
        }
        // F of the pair's table, one evaluation per lane whatever the pair
        // types in the wavefront: F_CH is read with (N_j, N_i) when i is
        // the carbon, and its two N derivatives swap back
        const R *ftab = tab + (ijpot == RB_CC ? RB_TAB_FCC : (ijpot == RB_HH ? RB_TAB_FHH : RB_TAB_FCH));
        const bool fsw = ijpot != RB_CC && ijpot != RB_HH && ki == RB_C;
        R fij = R(0.0), df1 = R(0.0), df2 = R(0.0), dfdncn = R(0.0);
        rb_table3(ftab, fsw ? ntj : nti, fsw ? nti : ntj, nconj, fij, df1, df2, dfdncn);
        const R dfdni = fsw ? df2 : df1, dfdnj = fsw ? df1 : df2;
        const R h = R(0.5) * fcar * fa;
        dfdncn = h * dfdncn;
        coef[s] = make_double4(h * dfdni, 2 * dfdncn * nconji, nconji, 1.0);
        coef[sji] = make_double4(h * dfdnj, 2 * dfdncn * nconjj, nconjj, 1.0);
        // pair energy and radial force (FP64 terms in the mixed form)
        const RbBond B64 = bond64[s];
        double fa2, dfa2, fr2, dfr2, hlf, df[3];
        rb_va<double>(*D64p, ijpot, B64.len, fa2, dfa2);
        rb_vr<double>(*D64p, ijpot, B64.len, fr2, dfr2);
        rb_radial(B64, fa2, dfa2, fr2, dfr2, 0.5 * (bi + bj + (double)fij), hlf, df);
        epair = hlf + hlf;
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x) wij[x + 3 * y] = (B64.len * B64.n[x]) * df[y];
        const double mdf[3] = {-df[0], -df[1], -df[2]};
        if (fpr) {
            // by directed bond (the half pair owns both slots): i's on i->j,
            // j's on j->i, summed by the owner gather
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                fpr[3 * (size_t)s + x] = df[x];
                fpr[3 * (size_t)sji + x] = mdf[x];
            }
        } else {
            fx_add3(facc, i, df, bad);
            fx_add3(facc, j, mdf, bad);
        }
        // per-bond properties on the list entry of the half pair's bond i->j
        // (bop_kernel_rebo2.f90:2546-2547, 2571-2572, 2762-2763); one thread
        // owns each half pair in each pass
        if (bsrc) {
            const size_t e = (size_t)bsrc[s];
            if (epb) epb[e] += epair;
            if (fpb)
                for (int x = 0; x < 3; ++x) fpb[3 * e + x] += df[x];
            if (wpb)
                for (int q = 0; q < 9; ++q) wpb[9 * e + q] += wij[q];
        }
        if (eacc) {
            fx_add(eacc + i, hlf, bad);
            fx_add(eacc + j, hlf, bad);
        }
        if (wacc) {
            for (int q = 0; q < 9; ++q) {
                fx_add(wacc + 9 * (size_t)i + q, wij[q] / 2, bad);
                fx_add(wacc + 9 * (size_t)j + q, wij[q] / 2, bad);
            }
        }
    }
    if (bad) atomicOr(flags, ((bad & 1) ? DERR_FIXED_RANGE : 0) | ((bad & 2) ? DERR_BOND_OVERFLOW : 0));
    rb_block_record(epair, wij, partial, blockIdx.x);
}

// Pass 2, thread per half pair with T != 0 (C-C with torsion): the torsion
// sum over k, l with its forces (bop_kernel_rebo2.f90:1939-2120, always on
// the FP64 records), its share T sum / 2 of the bond order in the pair
// energy and radial force, and its terms dT/dN sum, dT/dN^conj sum added to
// both directed bonds' coefficients before the neighbour pass.
// one (k, l) term of the torsion sum of pair i-j (the body of the double loop
// of bop_kernel_rebo2.f90:1939-2120): returns (1 - cos^2 w) fc_ik fc_jl;
// forces on i, j into fi, fj and on k, l into fk, fl; virial into w
__device__ __forceinline__ double rb_torsion_term(const RbBond &K, const RbBond &L, const double *rnij, double rlij,
                                                  double tij, double fa, double fcar, double *fi, double *fj,
                                                  double *fk, double *fl, double *w)
{
    const double tije = tij * fa * fcar;
    const double *rnik = K.n, *rnjl = L.n;
    const double rlik = K.len, rljl = L.len;
    double dot_ij_ik = 0.0, dot_ij_jl = 0.0, dot_ik_jl = 0.0;
    dot_ij_ik = dot_ij_ik + rnij[0] * rnik[0];
    dot_ij_ik = dot_ij_ik + rnij[1] * rnik[1];
    dot_ij_ik = dot_ij_ik + rnij[2] * rnik[2];
    dot_ij_jl = dot_ij_jl + rnij[0] * rnjl[0];
    dot_ij_jl = dot_ij_jl + rnij[1] * rnjl[1];
    dot_ij_jl = dot_ij_jl + rnij[2] * rnjl[2];
    dot_ik_jl = dot_ik_jl + rnik[0] * rnjl[0];
    dot_ik_jl = dot_ik_jl + rnik[1] * rnjl[1];
    dot_ik_jl = dot_ik_jl + rnik[2] * rnjl[2];
    const double dcik = 1.0 - dot_ij_ik * dot_ij_ik;
    const double dcjl = 1.0 - dot_ij_jl * dot_ij_jl;
    const double abs_dc = sqrt(dcik * dcjl);
    // the reciprocals once (the terms below are products with them)
    const double ra = 1.0 / abs_dc, rik_ = 1.0 / dcik, rjl_ = 1.0 / dcjl;
    const double cost = (dot_ij_ik * dot_ij_jl - dot_ik_jl) * ra;
    double bdhij = 1 - cost * cost;
    const double term = bdhij * K.fc * L.fc;
    bdhij = bdhij * tij * fa * fcar / 2;
    const double dbdhij = -2 * cost * tije * K.fc * L.fc / 2;
    const double cij = dbdhij / rlij, cik = dbdhij / rlik, cjl = dbdhij / rljl;
    const double a_ik = dot_ij_jl * ra + cost * dot_ij_ik * rik_;     // coefficient of n_ik in dcos/dn
    const double a_jl = dot_ij_ik * ra + cost * dot_ij_jl * rjl_;     // coefficient of n_jl
    double df[3], v[3];
#pragma unroll
    for (int x = 0; x < 3; ++x)
        df[x] = cij * (a_ik * rnik[x] + a_jl * rnjl[x] - (2 * dot_ik_jl * ra + cost * (rik_ + rjl_)) * rnij[x]);
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        fi[x] += df[x];
        fj[x] -= df[x];
        v[x] = rlij * rnij[x];
    }
    outer_add(w, v, df);
#pragma unroll
    for (int x = 0; x < 3; ++x)
        df[x] = cik * (-rik_ * cost * rnik[x] - ra * rnjl[x] + a_ik * rnij[x]) + bdhij * K.dfc * L.fc * rnik[x];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        fi[x] += df[x];
        fk[x] = -df[x];
        v[x] = rlik * rnik[x];
    }
    outer_add(w, v, df);
#pragma unroll
    for (int x = 0; x < 3; ++x)
        df[x] = cjl * (-rjl_ * cost * rnjl[x] - ra * rnik[x] + a_jl * rnij[x]) + bdhij * K.fc * L.dfc * rnjl[x];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        fj[x] += df[x];
        fl[x] = -df[x];
        v[x] = rljl * rnjl[x];
    }
    outer_add(w, v, df);
    return term;
}

// Pass 2, a group of G lanes (4, 8 or 16: a launch variant) per half pair
// with T != 0 (C-C with torsion, from the scanned list; the grid strides
// over the list): the torsion sum over k (bonds of i but j) and l (bonds of
// j but i, and not k at the same image) (bop_kernel_rebo2.f90:1939-2120,
// always on the FP64 records), lane q taking the bonds k = q, q + G, ...
// while the group walks the bonds l in order, with their forces on k and l;
// the group sums give T sum, the forces on i, j and the virial.  Then its share
// T sum / 2 of the bond order in the pair energy and radial force, and its
// dT/dN terms added to both directed bonds' coefficients before the
// neighbour pass.  Pairs are dealt to groups in a fixed order, so the block
// records do not depend on scheduling.
template <typename R, int G>
__global__ void __launch_bounds__(256)
k_rb_torsion(const int *__restrict__ list, const int *__restrict__ tlist, const int *__restrict__ tcount,
             const int *__restrict__ offsets, const int *__restrict__ bcnt, const int *__restrict__ bown,
             const int2 *__restrict__ bij, const int *__restrict__ rev, const Rb4<R> *__restrict__ tor,
             double4 *__restrict__ coef,
             unsigned long long *__restrict__ facc, unsigned long long *__restrict__ eacc,
             unsigned long long *__restrict__ wacc, double *__restrict__ partial, int *__restrict__ flags,
             const int *__restrict__ bsrc, double *__restrict__ epb, double *__restrict__ fpb,
             double *__restrict__ wpb, const RbDev *__restrict__ D64p, const RbBond *__restrict__ bond64,
             double *__restrict__ fpr)
{
    const int lane = threadIdx.x % G;
    const int ngroups = gridDim.x * (256 / G);
    int bad = 0;
    double e_run = 0.0, w_run[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w_run[q] = 0.0;
    const int ntor = tor ? *tcount : 0;
    // wavefront-uniform trip count (the group sums need every lane)
    for (int tt0 = blockIdx.x * (256 / G) + (threadIdx.x & ~63) / G; tt0 < ntor; tt0 += ngroups) {
        const int tt = tt0 + (threadIdx.x & 63) / G;
        Rb4<R> tv{R(0.0), R(0.0), R(0.0), R(0.0)};
        int t = 0;
        if (tt < ntor) {
            t = tlist[tt];
            tv = tor[t];
        }
        const bool on = tv.x != R(0.0);
        int s = 0, i = 0, j = 0, sji = 0, ib = 0, ni = 0, jb = 0, nj = 0, xs = 0, xr = 0, a0 = 0, b0 = 0, c0 = 0;
        RbBond B{};
        double fa = 0.0, dfa = 0.0;
        const double tij = (double)tv.x;
        double fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0}, bdh = 0.0, wij[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) wij[q] = 0.0;
        if (on) {
            s = list[t];
            i = bown[s];
            const int2 eij = bij[s];
            j = eij.x;
            sji = rev[s];
            B = bond64[s];
            rb_va<double>(*D64p, RB_CC, B.len, fa, dfa);
            ib = offsets[i];
            ni = bcnt[i] - 1;
            jb = offsets[j];
            nj = bcnt[j] - 1;
            xs = s - ib;                 // the bond i->j in i's row
            xr = sji - jb;               // j->i in j's row
            dc_unpack(eij.y, a0, b0, c0);
        }
        // lane q of the group owns the bonds k = q, q + G, ... of i (but j)
        // and walks the bonds l of j (but i) in order: its force on k is
        // summed in registers, the forces on l over the group, so each
        // atom k or l takes one fixed-point add per pair (per chunk of G
        // bonds k) in a fixed order
        const int nkc = on ? (ni + G - 1) / G : 0;      // the same for the whole group
        for (int kc = 0; kc < nkc; ++kc) {
            const int qk = kc * G + lane;
            const bool kv = qk < ni;
            const int kb = ib + qk + (qk >= xs ? 1 : 0);
            const int2 ek = kv ? bij[kb] : make_int2(-1, 0);
            int ak, bk, ck;
            dc_unpack(ek.y, ak, bk, ck);
            double fks[3] = {0, 0, 0};
            for (int ql = 0; ql < nj; ++ql) {
                const int lb = jb + ql + (ql >= xr ? 1 : 0);
                const int2 e = bij[lb];
                double fl[3] = {0, 0, 0};
                int al, bl, cl;
                dc_unpack(e.y, al, bl, cl);
                // l == k at the same image takes no term
                const bool same = e.x == ek.x && a0 + al == ak && b0 + bl == bk && c0 + cl == ck;
                if (kv && !same) {
                    double fk[3];
                    bdh += rb_torsion_term(bond64[kb], bond64[lb], B.n, B.len, tij, fa, B.fc, fi, fj, fk, fl,
                                           wij);
#pragma unroll
                    for (int x = 0; x < 3; ++x) fks[x] += fk[x];
                }
#pragma unroll
                for (int x = 0; x < 3; ++x) fl[x] = group_sum<G>(fl[x]);
                if (lane == 0) fx_add3(facc, e.x, fl, bad);
            }
            if (kv) fx_add3(facc, ek.x, fks, bad);
        }
        bdh = group_sum<G>(bdh);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            fi[x] = group_sum<G>(fi[x]);
            fj[x] = group_sum<G>(fj[x]);
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) wij[q] = group_sum<G>(wij[q]);
        if (on && lane == 0) {
            // coefficients of the neighbour pass
            const double h = 0.5 * B.fc * fa * bdh;
            const double4 ci = coef[s], cj = coef[sji];
            const double dn = h * (double)tv.w;
            coef[s] = make_double4(ci.x + h * (double)tv.y, ci.y + 2 * dn * ci.z, ci.z, 1.0);
            coef[sji] = make_double4(cj.x + h * (double)tv.z, cj.y + 2 * dn * cj.z, cj.z, 1.0);
            // the pair energy's torsion share: b += T sum / 2, no repulsion
            double hlf, df[3];
            rb_radial(B, fa, dfa, 0.0, 0.0, 0.5 * tij * bdh, hlf, df);
            const double epair = hlf + hlf;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                fi[x] += df[x];
                fj[x] -= df[x];
            }
#pragma unroll
            for (int y = 0; y < 3; ++y)
#pragma unroll
                for (int x = 0; x < 3; ++x) wij[x + 3 * y] += (B.len * B.n[x]) * df[y];
            if (fpr) {
                // the half pair's slots, written by the pair pass
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    fpr[3 * (size_t)s + x] += fi[x];
                    fpr[3 * (size_t)sji + x] += fj[x];
                }
            } else {
                fx_add3(facc, i, fi, bad);
                fx_add3(facc, j, fj, bad);
            }
            if (bsrc) {
                const size_t e = (size_t)bsrc[s];
                if (epb) epb[e] += epair;
                if (fpb)
                    for (int x = 0; x < 3; ++x) fpb[3 * e + x] += df[x];
                if (wpb)
                    for (int q = 0; q < 9; ++q) wpb[9 * e + q] += wij[q];
            }
            if (eacc) {
                fx_add(eacc + i, hlf, bad);
                fx_add(eacc + j, hlf, bad);
            }
            if (wacc) {
                for (int q = 0; q < 9; ++q) {
                    fx_add(wacc + 9 * (size_t)i + q, wij[q] / 2, bad);
                    fx_add(wacc + 9 * (size_t)j + q, wij[q] / 2, bad);
                }
            }
            e_run += epair;
#pragma unroll
            for (int q = 0; q < 9; ++q) w_run[q] += wij[q];
        }
    }
    if (bad) atomicOr(flags, ((bad & 1) ? DERR_FIXED_RANGE : 0) | ((bad & 2) ? DERR_BOND_OVERFLOW : 0));
    rb_block_record(e_run, w_run, partial, blockIdx.x);
}

// Pass 3, thread per directed bond c->o of a half pair (two threads per
// half pair: i->j, then j->i): the coordination, conjugation and bond-order
// forces of c's other neighbours k and their neighbours m
// (bop_kernel_rebo2.f90:2405-2520 and the dbidk/dbjdl parts of 1407-1920),
// with the coefficients of passes 1 and 2.  wdir (optional, for the per-bond
// virial) receives the virial of each directed bond.
template <typename R>
__global__ void __launch_bounds__(256)
k_rb_nbrs(const int *__restrict__ list, const int *__restrict__ count, const RbDevT<R> *__restrict__ Dp,
          const int *__restrict__ el, const int *__restrict__ offsets, const int *__restrict__ bcnt,
          const int *__restrict__ bown, const RbBondT<R> *__restrict__ bond, const int2 *__restrict__ bij,
          const double2 *__restrict__ nn, const Rb4<double> *__restrict__ bo, const int *__restrict__ rev,
          const double4 *__restrict__ coef, unsigned long long *__restrict__ facc,
          unsigned long long *__restrict__ wacc, double *__restrict__ partial, int *__restrict__ flags,
          double *__restrict__ wdir)
{
    const RbDevT<R> &D = *Dp;
    const int t2 = blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    double w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    if (t2 < 2 * *count) {
        const int s0 = list[t2 >> 1];
        const int s = (t2 & 1) ? rev[s0] : s0;
        const int c = bown[s];
        const int o = bij[s].x;
        const int kc = el[c], ko = el[o];
        const RbBondT<R> B = bond[s];
        const int ijpot = rb_pair(kc, ko);
        const double2 nnc = nn[c];
        R nC = nnc.x, nH = nnc.y;
        if (ko == RB_C) nC = nC - B.fc; else nH = nH - B.fc;
        nC = nC > R(4.0) ? R(4.0) : nC;
        nH = nH > R(4.0) ? R(4.0) : nH;
        const double4 cf = coef[s];
        const Rb4<double> bs = bo[s];
        double fc3[3] = {0, 0, 0}, fo3[3] = {0, 0, 0};
        rb_neighbours(D, c, kc, ijpot, B.len, R(1.0) / B.len, B.n, R(1.0), offsets[c], bcnt[c], s, bond, bij,
                      offsets, bcnt, el, nC + nH, (R)cf.x, (R)cf.y, (R)bs.y, (R)bs.z, (R)bs.w, fc3, fo3, w, facc,
                      bad);
        fx_add3(facc, c, fc3, bad);
        fx_add3(facc, o, fo3, bad);
        if (wacc) {
            for (int q = 0; q < 9; ++q) {
                fx_add(wacc + 9 * (size_t)c + q, w[q] / 2, bad);
                fx_add(wacc + 9 * (size_t)o + q, w[q] / 2, bad);
            }
        }
        if (wdir)
            for (int q = 0; q < 9; ++q) wdir[9 * (size_t)t2 + q] = w[q];
    }
    if (bad) atomicOr(flags, ((bad & 1) ? DERR_FIXED_RANGE : 0) | ((bad & 2) ? DERR_BOND_OVERFLOW : 0));
    rb_block_record(0.0, w, partial, blockIdx.x);
}

// Pass 3 without atomics (no per-atom or per-bond virial requested): a
// group of G lanes per centre slot c, lane = bond c->x.  The coordination
// and conjugation terms of every directed bond s = c->o act on the other
// neighbours k through scalars of s and the bond c->k only, so they are sums
// over the centre's bonds:
//   F_k = -dfc_ck n_ck [sum_{s != ck} (dE/dN_s + dfb_s dP/dN_type(k),s)
//                       + [k = C] F(x_ck) sum_{s != ck} dE/dN^conj_s]
// and the second-neighbour conjugation forces through k reduce to the bond
// scalar B_ck = [k = C] fc_ck dF(x_ck) sum_{s != ck} dE/dN^conj_s, applied in
// k_rb_conj_force.  The bond-order angular terms of each pair (s, k) are
// evaluated per lane s with the centre's bonds k walked in turn, the force on
// k summed over the group into lane k.  Per bond c->x the force on x goes to
// fnb[slot]; the centre's own force to own[c]; forces are gathered by the
// atoms afterwards (k_rb_gather), so the pass is free of atomics and its
// result does not depend on scheduling.
template <typename R, int G>
__device__ __forceinline__ void
rb_centre_chunk(int chunk, const int *__restrict__ clist, int ncls, const RbDevT<R> &D,
                const int *__restrict__ el, const int *__restrict__ offsets,
                const int *__restrict__ bcnt, const RbBondT<R> *__restrict__ bond, const int2 *__restrict__ bij,
                const double2 *__restrict__ nn, const Rb4<double> *__restrict__ bo, const double4 *__restrict__ coef,
                double *__restrict__ fnb, double *__restrict__ own, double *__restrict__ bconj, double *w)
{
    const int l = threadIdx.x % G;
    // centres chunk * 256 / G .. of the class list (ncls of them)
    const int gi = chunk * (256 / G) + threadIdx.x / G;
    const bool cv = gi < ncls;
    const int c = cv ? clist[gi] : 0;
    const int nb = cv ? bcnt[c] : 0;
    const int beg = cv ? offsets[c] : 0;
    const int kc = cv ? el[c] : 0;
    const bool has = l < nb;
    const int b = beg + l;
    RbBondT<R> B{};
    int kx = 0;
    double dfdn = 0.0, dfdncn = 0.0, dfb = 0.0, pc = 0.0, ph = 0.0;
    bool sv = false;      // the bond is a directed bond of a half pair (coef set)
    if (has) {
        B = bond[b];
        kx = el[bij[b].x];
        const double4 cf = coef[b];
        sv = cf.w != 0.0;
        if (sv) {
            const Rb4<double> bs = bo[b];
            dfdn = cf.x;
            dfdncn = cf.y;
            dfb = bs.y;
            pc = bs.y * bs.z;
            ph = bs.y * bs.w;
        }
    }
    const double S1 = group_sum<G>(dfdn), S2c = group_sum<G>(pc), S2h = group_sum<G>(ph);
    const double S3 = group_sum<G>(dfdncn);
    double fo[3] = {0, 0, 0}, fcen[3] = {0, 0, 0};
    // coordination / conjugation / P on this lane's neighbour x
    if (has) {
        const bool xC = kx == RB_C;
        const double A = (S1 - dfdn) + (xC ? (S2c - pc) + (double)B.fx * (S3 - dfdncn) : (S2h - ph));
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double fk = -A * (double)(B.dfc * B.n[x]);
            fo[x] = fk;
            fcen[x] = -fk;
        }
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x) w[x + 3 * y] -= (double)(B.len * B.n[x]) * fo[y];
        bconj[b] = xC ? (double)(B.fc * B.dfx) * (S3 - dfdncn) : 0.0;
    }
    // bond-order angular terms: this lane is s = c->o, k walks the bonds
    int xmax = nb;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) xmax = max(xmax, __shfl_xor(xmax, o, 64));
    R nt = R(0.0);
    int ijpot = 0;
    if (has) {
        const double2 nnc = nn[c];
        R nC = nnc.x, nH = nnc.y;
        if (kx == RB_C) nC = nC - B.fc; else nH = nH - B.fc;
        nC = nC > R(4.0) ? R(4.0) : nC;
        nH = nH > R(4.0) ? R(4.0) : nH;
        nt = nC + nH;
        ijpot = rb_pair(kc, kx);
    }
    const R rlij = B.len, rlijr = has ? R(1.0) / B.len : R(0.0);
    const R dfbR = (R)dfb;
    double fkx[3] = {0, 0, 0};
    for (int k = 0; k < xmax; ++k) {
        double dbk3[3] = {0, 0, 0};
        // bond k of the centre from the lane of the group that holds it
        // (cross-lane moves instead of a global load per step)
        RbBondT<R> K{};
#pragma unroll
        for (int x = 0; x < 3; ++x) K.n[x] = __shfl(B.n[x], k, G);
        K.len = __shfl(B.len, k, G);
        K.fc = __shfl(B.fc, k, G);
        K.dfc = __shfl(B.dfc, k, G);
        const int kxk = __shfl(kx, k, G);
        if (sv && k < nb && k != l) {
            const int ikpot = rb_pair(kc, kxk);
            const R rlik = K.len;
            if (rlik < D.cut_h[ikpot]) {
                R qf, qd, gf, gd, dgdn;
                rb_h(D, ijpot, ikpot, rlij - rlik, qf, qd);
                R costh = R(0.0);
                costh = costh + K.n[0] * B.n[0];
                costh = costh + K.n[1] * B.n[1];
                costh = costh + K.n[2] * B.n[2];
                rb_g(D, kc, costh, nt, gf, gd, dgdn);
                R dk[3], dd = R(0.0);
#pragma unroll
                for (int x = 0; x < 3; ++x) dk[x] = K.n[x] * rlik - B.n[x] * rlij;
                dd = dd + dk[0] * dk[0];
                dd = dd + dk[1] * dk[1];
                dd = dd + dk[2] * dk[2];
                const R dis = sqrt(dd);
                const R rdis = R(1.0) / dis, rlikr = R(1.0) / rlik;
#pragma unroll
                for (int x = 0; x < 3; ++x) dk[x] = dk[x] * rdis;
                const R dcs_io = rlikr - costh * rlijr;
                const R dcs_ik = rlijr - costh * rlikr;
                const R dcs_ok = -dis * rlijr * rlikr;
                const R dzfac = K.fc * gd * qf;
                const R dzdro = gf * K.fc * qd;
                const R dzdrk = gf * (K.dfc * qf - K.fc * qd);
                R dfo[3], dbk[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const R u = B.n[x];
                    const R dcs_c = -dcs_io * u - dcs_ik * K.n[x];
                    const R dcs_o = dcs_io * u - dcs_ok * dk[x];
                    const R dcs_k = dcs_ik * K.n[x] + dcs_ok * dk[x];
                    fcen[x] += (double)(-dfbR * (-dzdro * u - dzdrk * K.n[x] + dzfac * dcs_c));
                    dfo[x] = dzdro * u + dzfac * dcs_o;
                    fo[x] += (double)(-dfbR * dfo[x]);
                    dbk[x] = dzdrk * K.n[x] + dzfac * dcs_k;
                    dbk3[x] = (double)(-dfbR * dbk[x]);
                }
                const R rco[3] = {rlij * B.n[0], rlij * B.n[1], rlij * B.n[2]};
                const R rik[3] = {rlik * K.n[0], rlik * K.n[1], rlik * K.n[2]};
#pragma unroll
                for (int y = 0; y < 3; ++y)
#pragma unroll
                    for (int x = 0; x < 3; ++x)
                        w[x + 3 * y] += dfb * (double)(rco[x] * dfo[y] + rik[x] * dbk[y]);
            }
        }
        // the force on k, summed over the lanes s of the group
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double v = group_sum<G>(dbk3[x]);
            if (l == k) fkx[x] = v;
        }
    }
    if (has)
        for (int x = 0; x < 3; ++x) fnb[3 * (size_t)b + x] = fo[x] + fkx[x];
#pragma unroll
    for (int x = 0; x < 3; ++x) fcen[x] = group_sum<G>(fcen[x]);
    if (cv && l == 0)
        for (int x = 0; x < 3; ++x) own[3 * (size_t)c + x] = fcen[x];
}

// The second-neighbour conjugation forces: a group of G lanes per atom k,
// lane = bond k->m; with B_ck from k_rb_centre on the bonds c->k (the
// reverse slots), the force on m is -(sum_c B_ck - B_mk) dfc_km n_km (the
// centre c = m itself excluded, bop_kernel_rebo2.f90:2405-2520), its
// opposite on k.
template <typename R, int G>
__device__ __forceinline__ void
rb_conj_force_chunk(int chunk, const int *__restrict__ clist, int ncls, const int *__restrict__ offsets,
                    const int *__restrict__ bcnt, const RbBondT<R> *__restrict__ bond, const int *__restrict__ rev,
                    const double *__restrict__ bconj, double *__restrict__ fcj, double *__restrict__ own2, double *w)
{
    const int l = threadIdx.x % G;
    const int gi = chunk * (256 / G) + threadIdx.x / G;
    const bool cv = gi < ncls;
    const int k = cv ? clist[gi] : 0;
    const int nb = cv ? bcnt[k] : 0;
    const bool has = l < nb;
    const int b = (cv ? offsets[k] : 0) + l;
    const double Bv = has ? bconj[rev[b]] : 0.0;
    const double Bs = group_sum<G>(Bv);
    double fk[3] = {0, 0, 0};
    if (has) {
        const RbBondT<R> M = bond[b];
        const double a = Bs - Bv;
        double d2[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            d2[x] = -a * (double)(M.dfc * M.n[x]);
            fk[x] = -d2[x];
            fcj[3 * (size_t)b + x] = d2[x];
        }
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x) w[x + 3 * y] -= (double)(M.len * M.n[x]) * d2[y];
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) fk[x] = group_sum<G>(fk[x]);
    if (cv && l == 0)
        for (int x = 0; x < 3; ++x) own2[3 * (size_t)k + x] = fk[x];
}

// The per-centre passes of all three width classes in one launch: block
// chunks of 16 centres of 9-16 bonds, then of 32 centres of 5-8 bonds,
// then of 64 centres of up to 4 bonds (the class lists of
// k_rb_class_list), dealt to the blocks in turn, so the few long chunks of
// the wide centres run beside the others.  A block's virial sums its chunks
// in order, so the records do not depend on scheduling.  k_rb_conj_forces:
// the second-neighbour conjugation forces (needs every bond scalar of the
// centre pass).
__device__ __forceinline__ void rb_chunks(const int *__restrict__ ccount, int &c4, int &c8, int &c16)
{
    c4 = (ccount[0] + 63) / 64;
    c8 = (ccount[1] + 31) / 32;
    c16 = (ccount[2] + 15) / 16;
}

template <typename R>
__global__ void __launch_bounds__(256)
k_rb_centres(int nat, const int *__restrict__ clist, const int *__restrict__ ccount,
             const RbDevT<R> *__restrict__ Dp, const int *__restrict__ el, const int *__restrict__ offsets,
             const int *__restrict__ bcnt, const RbBondT<R> *__restrict__ bond, const int2 *__restrict__ bij,
             const double2 *__restrict__ nn, const Rb4<double> *__restrict__ bo, const double4 *__restrict__ coef,
             double *__restrict__ fnb, double *__restrict__ own, double *__restrict__ bconj,
             double *__restrict__ partial)
{
    int c4, c8, c16;
    rb_chunks(ccount, c4, c8, c16);
    double w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    for (int ch = blockIdx.x; ch < c4 + c8 + c16; ch += gridDim.x) {
        if (ch < c16)
            rb_centre_chunk<R, 16>(ch, clist + 2 * (size_t)nat, ccount[2], *Dp, el, offsets, bcnt, bond, bij, nn,
                                   bo, coef, fnb, own, bconj, w);
        else if (ch < c16 + c8)
            rb_centre_chunk<R, 8>(ch - c16, clist + nat, ccount[1], *Dp, el, offsets, bcnt, bond, bij, nn, bo,
                                  coef, fnb, own, bconj, w);
        else
            rb_centre_chunk<R, 4>(ch - c16 - c8, clist, ccount[0], *Dp, el, offsets, bcnt, bond, bij, nn, bo, coef,
                                  fnb, own, bconj, w);
    }
    rb_block_record(0.0, w, partial, blockIdx.x);
}

template <typename R>
__global__ void __launch_bounds__(256)
k_rb_conj_forces(int nat, const int *__restrict__ clist, const int *__restrict__ ccount,
                 const int *__restrict__ offsets, const int *__restrict__ bcnt, const RbBondT<R> *__restrict__ bond,
                 const int *__restrict__ rev, const double *__restrict__ bconj, double *__restrict__ fcj,
                 double *__restrict__ own2, double *__restrict__ partial)
{
    int c4, c8, c16;
    rb_chunks(ccount, c4, c8, c16);
    double w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    for (int ch = blockIdx.x; ch < c4 + c8 + c16; ch += gridDim.x) {
        if (ch < c16)
            rb_conj_force_chunk<R, 16>(ch, clist + 2 * (size_t)nat, ccount[2], offsets, bcnt, bond, rev, bconj, fcj,
                                       own2, w);
        else if (ch < c16 + c8)
            rb_conj_force_chunk<R, 8>(ch - c16, clist + nat, ccount[1], offsets, bcnt, bond, rev, bconj, fcj, own2,
                                      w);
        else
            rb_conj_force_chunk<R, 4>(ch - c16 - c8, clist, ccount[0], offsets, bcnt, bond, rev, bconj, fcj, own2,
                                      w);
    }
    rb_block_record(0.0, w, partial, blockIdx.x);
}

// f(atom order[p]) += own + own2 + the forces its neighbours' groups left on
// the bonds towards p (reverse slots) + the pair and torsion forces on p by
// its own bonds (fpr) + the fixed-point sums (torsion forces on k and l)
__global__ void k_rb_gather(int nat, const int *__restrict__ order, const int *__restrict__ offsets,
                            const int *__restrict__ bcnt, const int *__restrict__ rev, const double *__restrict__ own,
                            const double *__restrict__ own2, const double *__restrict__ fnb,
                            const double *__restrict__ fcj, const double *__restrict__ fpr,
                            unsigned long long *__restrict__ facc, unsigned long long *__restrict__ eacc,
                            double *__restrict__ f, double *__restrict__ epa)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    double a[3];
    for (int x = 0; x < 3; ++x) a[x] = own[3 * (size_t)p + x] + own2[3 * (size_t)p + x];
    const int beg = offsets[p], nb = bcnt[p];
    for (int b = beg; b < beg + nb; ++b) {
        const size_t r = (size_t)rev[b];
        for (int x = 0; x < 3; ++x) a[x] += (fnb[3 * r + x] + fcj[3 * r + x]) + fpr[3 * (size_t)b + x];
    }
    const size_t i = (size_t)order[p];
    for (int x = 0; x < 3; ++x) {
        f[3 * i + x] += a[x] + (double)(long long)facc[3 * (size_t)p + x] * RB_FX_INV;
    }
    if (epa) {
        epa[i] += (double)(long long)eacc[p] * RB_FX_INV;
    }
}

// per-bond virial: the two directed bonds' neighbour-pass virials of each
// half pair added to its list entry, in a fixed order
__global__ void k_rb_wpb(const int *__restrict__ list, const int *__restrict__ count, const int *__restrict__ bsrc,
                         const double *__restrict__ wdir, double *__restrict__ wpb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *count) return;
    const size_t e = (size_t)bsrc[list[t]];
    for (int q = 0; q < 9; ++q) wpb[9 * e + q] += wdir[9 * (size_t)(2 * t) + q] + wdir[9 * (size_t)(2 * t + 1) + q];
}

__global__ void k_rb_finish(int nat, const int *__restrict__ order, unsigned long long *__restrict__ facc,
                            unsigned long long *__restrict__ eacc, unsigned long long *__restrict__ wacc,
                            double *__restrict__ f, double *__restrict__ epa, double *__restrict__ wpa)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    const size_t i = (size_t)order[p];
    for (int x = 0; x < 3; ++x) {
        f[3 * i + x] += (double)(long long)facc[3 * (size_t)p + x] * RB_FX_INV;
    }
    if (epa) {
        epa[i] += (double)(long long)eacc[p] * RB_FX_INV;
    }
    if (wpa)
        for (int q = 0; q < 9; ++q) {
            wpa[9 * i + q] += (double)(long long)wacc[9 * (size_t)p + q] * RB_FX_INV;
        }
}

}  // namespace agpu

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
namespace agpu {

static void rebo2_free(void *p) { delete static_cast<Rebo2State *>(p); }

// rebo2_bind_to (rebo2_module.f90:57-110): element map C, H and the
// interaction ranges r2(CC), r2(CH), r2(HH) for the pairs present
int rebo2_bind(agpu_potential *pot)
{
    Rebo2State *st = static_cast<Rebo2State *>(pot->impl);
    agpu_particles *p = pot->p;
    for (int z = 0; z < 128; ++z) st->z2db[z] = -1;
    st->z2db[6] = RB_C;
    st->z2db[1] = RB_H;
    bool hasC = false, hasH = false;
    for (int z : p->h_Z) {
        hasC = hasC || z == 6;
        hasH = hasH || z == 1;
    }
    double cut = 0.0;
    const RbDev &D = st->hdev;
    if (hasC) {
        AGPU_CHECK(agpu_neighbors_request_interaction_range(pot->nl, D.cut_h[RB_CC]));
        cut = D.cut_h[RB_CC];
    }
    if (hasC && hasH) {
        AGPU_CHECK(agpu_neighbors_request_interaction_range(pot->nl, D.cut_h[RB_CH]));
        cut = cut > D.cut_h[RB_CH] ? cut : D.cut_h[RB_CH];
    }
    if (hasH) {
        AGPU_CHECK(agpu_neighbors_request_interaction_range(pot->nl, D.cut_h[RB_HH]));
        cut = cut > D.cut_h[RB_HH] ? cut : D.cut_h[RB_HH];
    }
    st->cutoff = cut;
    AGPU_CHECK(map_elements(pot, st->z2db, 128));
    st->Z_rev = p->Z_rev;
    return AGPU_OK;
}

// bonds, conjugation, bond orders (FP64 in both forms), the half-pair list
// and the pair kernel in precision R: FP64, or the mixed form (FP32 copies of
// the bond records for the per-neighbour loops, FP64 records for the pair's
// radial terms and the torsion, 64-bit fixed-point accumulation); the slot
// arrays and accumulators are ready
template <typename R>
static int rebo2_kernels(Rebo2State *st, agpu_potential *pot, const agpu_eval_args *a, const int *bsrc,
                         unsigned long long *eacc, unsigned long long *wacc, int *nblk_out, bool *gathered)
{
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    hipStream_t s = stream();
    const int nat = p->nat;
    const int64_t total = nl->total;
    const bool f64 = sizeof(R) == 8;
    AGPU_CHECK(st->bond.reserve(sizeof(RbBond) * (total + 1)));
    if (!f64) AGPU_CHECK(st->bond32.reserve(sizeof(RbBondT<float>) * (total + 1)));
    AGPU_CHECK(st->bo.reserve(sizeof(Rb4<double>) * (total + 1)));
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const RbDev *D64 = st->dev.as<RbDev>();
    const RbDevT<R> *D = f64 ? reinterpret_cast<const RbDevT<R> *>(st->dev.ptr)
                             : reinterpret_cast<const RbDevT<R> *>(st->dev32.ptr);
    const R *tab = f64 ? reinterpret_cast<const R *>(st->tab.ptr) : reinterpret_cast<const R *>(st->tab32.ptr);
    const int nb = (nat + 255) / 256;
    const int *order = nl->idx_sorted.as<int>();
    const bool lammps = p->natloc >= 0;
    const int *offs = nl_rows(nl).off;
    const int *els = st->el_s.as<int>();
    RbBond *bond = st->bond.as<RbBond>();
    RbBondT<float> *bond32 = f64 ? nullptr : st->bond32.as<RbBondT<float>>();
    const RbBondT<R> *bondR = f64 ? reinterpret_cast<const RbBondT<R> *>(bond)
                                  : reinterpret_cast<const RbBondT<R> *>(bond32);
    Rb4<double> *bo = st->bo.as<Rb4<double>>();
    int *dstats = st->stats.as<int>();
    // neighbour pass: per centre groups and an owner gather when no per-atom
    // or per-bond virial is asked for (those split each pair's virial over
    // its atoms and bond, which the centre sums do not keep); otherwise a
    // thread per directed bond with fixed-point accumulation.  Centres are
    // taken by groups of 4 lanes (up to 4 bonds), then 8 and 16 lanes for
    // the wider rows, one launch per width class over that class's list
    const int mr = nl->maxrow;
    const bool by_centre = !wacc && !a->wpot_per_bond && mr <= 16 && nat < (1 << RB_CLS_BITS);
    unsigned long long *cflag = nullptr, *cpos = nullptr;
    // pair and torsion forces on i, j by directed bond (the owner gather sums
    // them), instead of fixed-point adds, on the per-centre pass
    double *fpr = nullptr;
    if (by_centre) {
        AGPU_CHECK(st->fpr.reserve(sizeof(double) * 3 * (size_t)(total + 1)));
        fpr = st->fpr.as<double>();
        AGPU_CHECK(st->cls.reserve(sizeof(unsigned long long) * 2 * (size_t)(nat + 1) + sizeof(int) * (3 * (size_t)nat + 4)));
        cflag = st->cls.as<unsigned long long>();
        cpos = cflag + nat + 1;
    }
    AGPU_LAUNCH("rebo2_bonds", k_rb_bonds, dim3(nb), dim3(256), 0, s, nat, order, els, C, D64, nl_rows(nl), offs,
                nl->offsets.as<int>(), bond, bond32, st->bij.as<int2>(), st->bown.as<int>(), st->bcnt.as<int>(),
                st->nn.as<double2>(), const_cast<int *>(bsrc), cflag);
    int *clist = nullptr, *ccount = nullptr;
    if (by_centre) {
        clist = reinterpret_cast<int *>(cpos + nat + 1);
        ccount = clist + 3 * (size_t)nat;
        size_t tmp = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, cflag, cpos, 0ull, (size_t)nat + 1,
                                         rocprim::plus<unsigned long long>(), s));
        AGPU_CHECK(st->scan_tmp.reserve(tmp + 16));
        AGPU_HIP(rocprim::exclusive_scan(st->scan_tmp.ptr, tmp, cflag, cpos, 0ull, (size_t)nat + 1,
                                         rocprim::plus<unsigned long long>(), s));
        AGPU_LAUNCH("rebo2_classes", k_rb_class_list, dim3((nat + 256) / 256), dim3(256), 0, s, nat,
                    st->bcnt.as<int>(), cpos, clist, ccount);
    }
    AGPU_LAUNCH("rebo2_conj", k_rb_conj, dim3(nb), dim3(256), 0, s, nat, els, offs, st->bcnt.as<int>(),
                st->bij.as<int2>(), st->nn.as<double2>(), bond, bond32, st->nconj.as<double>());
    const int ntot = (int)total;
    int nblk = (ntot + 255) / 256;
    if (nblk < 1) nblk = 1;
    *nblk_out = nblk;
    // per-centre pass grid: the chunks of the class sizes seen on the
    // previous call (+25 %), else of every atom in each class the longest
    // row allows, at most 2048 blocks striding over the chunks
    int nbw[3] = {(nat + 63) / 64, mr > 4 ? (nat + 31) / 32 : 0, mr > 8 ? (nat + 15) / 16 : 0};
    for (int w = 0; w < 3; ++w) {
        const int cpb = 64 >> w;
        if (st->chint[w] >= 0) nbw[w] = std::min(nbw[w], (st->chint[w] + st->chint[w] / 4 + cpb - 1) / cpb + 8);
    }
    const int nbc = std::min(nbw[0] + nbw[1] + nbw[2], 2048);
    // block records: pair, torsion, then the neighbour pass (2 nblk, or
    // centre and conjugation passes)
    // torsion groups: 64 per block, the grid striding over the pairs
    // torsion groups: 256 / G per block, at most 2^17 groups, the grid
    // striding over the pairs
    const int tg = st->tor_g, tgpb = 256 / tg;
    const int ntb = std::max(1, std::min((ntot + tgpb - 1) / tgpb, (1 << 17) / tgpb));
    const size_t nrec = (size_t)nblk + (size_t)ntb + (by_centre ? 2 * (size_t)nbc : 2 * (size_t)nblk);
    AGPU_CHECK(st->partial.reserve(sizeof(double) * 10 * nrec));
    AGPU_CHECK(st->coef.reserve(sizeof(double4) * (total + 1)));
    AGPU_CHECK(st->rev.reserve(sizeof(int) * (total + 1)));
    AGPU_CHECK(st->half.reserve(sizeof(int) * (total + 1)));
    AGPU_CHECK(st->hpos.reserve(sizeof(int) * (total + 1)));
    AGPU_CHECK(st->list.reserve(sizeof(int) * (total + 2)));
    int *count = st->list.as<int>() + total + 1;
    if (ntot > 0) {
        AGPU_LAUNCH("rebo2_bo", k_rb_bo<double>, dim3(nblk), dim3(256), 0, s, ntot, D64, st->tab.as<double>(),
                    els, offs, st->bcnt.as<int>(), st->bown.as<int>(), bond, st->bij.as<int2>(),
                    st->nn.as<double2>(), bo, st->rev.as<int>(), st->half.as<int>(), p->natloc,
                    lammps ? st->tag_s.as<int>() : nullptr, order, dstats, st->coef.as<double4>(), fpr);
        size_t tmp = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, st->half.as<int>(), st->hpos.as<int>(), 0, (size_t)ntot,
                                         rocprim::plus<int>(), s));
        AGPU_CHECK(st->scan_tmp.reserve(tmp + 16));
        AGPU_HIP(rocprim::exclusive_scan(st->scan_tmp.ptr, tmp, st->half.as<int>(), st->hpos.as<int>(), 0,
                                         (size_t)ntot, rocprim::plus<int>(), s));
        AGPU_LAUNCH("rebo2_list", k_rb_list, dim3(nblk), dim3(256), 0, s, ntot, st->half.as<int>(),
                    st->hpos.as<int>(), st->list.as<int>(), count);
    } else {
        AGPU_CHECK(zero_dev(count, sizeof(int)));
    }
    // pass records: pair (nblk blocks), torsion (nblk), neighbours (2 nblk)
    const bool tors = st->hdev.with_dihedral != 0;
    if (tors) {
        AGPU_CHECK(st->tor.reserve(sizeof(Rb4<R>) * (total + 1)));
        AGPU_CHECK(st->tflag.reserve(sizeof(int) * (total + 1)));
        AGPU_CHECK(st->tpos.reserve(sizeof(int) * (total + 1)));
        AGPU_CHECK(st->tlist.reserve(sizeof(int) * (total + 2)));
    }
    Rb4<R> *tor = tors ? st->tor.as<Rb4<R>>() : nullptr;
    int *tflag = tors ? st->tflag.as<int>() : nullptr;
    int *tlist = tors ? st->tlist.as<int>() : nullptr;
    int *tcount = tors ? tlist + total + 1 : count;
    double4 *coef = st->coef.as<double4>();
    double *part = st->partial.as<double>();
    double *wdir = nullptr;
    if (a->wpot_per_bond) {
        AGPU_CHECK(st->wdir.reserve(sizeof(double) * 9 * (2 * (size_t)total + 2)));
        wdir = st->wdir.as<double>();
    }
    const int *list = st->list.as<int>();
    AGPU_LAUNCH(f64 ? "rebo2_pair" : "rebo2_pair_mixed", k_rb_pair<R>, dim3(nblk), dim3(256), 0, s, list, count,
                D, tab, els, offs, st->bcnt.as<int>(), st->bown.as<int>(), bondR, st->bij.as<int2>(),
                st->nn.as<double2>(), st->nconj.as<double>(), bo, st->rev.as<int>(), coef, tor, tflag, ntot,
                st->facc.as<unsigned long long>(), eacc, wacc, part, dstats, bsrc, a->epot_per_bond,
                a->f_per_bond, a->wpot_per_bond, D64, bond, fpr);
    if (tors && ntot > 0) {
        // the half pairs with T != 0, in list order (flags + scan)
        size_t tmp = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, tflag, st->tpos.as<int>(), 0, (size_t)ntot,
                                         rocprim::plus<int>(), s));
        AGPU_CHECK(st->scan_tmp.reserve(tmp + 16));
        AGPU_HIP(rocprim::exclusive_scan(st->scan_tmp.ptr, tmp, tflag, st->tpos.as<int>(), 0, (size_t)ntot,
                                         rocprim::plus<int>(), s));
        AGPU_LAUNCH("rebo2_tlist", k_rb_list, dim3(nblk), dim3(256), 0, s, ntot, tflag, st->tpos.as<int>(),
                    tlist, tcount);
    }
    // without torsion (tor == nullptr) the pass only writes zero records
    auto ktor = tg == 16 ? k_rb_torsion<R, 16> : (tg == 8 ? k_rb_torsion<R, 8> : k_rb_torsion<R, 4>);
    AGPU_LAUNCH("rebo2_torsion", ktor, dim3(ntb), dim3(256), 0, s, list, tlist, tcount, offs,
                st->bcnt.as<int>(), st->bown.as<int>(), st->bij.as<int2>(), st->rev.as<int>(), tor, coef,
                st->facc.as<unsigned long long>(), eacc, wacc, part + 10 * (size_t)nblk, dstats, bsrc,
                a->epot_per_bond, a->f_per_bond, a->wpot_per_bond, D64, bond, fpr);
    if (by_centre) {
        AGPU_CHECK(st->fnb.reserve(sizeof(double) * 3 * (total + 1)));
        AGPU_CHECK(st->fcj.reserve(sizeof(double) * 3 * (total + 1)));
        AGPU_CHECK(st->bconj.reserve(sizeof(double) * (total + 1)));
        AGPU_CHECK(st->own.reserve(sizeof(double) * 6 * (nat + 1)));
        double *own = st->own.as<double>(), *own2 = own + 3 * (size_t)nat;
        double *pc = part + 10 * ((size_t)nblk + ntb), *pj = pc + 10 * (size_t)nbc;
        AGPU_LAUNCH(f64 ? "rebo2_centre" : "rebo2_centre_mixed", k_rb_centres<R>, dim3(nbc), dim3(256), 0, s, nat,
                    clist, ccount, D, els, offs, st->bcnt.as<int>(), bondR, st->bij.as<int2>(), st->nn.as<double2>(),
                    bo, coef, st->fnb.as<double>(), own, st->bconj.as<double>(), pc);
        // the bond scalars of every centre are in place before any conjugation force
        AGPU_LAUNCH(f64 ? "rebo2_conj_force" : "rebo2_conj_force_mixed", k_rb_conj_forces<R>, dim3(nbc), dim3(256), 0,
                    s, nat, clist, ccount, offs, st->bcnt.as<int>(), bondR, st->rev.as<int>(),
                    st->bconj.as<double>(), st->fcj.as<double>(), own2, pj);
        // the class sizes, read back with the flags: the next call's grids
        AGPU_HIP(hipMemcpyAsync(pot->stats.h + 8, ccount, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        st->chint_pending = true;
        AGPU_LAUNCH("rebo2_gather", k_rb_gather, dim3(nb), dim3(256), 0, s, nat, order, offs, st->bcnt.as<int>(),
                    st->rev.as<int>(), own, own2, st->fnb.as<double>(), st->fcj.as<double>(), fpr,
                    st->facc.as<unsigned long long>(), eacc, a->f, a->epot_per_at);
        *nblk_out = nblk + ntb + 2 * nbc;
        *gathered = true;
        return AGPU_OK;
    }
    AGPU_LAUNCH(f64 ? "rebo2_nbrs" : "rebo2_nbrs_mixed", k_rb_nbrs<R>, dim3(2 * nblk), dim3(256), 0, s, list,
                count, D, els, offs, st->bcnt.as<int>(), st->bown.as<int>(), bondR, st->bij.as<int2>(),
                st->nn.as<double2>(), bo, st->rev.as<int>(), coef, st->facc.as<unsigned long long>(), wacc,
                part + 10 * ((size_t)nblk + ntb), dstats, wdir);
    if (wdir)
        AGPU_LAUNCH("rebo2_wpb", k_rb_wpb, dim3(nblk), dim3(256), 0, s, list, count, bsrc, wdir,
                    a->wpot_per_bond);
    *nblk_out = 3 * nblk + ntb;
    *gathered = false;
    return AGPU_OK;
}

int rebo2_variant(agpu_potential *pot, int k, int *g)
{
    constexpr int TG[3] = {4, 8, 16};
    Rebo2State *st = static_cast<Rebo2State *>(pot->impl);
    if (k >= 0 && k < 3) st->tor_g = TG[k];
    if (g) *g = st->tor_g;
    return 3;
}

int rebo2_compute(agpu_potential *pot, const agpu_eval_args *a)
{
    Rebo2State *st = static_cast<Rebo2State *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    if (a->mask)
        return set_error(AGPU_ERR_UNSUPPORTED, "REBO2: the reference kernel takes no mask");
    if (st->Z_rev != p->Z_rev) AGPU_CHECK(rebo2_bind(pot));
    // the rows are read where the builder left them (slot order)
    AGPU_CHECK(nl_update(nl, p, 0, false));
    const int nat = p->nat;
    hipStream_t s = stream();
    if (nat == 0) return zero_dev(a->epot, sizeof(double));
    const int64_t total = nl->total;
    AGPU_CHECK(st->bij.reserve(sizeof(int2) * (total + 1)));
    AGPU_CHECK(st->bown.reserve(sizeof(int) * (total + 1)));
    AGPU_CHECK(st->bcnt.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(st->nn.reserve(sizeof(double2) * (nat + 1)));
    AGPU_CHECK(st->nconj.reserve(sizeof(double) * (nat + 1)));
    AGPU_CHECK(st->facc.reserve(sizeof(unsigned long long) * 3 * (nat + 1)));
    AGPU_CHECK(st->stats.reserve(sizeof(int) * 4));
    AGPU_CHECK(pot->stats.init());
    unsigned long long *eacc = nullptr, *wacc = nullptr;
    if (a->epot_per_at) {
        AGPU_CHECK(st->eacc.reserve(sizeof(unsigned long long) * (nat + 1)));
        eacc = st->eacc.as<unsigned long long>();
    }
    if (a->wpot_per_at) {
        AGPU_CHECK(st->wacc.reserve(sizeof(unsigned long long) * 9 * (nat + 1)));
        wacc = st->wacc.as<unsigned long long>();
    }
    int *dstats = st->stats.as<int>();
    const int *el = pot->d_el.as<int>();
    const int nb = (nat + 255) / 256;
    const bool per_bond = a->epot_per_bond || a->f_per_bond || a->wpot_per_bond;
    if (per_bond) AGPU_CHECK(st->bsrc.reserve(sizeof(int) * (total + 1)));
    // the bonds' entries in the caller-order list (the reference's indexing)
    if (per_bond) AGPU_CHECK(nl_caller_offsets(nl));
    int *bsrc = per_bond ? st->bsrc.as<int>() : nullptr;
    // slot order: elements, tags; the bond slots of slot p are its row's
    // entries (the list's slot-order offsets); accumulators, flags and
    // energy cleared in the same pass
    const int *order = nl->idx_sorted.as<int>();
    const bool lammps = p->natloc >= 0;
    AGPU_CHECK(st->el_s.reserve(sizeof(int) * (nat + 1)));
    if (lammps) AGPU_CHECK(st->tag_s.reserve(sizeof(int) * (nat + 1)));
    AGPU_LAUNCH("rebo2_slots", k_rb_slots, dim3((std::max(nat, 4) + 255) / 256), dim3(256), 0, s, nat, order, el,
                lammps ? p->d_tag.as<int>() : nullptr, st->el_s.as<int>(), lammps ? st->tag_s.as<int>() : nullptr,
                st->facc.as<unsigned long long>(), eacc, wacc, dstats, a->epot);
    int nblk = 0;
    bool gathered = false;
    if (pot->precision == AGPU_PREC_MIXED)
        AGPU_CHECK(rebo2_kernels<float>(st, pot, a, bsrc, eacc, wacc, &nblk, &gathered));
    else
        AGPU_CHECK(rebo2_kernels<double>(st, pot, a, bsrc, eacc, wacc, &nblk, &gathered));
    if (!gathered)
        AGPU_LAUNCH("rebo2_finish", k_rb_finish, dim3(nb), dim3(256), 0, s, nat, order,
                    st->facc.as<unsigned long long>(), eacc, wacc, a->f, a->epot_per_at, a->wpot_per_at);
    AGPU_CHECK(reduce_ev(st->partial.as<double>(), nblk, 10, a->epot, a->wpot, nullptr));
    AGPU_HIP(hipMemcpyAsync(pot->stats.h, dstats, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (st->chint_pending) {
        for (int w = 0; w < 3; ++w) st->chint[w] = pot->stats.h[8 + w];
        st->chint_pending = false;
    }
    if (pot->stats.h[0] & DERR_FIXED_RANGE)
        return set_error(AGPU_ERR_DOMAIN, "REBO2: a force contribution exceeds the fixed-point range (2^21 eV/A)");
    if (pot->stats.h[0] & DERR_BOND_OVERFLOW)
        return set_error(AGPU_ERR_NEB, "REBO2: bond without its reverse bond (inconsistent list)");
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_rebo2_create(agpu_potential **pot, int flags)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    if (flags & ~(AGPU_REBO2_ZERO_TABLES | AGPU_REBO2_DIHEDRAL))
        return set_error(AGPU_ERR_ARG, "REBO2: unknown flag");
    Rebo2State *st = new Rebo2State();
    st->zero_tables = (flags & AGPU_REBO2_ZERO_TABLES) != 0;
    st->with_dihedral = (flags & AGPU_REBO2_DIHEDRAL) != 0;
    int rc = rebo2_setup(st);
    if (rc != AGPU_OK) {
        delete st;
        return rc;
    }
    agpu_potential *P = new agpu_potential();
    P->kind = AGPU_KIND_REBO2;
    P->impl = st;
    P->free_impl = rebo2_free;
    P->cutoff = st->hdev.cut_h[RB_CC];
    *pot = P;
    return AGPU_OK;
}
