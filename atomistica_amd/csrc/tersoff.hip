// Tersoff bond-order potential: energy, forces and virial on gfx950.
//
// Semantics follow the reference's generic BOP kernel with the Tersoff forms
// (src/potentials/bop/bop_kernel.f90:563-1529, tersoff_func.f90:32-211);
// the work decomposition is this library's own:
//
//   k_bop_bonds   compact the neighbour row of every atom to its bonds
//                 (r < r2 of the element pair, loop1 of bop_kernel.f90:997-1058)
//   k_bop_rev     for each bond i->j (image n) the slot of j->i (image -n)
//   k_tersoff     a group of G lanes per atom, lane = bond ij: zeta_ij over the
//                 group's bonds through LDS, b_ij, then every triplet
//                 derivative once.  Forces on k are reduced across the group
//                 into the lane that owns bond ik (a fixed butterfly, no
//                 atomics); the result per bond slot is "force on the
//                 neighbour from the terms centred at i".
//   k_bop_gather  owner-computes assembly: f_a = f_own(a) + sum over a's bonds
//                 of the slot force its neighbour computed for a.
// Energy and virial are summed per block and reduced in a fixed order, so
// every result is bitwise reproducible run to run.
//
// Precision: R = double (FP64) or float (mixed: FP32 pair/triplet arithmetic,
// FP64 position differences and FP64 accumulation of forces, energy, virial).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

struct TersoffDev {
    int nel;
    double A[6], B[6], xi[6], lambda[6], mu[6], omega[6], mubo[6];
    int m[6];
    double beta[3], n[3], c[3], d[3], h[3];
    double r1[9], r2[9];          // 9: the ordered pairs of the Juslin family
    // derived
    double c2[3], d2[3], cd2[3], bn[3], e[3], fac[9];   // cd2 = c^2 / d^2
    // Kumagai family (kumagai_func.f90): g, bond order and h forms; A, B,
    // lambda (= lambda1), mu (= lambda2), h, r1, r2 are shared with the above
    int fam;                      // 0 Tersoff, 1 Kumagai, 2 Brenner
    double kc1[3], kc2[3], kc3[3], kc4[3], kc5[3], keta[3], kdelta[3];
    double kalpha[6];
    int kbeta[6];
    // Brenner family (brenner_func.f90, per element pair; h-term mu, m in
    // mubo, m above): V_R = VR_f exp(-expR (r - r0)), V_A = -VA_f exp(-expA
    // (r - r0)), g = gamma (1 + c^2/d^2 - c^2/(d^2 + (h + cos)^2)),
    // b = (1 + zeta^n)^(-1/2n)
    double br0[9], bVRf[9], bVAf[9], bexpR[9], bexpA[9], bgam[9], bc2[9], bd2[9], bh[9], bnn[9], bboe[9];
    double bcd2[9];               // c^2 / d^2
    // Juslin family (juslin_func.f90): the Brenner forms above over ordered
    // element pairs j + i nel, and h = omega exp((alpha dr)^m) per ordered
    // triplet k + nel (j + nel i) of centre i
    double jal[27], jom[27];
    int jm[27];
};

struct ScrDev;

struct TersoffState {
    agpu_tersoff_params prm;
    TersoffDev dev;
    double cutoff = 0.0;
    int z2db[128];
    uint64_t Z_rev = ~0ull;
    DevBuf bslot, bcnt, rev, fslot, fown, eslot, wslot, partial, stats, el_s;
    // screened variants: parameters (host copy of ScrDev) and per-entry /
    // per-record buffers
    bool scr = false;
    agpu_screening_params sp;
    DevBuf keep, rcnt, roff, fcv, rkn, rco, frec, fsl, nlrev, scan_tmp;
    int64_t nlrev_upd = -1;
    const agpu_neighbors *nlrev_nl = nullptr;
    // take the bond-compacting path even for rows of at most 64 entries
    // (agpu_potential_set_launch variant 1 of a bond-order class)
    bool compact = false;
    // lane-group width that a row of an unchecked build outgrew: calls over
    // unchecked builds use at least this width from then on
    int row_g = 0;
};

// Device pointers of the per-call scratch of the unscreened bond-order path:
// bond slots and counts, reverse slots, slot forces, own forces, per-block
// records (10 doubles per block), second-level reduction records (NULL: the
// library's own) and 4 flag words.
struct BopBufs {
    int *bslot, *bcnt, *rev, *stats;
    double *fslot, *fown, *partial, *red;
};

__host__ __device__ __forceinline__ int pair_index(int i, int j, int nel)
{
    int a = i + j * nel, b = j + i * nel;
    int c = i * (i + 1) / 2, d = j * (j + 1) / 2;
    return (a < b ? a : b) - (c < d ? c : d);
}

// element pair of bond i-j: ordered j + i nel for the Juslin family
// (PAIR_INDEX_NS, macros.inc:139), else the symmetric PAIR_INDEX
__host__ __device__ __forceinline__ int bop_pair(const TersoffDev &P, int i, int j)
{
    return P.fam == 3 ? j + i * P.nel : pair_index(i, j, P.nel);
}

static void derive(TersoffDev &D, const agpu_tersoff_params &p)
{
    std::memset(&D, 0, sizeof D);
    D.nel = p.nel;
    for (int k = 0; k < 6; ++k) {
        D.A[k] = p.A[k];
        D.B[k] = p.B[k];
        D.xi[k] = p.xi[k];
        D.lambda[k] = p.lambda[k];
        D.mu[k] = p.mu[k];
        D.omega[k] = p.omega[k];
        D.mubo[k] = p.mubo[k];
        D.m[k] = p.m[k];
        D.r1[k] = p.r1[k];
        D.r2[k] = p.r2[k];
        D.fac[k] = M_PI / (p.r2[k] - p.r1[k]);
    }
    for (int k = 0; k < 3; ++k) {
        D.beta[k] = p.beta[k];
        D.n[k] = p.n[k];
        D.c[k] = p.c[k];
        D.d[k] = p.d[k];
        D.h[k] = p.h[k];
        D.c2[k] = p.c[k] * p.c[k];
        D.d2[k] = p.d[k] * p.d[k];
        D.cd2[k] = D.d2[k] != 0.0 ? D.c2[k] / D.d2[k] : 0.0;
        D.bn[k] = std::pow(p.beta[k], p.n[k]);
        D.e[k] = -0.5 / p.n[k];
    }
}

// -------------------------------------------------------------------------
// bonds: compact each row to the pairs inside r2 of their element pair
// -------------------------------------------------------------------------
__global__ void k_bop_bonds(int nat, const double *__restrict__ r, const int *__restrict__ el,
                            CellArg C, TersoffDev P, const int *__restrict__ offsets,
                            const int2 *__restrict__ nbr, int *__restrict__ bslot,
                            int *__restrict__ bcnt, int *__restrict__ stats)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int eli = i < nat ? el[i] : -1;
    int nb = 0;
    if (eli >= 0) {
        const double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
        int beg = offsets[i], end = offsets[i + 1];
        for (int s = beg; s < end; ++s) {
            int2 e = nbr[s];
            int elj = el[e.x];
            if (elj < 0) continue;
            int a, b, c;
            dc_unpack(e.y, a, b, c);
            const double rj[3] = {r[3 * e.x], r[3 * e.x + 1], r[3 * e.x + 2]};
            double d[3];
            delta_r(C, ri, rj, a, b, c, d);
            double rl = dot3_nc(d);
            int t = bop_pair(P, eli, elj);
            if (rl < P.r2[t] * P.r2[t]) bslot[beg + nb++] = s;
        }
    }
    if (i < nat) bcnt[i] = nb;
    (void)stats;   // the bond maximum is reduced from bcnt afterwards
}

// Thread per (atom a, bond k): position of the reverse bond j->a (image -n)
// in j's compacted bond list.
__global__ void k_bop_rev(int nat, int G, const int *__restrict__ offsets, const int2 *__restrict__ nbr,
                          const int *__restrict__ bslot, const int *__restrict__ bcnt,
                          int *__restrict__ rev, int *__restrict__ stats)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int a = (int)(t / G), k = (int)(t % G);
    if (a >= nat || k >= bcnt[a]) return;
    const int s = bslot[offsets[a] + k];
    const int2 e = nbr[s];
    int x, y, z;
    dc_unpack(e.y, x, y, z);
    const int want = dc_pack(-x, -y, -z);
    const int j = e.x, jb = offsets[j], jn = bcnt[j];
    int found = -1;
    for (int q = 0; q < jn; ++q) {
        const int sj = bslot[jb + q];
        const int2 ej = nbr[sj];
        if (ej.x == a && ej.y == want) {
            found = sj;
            break;
        }
    }
    rev[s] = found;
    if (found < 0) atomicOr(stats + 1, DERR_BOND_OVERFLOW);
}

// -------------------------------------------------------------------------
// the bond-order kernel
// -------------------------------------------------------------------------
template <typename R> __device__ __forceinline__ R rexp(R x);
template <> __device__ __forceinline__ double rexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float rexp<float>(float x) { return expf(x); }
template <typename R> __device__ __forceinline__ R rpow(R x, R y);
template <> __device__ __forceinline__ double rpow<double>(double x, double y) { return pow(x, y); }
template <> __device__ __forceinline__ float rpow<float>(float x, float y) { return powf(x, y); }
template <typename R> __device__ __forceinline__ R rsqrt_(R x);
template <> __device__ __forceinline__ double rsqrt_<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float rsqrt_<float>(float x) { return sqrtf(x); }
template <typename R> __device__ __forceinline__ void rsincos(R x, R *s, R *c);
template <> __device__ __forceinline__ void rsincos<double>(double x, double *s, double *c) { sincos(x, s, c); }
template <> __device__ __forceinline__ void rsincos<float>(float x, float *s, float *c) { sincosf(x, s, c); }
template <typename R> __device__ __forceinline__ R rlog(R x);
template <> __device__ __forceinline__ double rlog<double>(double x) { return log(x); }
template <> __device__ __forceinline__ float rlog<float>(float x) { return logf(x); }
// x^y for x > 0 as exp(y log x): one log and one exp instead of the
// general pow; the relative error stays ~(1 + |y log x|) ulp
template <typename R> __device__ __forceinline__ R rpow_pos(R x, R y) { return rexp<R>(y * rlog<R>(x)); }

// DPP lane moves of both halves of a double (quad_perm and row mirrors stay
// inside a 16-lane row; no LDS traffic).  These patterns read no lane
// outside the row, so no lane keeps an old value: a plain DPP move, without
// the initialisation of its destination
template <int CTRL> __device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over aligned groups of G lanes; every lane of a group ends with the
// same value.  Pairings: lane ^ 1 and lane ^ 2 (quad_perm), lane <-> 7 - lane
// and lane <-> 15 - lane inside the row (half mirror, mirror), then lane ^ 16
// and lane ^ 32 across rows.  A fixed pairing, so sums are reproducible.
template <int G> __device__ __forceinline__ double gsum(double v)
{
    if (G >= 2) v += dpp_d<0xB1>(v);
    if (G >= 4) v += dpp_d<0x4E>(v);
    if (G >= 8) v += dpp_d<0x141>(v);
    if (G >= 16) v += dpp_d<0x140>(v);
    if (G >= 32) v += __shfl_xor(v, 16, 64);
    if (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// h(r_ij - r_ik): Tersoff (tersoff_func.f90:168-211) or Kumagai
// (kumagai_func.f90:179-221, including its beta = 3 branch, which leaves
// alpha out of the exponent as the reference does)
template <typename R, int FAM>
__device__ __forceinline__ void h_fn(const TersoffDev &P, int ijpot, int ikpot, R dr, R &val, R &dval)
{
    if (FAM == 3) {
        // juslin_func.f90:272-298: centre i = ikpot / nel, j = ijpot % nel,
        // k = ikpot % nel
        const int t = ikpot % P.nel + P.nel * (ijpot % P.nel + P.nel * (ikpot / P.nel));
        const R alpha = (R)P.jal[t], omega = (R)P.jom[t];
        const int m = P.jm[t];
        if (m == 1) {
            val = omega * rexp<R>(alpha * dr);
            dval = alpha * val;
        } else if (m == 3) {
            const R arg = alpha * dr;
            val = omega * rexp<R>(arg * arg * arg);
            dval = (R)3 * alpha * arg * arg * val;
        } else {
            const R arg = alpha * dr;
            val = omega * rexp<R>(rpow<R>(arg, (R)m));
            dval = (R)m * rpow<R>(arg, (R)(m - 1)) * alpha * val;
        }
        return;
    }
    if (FAM == 1) {
        const R alpha = (R)P.kalpha[ikpot];
        if (alpha == (R)0) {
            val = (R)1;
            dval = (R)0;
            return;
        }
        const int beta = P.kbeta[ikpot];
        if (beta == 1) {
            val = rexp<R>(alpha * dr);
            dval = alpha * val;
        } else if (beta == 3) {
            val = rexp<R>(dr * dr * dr);
            dval = (R)3 * alpha * dr * dr * val;
        } else {
            val = rexp<R>(alpha * rpow<R>(dr, (R)beta));
            dval = (R)beta * alpha * rpow<R>(dr, (R)(beta - 1)) * val;
        }
        return;
    }
    R mu = (R)P.mubo[ikpot];
    if (mu == (R)0) {
        val = (R)1;
        dval = (R)0;
        return;
    }
    int m = P.m[ikpot];
    if (m == 1) {
        val = rexp<R>((R)2 * mu * dr);
        dval = (R)2 * mu * val;
    } else if (m == 3) {
        R arg = (R)2 * mu * dr;
        val = rexp<R>(arg * arg * arg);
        dval = (R)2 * mu * (R)m * arg * arg * val;
    } else {
        val = rexp<R>(rpow<R>((R)2 * mu * dr, (R)m));
        dval = (R)2 * mu * (R)m * rpow<R>((R)2 * mu * dr, (R)(m - 1)) * val;
    }
}

// Tersoff's g and dg/dcos from its centre parameters c^2, c^2/d^2, d^2, h
// and the pair's omega (one reciprocal of d^2 + (h - cos)^2 for both)
template <typename R>
__device__ __forceinline__ void g_tersoff(R om, R c2, R cd2, R d2, R h, R costh, R &g, R &dg)
{
    const R h_c = h - costh;
    const R hh = d2 + h_c * h_c;
    const R rh = (R)1 / hh;
    g = om * ((R)1 + cd2 * h_c * h_c * rh);
    dg = (R)-2 * om * c2 * h_c * rh * rh;
}

// g(cos theta) and dg/dcos.  Tersoff (tersoff_func.f90:86-114) in the
// cancellation-free form omega (1 + c^2 t^2 / (d^2 (d^2 + t^2))), t = h - cos;
// Kumagai (kumagai_func.f90:86-131): c1 + t go (1 + ga1), go = c2 t / (c3 +
// t^2), ga1 = c4 exp(-c5 t^2), all of the centre element.
template <typename R, int FAM>
__device__ __forceinline__ void g_fn(const TersoffDev &P, int eli, int ikpot, R costh, R &g, R &dg)
{
    if (FAM >= 2) {
        // brenner_func.f90:77-101 in the cancellation-free form
        // gamma (1 + c^2 t^2 / (d^2 (d^2 + t^2))), t = h + cos, of the pair ik
        // (one reciprocal of hh for both)
        const R c2 = (R)P.bc2[ikpot], d2 = (R)P.bd2[ikpot], gam = (R)P.bgam[ikpot];
        const R t = (R)P.bh[ikpot] + costh;
        const R rh = (R)1 / (d2 + t * t);
        g = gam * ((R)1 + (R)P.bcd2[ikpot] * t * t * rh);
        dg = (R)2 * gam * c2 * t * rh * rh;
        return;
    }
    if (FAM == 1) {
        const R t = (R)P.h[eli] - costh, tt = t * t;
        const R tmp = t / ((R)P.kc3[eli] + tt);
        const R go = (R)P.kc2[eli] * tmp;
        const R ga1 = (R)P.kc4[eli] * rexp<R>(-(R)P.kc5[eli] * tt);
        const R v = go * ((R)1 + ga1);
        dg = (R)-2 * ((R)1 - t * tmp) * v + (R)2 * (R)P.kc5[eli] * tt * go * ga1;
        g = (R)P.kc1[eli] + t * v;
        return;
    }
    g_tersoff<R>((R)P.omega[ikpot], (R)P.c2[eli], (R)P.cd2[eli], (R)P.d2[eli], (R)P.h[eli], costh, g, dg);
}

// bond order b_ij and Pz = 1/2 fc V_A db/dzeta (V_A already carries the mask
// factor): Tersoff xi (1 + (beta zeta)^n)^(-1/2n) (tersoff_func.f90:122-160),
// Kumagai (1 + zeta^eta)^(-delta) (kumagai_func.f90:139-171)
template <typename R, int FAM>
__device__ __forceinline__ void bo_fn(const TersoffDev &P, int eli, int typ, R zij, R fc, R VA, R &bij, R &Pz)
{
    bij = (R)1;
    Pz = (R)0;
    if (FAM >= 2) {
        // brenner_func.f90:107-147: n = 1 needs no zeta > 0 guard
        const R n = (R)P.bnn[typ], boe = (R)P.bboe[typ];
        if (n == (R)1) {
            const R arg = (R)1 + zij;
            const R a1 = rpow_pos<R>(arg, boe - (R)1);
            bij = a1 * arg;
            Pz = (R)0.5 * boe * n * fc * VA * a1;
        } else if (zij > (R)0) {
            const R zn1 = rpow_pos<R>(zij, n - (R)1);
            const R arg = (R)1 + zn1 * zij;
            const R a1 = rpow_pos<R>(arg, boe - (R)1);
            bij = a1 * arg;
            Pz = (R)0.5 * boe * n * fc * VA * zn1 * a1;
        }
        return;
    }
    if (!(zij > (R)0)) return;
    if (FAM == 1) {
        const R eta = (R)P.keta[eli], delta = (R)P.kdelta[eli];
        const R zn1 = rpow_pos<R>(zij, eta - (R)1);
        const R arg = (R)1 + zn1 * zij;
        const R ad1 = rpow_pos<R>(arg, -delta - (R)1);
        bij = ad1 * arg;
        Pz = (R)0.5 * fc * VA * eta * zn1 * (-delta) * ad1;
        return;
    }
    R nn = (R)P.n[eli];
    R e = (R)P.e[eli];
    R b = (R)P.bn[eli];
    R zn1 = rpow_pos<R>(zij, nn - (R)1);
    R arg = (R)1 + b * zn1 * zij;
    R ae1 = rpow_pos<R>(arg, e - (R)1);
    bij = (R)P.xi[typ] * ae1 * arg;
    Pz = (R)-0.25 * fc * VA * (R)P.xi[typ] * b * zn1 * ae1;
}

template <typename R, int G, int FAM>
__global__ void __launch_bounds__(256)
k_tersoff(TersoffDev P, int nat, const int *__restrict__ order, const double *__restrict__ r,
          const int *__restrict__ el, CellArg C, const int *__restrict__ offsets,
          const int2 *__restrict__ nbr, const int *__restrict__ bslot, const int *__restrict__ bcnt,
          const int *__restrict__ mask, double *__restrict__ fslot, double *__restrict__ fown,
          double *__restrict__ eslot, double *__restrict__ wslot, double *__restrict__ partial,
          double *__restrict__ epb, double *__restrict__ fpb, double *__restrict__ wpb)
{
    constexpr int GPB = 256 / G;
    __shared__ R s_rn[3][256];
    __shared__ R s_rl[256], s_fc[256], s_dfc[256];
    __shared__ int s_typ[256];
    __shared__ double s_red[4][10];

    const int tid = threadIdx.x;
    const int lane = tid % G;
    const int grp = tid / G;
    const int gbase = grp * G;
    const int ai = xcd_block() * GPB + grp;
    const bool valid = ai < nat;
    const int i = valid ? order[ai] : 0;
    const int eli = valid ? el[i] : -1;
    // G is chosen >= every bond count; the clamp only keeps lanes in range
    const int nb = (eli >= 0) ? min(bcnt[i], G) : 0;
    const bool act = lane < nb;

    int s = -1, j = 0, typ = 0;
    R rn[3] = {(R)0, (R)0, (R)0}, rl = (R)1, fc = (R)0, dfc = (R)0;
    bool ok = false;
    if (act) {
        s = bslot[offsets[i] + lane];
        int2 e = nbr[s];
        j = e.x;
        int a, b, c;
        dc_unpack(e.y, a, b, c);
        // rij = r_j - r_i - A.dc   (bop_kernel.f90:611)
        double rij[3];
        for (int x = 0; x < 3; ++x)
            rij[x] = (r[3 * j + x] - r[3 * i + x]) - image_x(C, x, a, b, c);
        double rl2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
        double rld = sqrt(rl2);
        typ = FAM == 3 ? el[j] + eli * P.nel : pair_index(eli, el[j], P.nel);
        for (int x = 0; x < 3; ++x) rn[x] = (R)(rij[x] / rld);
        rl = (R)rld;
        ok = rld < P.r2[typ];
        if (rl2 < P.r1[typ] * P.r1[typ]) {
            fc = (R)1;
            dfc = (R)0;
        } else {
            R fac = (R)P.fac[typ];
            R x = fac * (rl - (R)P.r1[typ]);
            R sn, cs;
            rsincos<R>(x, &sn, &cs);
            fc = (R)0.5 * ((R)1 + cs);
            dfc = (R)-0.5 * fac * sn;
            if (!(rld > P.r1[typ])) { fc = (R)1; dfc = (R)0; }
            if (!(rld < P.r2[typ])) { fc = (R)0; dfc = (R)0; }
        }
    }
    s_rn[0][tid] = rn[0];
    s_rn[1][tid] = rn[1];
    s_rn[2][tid] = rn[2];
    s_rl[tid] = rl;
    s_fc[tid] = fc;
    s_dfc[tid] = dfc;
    s_typ[tid] = ok ? typ : -1;
    __syncthreads();

    int maskfac = 2;
    if (mask && act) {
        int mi = mask[i], mj = mask[j];
        if (mi == 0 && mj == 0) maskfac = 0;
        else if (mi == 0 || mj == 0) maskfac = 1;
    }
    const bool on = act && ok && maskfac > 0;

    // ---- pass 1: zeta_ij ----
    const int elc = eli < 0 ? 0 : eli;
    R zij = (R)0;
    // the k loops stop at the wavefront's largest bond count (uniform, so the
    // group shuffles below always see their whole group)
    int nbw = nb;
    for (int o = 32; o >= 1; o >>= 1) nbw = max(nbw, __shfl_xor(nbw, o, 64));
    constexpr int UNR = G <= 8 ? G : 1;
#pragma unroll UNR
    for (int k = 0; k < G; ++k) {
        if (k >= nbw) break;
        const int tk = gbase + k;
        const int tpk = s_typ[tk];
        if (on && k != lane && k < nb && tpk >= 0) {
            R rnk[3] = {s_rn[0][tk], s_rn[1][tk], s_rn[2][tk]};
            R costh = rnk[0] * rn[0] + rnk[1] * rn[1] + rnk[2] * rn[2];
            R g, dg;
            g_fn<R, FAM>(P, elc, tpk, costh, g, dg);
            R hv, dh;
            h_fn<R, FAM>(P, typ, tpk, rl - s_rl[tk], hv, dh);
            zij += s_fc[tk] * g * hv;
        }
    }

    // ---- bond order ----
    R VA = (R)0, dVA = (R)0, VR = (R)0, dVR = (R)0, bij = (R)1, Pz = (R)0;
    if (on) {
        if (FAM >= 2) {
            // brenner_func.f90:27-71 (juslin_func.f90:69-99 the same)
            const R dr0 = rl - (R)P.br0[typ];
            const R ea = rexp<R>(-(R)P.bexpA[typ] * dr0);
            VA = -(R)P.bVAf[typ] * ea;
            dVA = (R)P.bVAf[typ] * (R)P.bexpA[typ] * ea;
            const R er = rexp<R>(-(R)P.bexpR[typ] * dr0);
            VR = (R)P.bVRf[typ] * er;
            dVR = -(R)P.bVRf[typ] * (R)P.bexpR[typ] * er;
        } else {
            R ea = rexp<R>(-(R)P.mu[typ] * rl);
            VA = -(R)P.B[typ] * ea;
            dVA = (R)P.B[typ] * (R)P.mu[typ] * ea;
            R er = rexp<R>(-(R)P.lambda[typ] * rl);
            VR = (R)P.A[typ] * er;
            dVR = -(R)P.A[typ] * (R)P.lambda[typ] * er;
        }
        R mf = (R)0.5 * (R)maskfac;
        VA *= mf;
        dVA *= mf;
        VR *= mf;
        dVR *= mf;
        bo_fn<R, FAM>(P, eli, typ, zij, fc, VA, bij, Pz);
    }

    // ---- pass 2: derivatives; forces on k reduced into lane k ----
    R dbidi[3] = {(R)0, (R)0, (R)0}, dbidj[3] = {(R)0, (R)0, (R)0};
    R wijb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) wijb[q] = (R)0;
    double recv[3] = {0.0, 0.0, 0.0};
    const R rlijr = (R)1 / rl;
    const R rij[3] = {rl * rn[0], rl * rn[1], rl * rn[2]};
#pragma unroll UNR
    for (int k = 0; k < G; ++k) {
        if (k >= nbw) break;
        const int tk = gbase + k;
        const int tpk = s_typ[tk];
        R dk[3] = {(R)0, (R)0, (R)0};
        if (on && k != lane && k < nb && tpk >= 0) {
            R rnk[3] = {s_rn[0][tk], s_rn[1][tk], s_rn[2][tk]};
            R rlik = s_rl[tk], fcik = s_fc[tk], dfcik = s_dfc[tk];
            R hv, dh;
            h_fn<R, FAM>(P, typ, tpk, rl - rlik, hv, dh);
            R costh = rnk[0] * rn[0] + rnk[1] * rn[1] + rnk[2] * rn[2];
            R g, dg;
            g_fn<R, FAM>(P, elc, tpk, costh, g, dg);
            // r_jk = r_k - r_j; its length cancels in d cos / d r_k and d r_j
            R dkc[3];
            for (int x = 0; x < 3; ++x) dkc[x] = rnk[x] * rlik - rn[x] * rl;
            R rlikr = (R)1 / rlik;
            R dcsdij = rlikr - costh * rlijr;
            R dcsdik = rlijr - costh * rlikr;
            R dcsdjk = -rlijr * rlikr;
            R dzfac = fcik * dg * hv;
            R dzdrij = g * fcik * dh;
            R dzdrik = g * (dfcik * hv - fcik * dh);
            R df[3];
            for (int x = 0; x < 3; ++x) {
                R dcsdi = -dcsdij * rn[x] - dcsdik * rnk[x];
                R dcsdj = dcsdij * rn[x] - dcsdjk * dkc[x];
                R dcsdk = dcsdik * rnk[x] + dcsdjk * dkc[x];
                dbidi[x] += -dzdrij * rn[x] - dzdrik * rnk[x] + dzfac * dcsdi;
                df[x] = dzdrij * rn[x] + dzfac * dcsdj;
                dbidj[x] += df[x];
                dk[x] = dzdrik * rnk[x] + dzfac * dcsdk;
            }
            R rik[3] = {rlik * rnk[0], rlik * rnk[1], rlik * rnk[2]};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) wijb[a + 3 * b] -= rij[a] * df[b] + rik[a] * dk[b];
        }
        // force on k from bond ij: -P_ij * dzeta_ij/dr_k, summed over the group
        double c0 = -(double)Pz * (double)dk[0];
        double c1 = -(double)Pz * (double)dk[1];
        double c2v = -(double)Pz * (double)dk[2];
        c0 = group_sum<G>(c0);
        c1 = group_sum<G>(c1);
        c2v = group_sum<G>(c2v);
        if (lane == k) {
            recv[0] = c0;
            recv[1] = c1;
            recv[2] = c2v;
        }
    }

    double fi[3] = {0.0, 0.0, 0.0}, fj[3] = {0.0, 0.0, 0.0}, dfp[3] = {0.0, 0.0, 0.0};
    double e_ij = 0.0;
    double w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    if (on) {
        e_ij = (double)((R)0.5 * fc * (VR + bij * VA));
        R dffac = (R)0.5 * (dVR * fc + bij * dVA * fc + VR * dfc + bij * VA * dfc);
        for (int x = 0; x < 3; ++x) {
            double df = (double)(dffac * rn[x]);
            dfp[x] = df;
            fi[x] = df - (double)Pz * (double)dbidi[x];
            fj[x] = -df - (double)Pz * (double)dbidj[x];
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                w[a + 3 * b] = (double)rij[a] * (double)(dffac * rn[b]) - (double)Pz * (double)wijb[a + 3 * b];
    }
    if (act) {
        fslot[3 * (size_t)s + 0] = fj[0] + recv[0];
        fslot[3 * (size_t)s + 1] = fj[1] + recv[1];
        fslot[3 * (size_t)s + 2] = fj[2] + recv[2];
        if (eslot) eslot[s] = e_ij;
        if (wslot)
            for (int q = 0; q < 9; ++q) wslot[9 * (size_t)s + q] = w[q];
        // per-bond properties of bond s (bop_kernel.f90:1370-1373,1406-1408,1511-1512)
        if (on) {
            if (epb) epb[s] += e_ij;
            if (fpb)
                for (int x = 0; x < 3; ++x) fpb[3 * (size_t)s + x] += dfp[x];
            if (wpb)
                for (int q = 0; q < 9; ++q) wpb[9 * (size_t)s + q] += w[q];
        }
    }
    // own force of i: sum over the group
    double f0 = group_sum<G>(fi[0]);
    double f1 = group_sum<G>(fi[1]);
    double f2 = group_sum<G>(fi[2]);
    if (valid && lane == 0) {
        fown[3 * (size_t)i + 0] = f0;
        fown[3 * (size_t)i + 1] = f1;
        fown[3 * (size_t)i + 2] = f2;
    }
    // energy + virial: wave butterfly, then the four waves in order
    double red[10];
    red[0] = e_ij;
#pragma unroll
    for (int q = 0; q < 9; ++q) red[1 + q] = w[q];
#pragma unroll
    for (int q = 0; q < 10; ++q) red[q] = group_sum<64>(red[q]);
    const int wave = tid / 64;
    if ((tid & 63) == 0)
        for (int q = 0; q < 10; ++q) s_red[wave][q] = red[q];
    __syncthreads();
    if (tid < 10)
        partial[(size_t)blockIdx.x * 10 + tid] =
            ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
}

// Owner-computes assembly, G lanes per atom (lane = bond): the slot forces
// the neighbours computed for a, summed over the group in a fixed butterfly.
template <int G>
__global__ void __launch_bounds__(256)
k_bop_gather(int nat, const int *__restrict__ offsets, const int *__restrict__ bslot,
             const int *__restrict__ bcnt, const int *__restrict__ rev,
             const double *__restrict__ fslot, const double *__restrict__ fown,
             const double *__restrict__ eslot, const double *__restrict__ wslot,
             double *__restrict__ f, double *__restrict__ epa, double *__restrict__ wpa)
{
    const int lane = threadIdx.x % G;
    const int a = blockIdx.x * (256 / G) + threadIdx.x / G;
    const bool ok = a < nat;
    double acc[3] = {0.0, 0.0, 0.0};
    double e = 0.0, w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    if (ok && lane < bcnt[a]) {
        const int s = bslot[offsets[a] + lane];
        const int t = rev[s];
        if (t >= 0) {
            acc[0] = fslot[3 * (size_t)t];
            acc[1] = fslot[3 * (size_t)t + 1];
            acc[2] = fslot[3 * (size_t)t + 2];
            if (epa) e = eslot[s] + eslot[t];
            if (wpa)
                for (int q = 0; q < 9; ++q) w[q] = wslot[9 * (size_t)s + q] + wslot[9 * (size_t)t + q];
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) acc[x] = group_sum<G>(acc[x]);
    if (epa) e = group_sum<G>(e);
    if (wpa)
        for (int q = 0; q < 9; ++q) w[q] = group_sum<G>(w[q]);
    if (ok && lane == 0) {
        f[3 * (size_t)a] += fown[3 * (size_t)a] + acc[0];
        f[3 * (size_t)a + 1] += fown[3 * (size_t)a + 1] + acc[1];
        f[3 * (size_t)a + 2] += fown[3 * (size_t)a + 2] + acc[2];
        if (epa) epa[a] += 0.5 * e;
        if (wpa)
            for (int q = 0; q < 9; ++q) wpa[9 * (size_t)a + q] += 0.5 * w[q];
    }
}

// -------------------------------------------------------------------------
// Bond-order kernel over the neighbour rows as the list builder left them
// (staged in bin order, or CSR), with no bond compaction and no reverse-slot
// pass: a group of G lanes per atom (G >= the longest row), lane = row
// entry; entries past r2 of their pair, or towards atoms of no element, take
// part as zero-force lanes.  Each lane finds its reverse entry j -> i (image
// -n) in j's row and writes there the total force the terms centred at i put
// on j, so the assembly only sums every atom's own row (k_bop_rows_gather).
// g(cos), dg and h of pass 1 are kept for pass 2 when G <= 8; the virial
// of the terms centred at i is -sum_j r_ij (x) F_j unless per-bond virials
// are needed (WB: the reference's per-bond wij, bop_kernel.f90:1325-1512).
// HT = false: every h-term of the parameter set is 1 (mubo = 0), so h and
// dh are constants.  The reverse entry is looked up first, so its loads
// overlap the arithmetic.  Over a list build whose flags the host has not
// read (nlflags non-NULL), a build that proved unusable, or a row longer
// than G, sets DERR_REDO in flags[0]: the gather and the reduction then
// leave the caller's outputs alone and the host redoes the call.
// -------------------------------------------------------------------------
constexpr int DERR_REV_MISSING = DERR_BOND_OVERFLOW;

template <typename R, int G, int FAM, bool WB, bool HT = true>
__global__ void __launch_bounds__(256)
k_bop_rows(TersoffDev P, int nat, const int *__restrict__ order, const int *__restrict__ el_s, CellArg C,
           NlRows L, const int *__restrict__ offsets, const int *__restrict__ mask, double *__restrict__ frev,
           double *__restrict__ fown, double *__restrict__ eslot, double *__restrict__ erev,
           double *__restrict__ wslot, double *__restrict__ wrev, double *__restrict__ partial,
           double *__restrict__ epb, double *__restrict__ fpb, double *__restrict__ wpb, int *__restrict__ flags,
           const int *__restrict__ nlflags, int vrev)
{
    if (nlflags && (*nlflags & DERR_NL_UNUSABLE)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flags, DERR_REDO);
        return;
    }
    constexpr int GPB = 256 / G;
    constexpr bool CACHE = G <= 8;
    constexpr int NC = CACHE ? G : 1;
    __shared__ R s_rn[3][256];
    __shared__ R s_rl[256], s_ri[256], s_fc[256], s_dfc[256];
    __shared__ int s_typ[256];
    __shared__ double s_acc[10][256 / 16];

    const int tid = threadIdx.x;
    const int lane = tid % G;
    const int grp = tid / G;
    const int gbase = grp * G;
    const int ai = xcd_block() * GPB + grp;
    const bool valid = ai < nat;
    const int i = valid ? order[ai] : 0;           // caller index (mask, per-bond arrays)
    const int eli = valid ? el_s[ai] : -1;
    int n = 0;
    const int2 *row = valid ? nl_row(L, ai, n) : nullptr;
    if (n > G) atomicOr(flags, DERR_REDO);  // only over an unchecked build
    n = min(n, G);
    const bool act = lane < n;

    int j = 0, code = 0, typ = 0;       // j: the neighbour's slot
    R rn[3] = {(R)0, (R)0, (R)0}, rl = (R)1, fc = (R)0, dfc = (R)0;
    bool ok = false;
    long long rt = -1;                  // entry index of the reverse entry j -> i
    if (act) {
        const int2 e = row[lane];
        j = e.x;
        code = e.y;
        // reverse entry j -> i (image -n) in j's row
        {
            int a0, b0, c0;
            dc_unpack(code, a0, b0, c0);
            const int want = dc_pack(-a0, -b0, -c0);
            int nj = 0;
            const int2 *rowj = nl_row(L, j, nj);
            int t = -1;
            if (G <= 8) {
                // every candidate load issued before the first compare; staged
                // rows of an even stride of at least G (vrev) are read as G/2
                // aligned 16-byte vectors, entries past the row masked
                int2 ej[G];
                if (vrev) {
                    const int4 *r4 = reinterpret_cast<const int4 *>(rowj);
#pragma unroll
                    for (int q = 0; q < G / 2; ++q) {
                        const int4 v = r4[q];
                        ej[2 * q] = make_int2(2 * q < nj ? v.x : -1, v.y);
                        ej[2 * q + 1] = make_int2(2 * q + 1 < nj ? v.z : -1, v.w);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < G; ++q) ej[q] = q < nj ? rowj[q] : make_int2(-1, 0);
                }
#pragma unroll
                for (int q = G - 1; q >= 0; --q)
                    if (ej[q].x == ai && ej[q].y == want) t = q;
                if (nj > G) t = -1;     // longer rows only on hosts that chose a wider G
            } else {
                for (int q = 0; q < nj; ++q) {
                    const int2 eq = rowj[q];
                    if (eq.x == ai && eq.y == want) {
                        t = q;
                        break;
                    }
                }
            }
            rt = t < 0 ? -1 : (long long)nl_at(L, j, t);
        }
        const int elj = el_s[j];
        if (eli >= 0 && elj >= 0) {
            int a, b, c;
            dc_unpack(code, a, b, c);
            const SortedAtom Ai = L.atoms[ai], Aj = L.atoms[j];
            const double ri[3] = {Ai.x, Ai.y, Ai.z}, rj[3] = {Aj.x, Aj.y, Aj.z};
            // rij = r_j - r_i - A.dc   (bop_kernel.f90:611)
            double rij[3];
            for (int x = 0; x < 3; ++x) rij[x] = (rj[x] - ri[x]) - image_x(C, x, a, b, c);
            const double rl2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
            const double rld = sqrt(rl2);
            typ = FAM == 3 ? elj + eli * P.nel : pair_index(eli, elj, P.nel);
            ok = rld < P.r2[typ];
            if (ok) {
                const double inv = 1.0 / rld;
                for (int x = 0; x < 3; ++x) rn[x] = (R)(rij[x] * inv);
                rl = (R)rld;
                if (rl2 < P.r1[typ] * P.r1[typ]) {
                    fc = (R)1;
                    dfc = (R)0;
                } else {
                    // trig_off (cutoff.f90:171-196)
                    const R fac = (R)P.fac[typ];
                    R sn, cs;
                    rsincos<R>(fac * (rl - (R)P.r1[typ]), &sn, &cs);
                    fc = (R)0.5 * ((R)1 + cs);
                    dfc = (R)-0.5 * fac * sn;
                    if (!(rld > P.r1[typ])) { fc = (R)1; dfc = (R)0; }
                }
            }
        }
    }
    s_rn[0][tid] = rn[0];
    s_rn[1][tid] = rn[1];
    s_rn[2][tid] = rn[2];
    s_rl[tid] = rl;
    s_ri[tid] = (R)1 / rl;
    s_fc[tid] = fc;
    s_dfc[tid] = dfc;
    s_typ[tid] = ok ? typ : -1;
    __syncthreads();

    int maskfac = 2;
    if (mask && act) {
        const int mi = mask[i], mj = mask[order[j]];
        if (mi == 0 && mj == 0) maskfac = 0;
        else if (mi == 0 || mj == 0) maskfac = 1;
    }
    const bool on = ok && maskfac > 0;

    // ---- pass 1: zeta_ij (g, dg, h, dh kept for pass 2) ----
    const int elc = eli < 0 ? 0 : eli;
    // the centre's parameters of g, read once rather than per triplet
    const R gc2 = FAM == 0 ? (R)P.c2[elc] : (R)0, gcd2 = FAM == 0 ? (R)P.cd2[elc] : (R)0;
    const R gd2 = FAM == 0 ? (R)P.d2[elc] : (R)1, gh = FAM == 0 ? (R)P.h[elc] : (R)0;
    R zij = (R)0;
    R cg[NC], cdg[NC], chv[NC], cdh[NC];
    int nbw = n;
    for (int o = 32; o >= 1; o >>= 1) nbw = max(nbw, __shfl_xor(nbw, o, 64));
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (!CACHE && k >= nbw) break;
        const int tk = gbase + k;
        const int tpk = s_typ[tk];
        R g = (R)0, dg = (R)0, hv = (R)0, dh = (R)0;
        if (on && k != lane && tpk >= 0) {
            const R costh = s_rn[0][tk] * rn[0] + s_rn[1][tk] * rn[1] + s_rn[2][tk] * rn[2];
            if (FAM == 0)
                g_tersoff<R>((R)P.omega[tpk], gc2, gcd2, gd2, gh, costh, g, dg);
            else
                g_fn<R, FAM>(P, elc, tpk, costh, g, dg);
            if (HT) {
                h_fn<R, FAM>(P, typ, tpk, rl - s_rl[tk], hv, dh);
            } else {
                hv = (R)1;
                dh = (R)0;
            }
            zij += s_fc[tk] * g * hv;
        }
        if (CACHE) {
            cg[CACHE ? k : 0] = g;
            cdg[CACHE ? k : 0] = dg;
            chv[CACHE ? k : 0] = hv;
            cdh[CACHE ? k : 0] = dh;
        }
    }

    // ---- bond order ----
    R VA = (R)0, dVA = (R)0, VR = (R)0, dVR = (R)0, bij = (R)1, Pz = (R)0;
    if (on) {
        if (FAM >= 2) {
            // brenner_func.f90:27-71 (juslin_func.f90:69-99 the same)
            const R dr0 = rl - (R)P.br0[typ];
            const R ea = rexp<R>(-(R)P.bexpA[typ] * dr0);
            VA = -(R)P.bVAf[typ] * ea;
            dVA = (R)P.bVAf[typ] * (R)P.bexpA[typ] * ea;
            const R er = rexp<R>(-(R)P.bexpR[typ] * dr0);
            VR = (R)P.bVRf[typ] * er;
            dVR = -(R)P.bVRf[typ] * (R)P.bexpR[typ] * er;
        } else {
            const R ea = rexp<R>(-(R)P.mu[typ] * rl);
            VA = -(R)P.B[typ] * ea;
            dVA = (R)P.B[typ] * (R)P.mu[typ] * ea;
            const R er = rexp<R>(-(R)P.lambda[typ] * rl);
            VR = (R)P.A[typ] * er;
            dVR = -(R)P.A[typ] * (R)P.lambda[typ] * er;
        }
        const R mf = (R)0.5 * (R)maskfac;
        VA *= mf;
        dVA *= mf;
        VR *= mf;
        dVR *= mf;
        bo_fn<R, FAM>(P, eli, typ, zij, fc, VA, bij, Pz);
    }

    // ---- pass 2: derivatives; forces on k summed over the group into lane k ----
    // zeta_ij depends on r_ij and the r_ik only, so its gradient in r_i is
    // minus the sum of those in r_j and the r_k (dksum: the latter)
    R dksum[3] = {(R)0, (R)0, (R)0}, dbidj[3] = {(R)0, (R)0, (R)0};
    R wijb[WB ? 9 : 1];
#pragma unroll
    for (int q = 0; q < (WB ? 9 : 1); ++q) wijb[q] = (R)0;
    double recv[3] = {0.0, 0.0, 0.0};
    const R rlijr = (R)1 / rl;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (!CACHE && k >= nbw) break;
        const int tk = gbase + k;
        const int tpk = s_typ[tk];
        R dk[3] = {(R)0, (R)0, (R)0};
        if (on && k != lane && tpk >= 0) {
            const R rnk[3] = {s_rn[0][tk], s_rn[1][tk], s_rn[2][tk]};
            const R rlik = s_rl[tk], rlikr = s_ri[tk], fcik = s_fc[tk], dfcik = s_dfc[tk];
            R g, dg, hv, dh;
            if (CACHE) {
                g = cg[CACHE ? k : 0];
                dg = cdg[CACHE ? k : 0];
                hv = chv[CACHE ? k : 0];
                dh = cdh[CACHE ? k : 0];
            } else {
                const R cs0 = rnk[0] * rn[0] + rnk[1] * rn[1] + rnk[2] * rn[2];
                g_fn<R, FAM>(P, elc, tpk, cs0, g, dg);
                if (HT) {
                    h_fn<R, FAM>(P, typ, tpk, rl - rlik, hv, dh);
                } else {
                    hv = (R)1;
                    dh = (R)0;
                }
            }
            if (!HT) {
                hv = (R)1;
                dh = (R)0;
            }
            const R costh = rnk[0] * rn[0] + rnk[1] * rn[1] + rnk[2] * rn[2];
            // r_jk = r_k - r_j; d cos / d r_jk along the unit vector r_jk / |r_jk|
            // is -|r_jk| / (r_ij r_ik), so with the vector itself the length
            // cancels (no square root or division per triplet)
            R dkc[3];
            for (int x = 0; x < 3; ++x) dkc[x] = rnk[x] * rlik - rn[x] * rl;
            const R dcsdij = rlikr - costh * rlijr;
            const R dcsdik = rlijr - costh * rlikr;
            const R dcsdjk = -rlijr * rlikr;
            const R dzfac = fcik * dg * hv;
            const R dzdrij = g * fcik * dh;
            const R dzdrik = g * (dfcik * hv - fcik * dh);
            R df[3];
            for (int x = 0; x < 3; ++x) {
                const R dcsdj = dcsdij * rn[x] - dcsdjk * dkc[x];
                const R dcsdk = dcsdik * rnk[x] + dcsdjk * dkc[x];
                df[x] = dzdrij * rn[x] + dzfac * dcsdj;
                dbidj[x] += df[x];
                dk[x] = dzdrik * rnk[x] + dzfac * dcsdk;
                dksum[x] += dk[x];
            }
            if (WB) {
                const R rij[3] = {rl * rn[0], rl * rn[1], rl * rn[2]};
                const R rik[3] = {rlik * rnk[0], rlik * rnk[1], rlik * rnk[2]};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) wijb[WB ? a + 3 * b : 0] -= rij[a] * df[b] + rik[a] * dk[b];
            }
        }
        // force on k from bond ij: -P_ij dzeta_ij/dr_k, summed over the group
        const double c0 = gsum<G>(-(double)Pz * (double)dk[0]);
        const double c1 = gsum<G>(-(double)Pz * (double)dk[1]);
        const double c2v = gsum<G>(-(double)Pz * (double)dk[2]);
        if (lane == k) {
            recv[0] = c0;
            recv[1] = c1;
            recv[2] = c2v;
        }
    }

    double fi[3] = {0.0, 0.0, 0.0}, F[3] = {0.0, 0.0, 0.0}, dfp[3] = {0.0, 0.0, 0.0};
    double e_ij = 0.0;
    double w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    const double rijd[3] = {(double)(rl * rn[0]), (double)(rl * rn[1]), (double)(rl * rn[2])};
    if (on) {
        e_ij = (double)((R)0.5 * fc * (VR + bij * VA));
        const R dffac = (R)0.5 * (dVR * fc + bij * dVA * fc + VR * dfc + bij * VA * dfc);
        for (int x = 0; x < 3; ++x) {
            dfp[x] = (double)(dffac * rn[x]);
            const R dbidi = -(dbidj[x] + dksum[x]);
            fi[x] = dfp[x] - (double)Pz * (double)dbidi;
            F[x] = -dfp[x] - (double)Pz * (double)dbidj[x];
        }
        if (WB)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    w[a + 3 * b] = rijd[a] * dfp[b] - (double)Pz * (double)wijb[WB ? a + 3 * b : 0];
    }
    for (int x = 0; x < 3; ++x) F[x] += recv[x];
    if (!WB)
        // virial of the terms centred at i: -sum_j r_ij (x) F_j (their
        // energy depends on the r_ij only)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) w[a + 3 * b] = -rijd[a] * F[b];
    if (act) {
        const size_t sown = nl_at(L, ai, lane);
        // the caller's list entry (per-bond arrays only)
        const size_t sbond = (epb || fpb || wpb) ? (size_t)offsets[i] + lane : 0;
        if (rt < 0) {
            atomicOr(flags, DERR_REV_MISSING);
        } else {
            frev[3 * (size_t)rt + 0] = F[0];
            frev[3 * (size_t)rt + 1] = F[1];
            frev[3 * (size_t)rt + 2] = F[2];
            if (erev) erev[rt] = e_ij;
            if (wrev)
                for (int q = 0; q < 9; ++q) wrev[9 * (size_t)rt + q] = w[q];
        }
        if (eslot) eslot[sown] = e_ij;
        if (wslot)
            for (int q = 0; q < 9; ++q) wslot[9 * sown + q] = w[q];
        // per-bond properties (bop_kernel.f90:1370-1373,1406-1408,1511-1512)
        if (on) {
            if (epb) epb[sbond] += e_ij;
            if (fpb)
                for (int x = 0; x < 3; ++x) fpb[3 * sbond + x] += dfp[x];
            if (wpb)
                for (int q = 0; q < 9; ++q) wpb[9 * sbond + q] += w[q];
        }
    }
    // own force of i: sum over the group
    const double f0 = gsum<G>(fi[0]);
    const double f1 = gsum<G>(fi[1]);
    const double f2 = gsum<G>(fi[2]);
    if (valid && lane == 0) {
        fown[3 * (size_t)ai + 0] = f0;
        fown[3 * (size_t)ai + 1] = f1;
        fown[3 * (size_t)ai + 2] = f2;
    }
    // energy + virial of the block: 16-lane sums (DPP), then 16 partial
    // records per value summed in order
    double red[10];
    red[0] = e_ij;
#pragma unroll
    for (int q = 0; q < 9; ++q) red[1 + q] = w[q];
#pragma unroll
    for (int q = 0; q < 10; ++q) red[q] = gsum<16>(red[q]);
    if ((tid & 15) == 0)
        for (int q = 0; q < 10; ++q) s_acc[q][tid >> 4] = red[q];
    __syncthreads();
    if (tid < 10) {
        double t = 0.0;
        for (int c = 0; c < 16; ++c) t += s_acc[tid][c];
        partial[(size_t)blockIdx.x * 10 + tid] = t;
    }
}

// Assembly for k_bop_rows: f_a += own force + the forces a's neighbours'
// terms put on a (stored in a's own row), in row order; per-atom energy and
// virial halves of a's bonds in both directions (pe(i), pe(j) of
// bop_kernel.f90:1366-1368).  A thread per slot x (caller atom order[x]).
__global__ void __launch_bounds__(256)
k_bop_rows_gather(int nat, const int *__restrict__ order, NlRows L, const double *__restrict__ frev,
                  const double *__restrict__ fown, const double *__restrict__ eslot, const double *__restrict__ erev,
                  const double *__restrict__ wslot, const double *__restrict__ wrev, double *__restrict__ f,
                  double *__restrict__ epa, double *__restrict__ wpa, const int *__restrict__ flags)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nat || (*flags & DERR_REDO)) return;
    const int a = order[x];
    int n;
    nl_row(L, x, n);
    const size_t s0 = nl_at(L, x, 0);
    double acc[3] = {fown[3 * (size_t)x], fown[3 * (size_t)x + 1], fown[3 * (size_t)x + 2]};
    double e = 0.0;
    for (size_t s = s0; s < s0 + n; ++s) {
        acc[0] += frev[3 * s];
        acc[1] += frev[3 * s + 1];
        acc[2] += frev[3 * s + 2];
        if (epa) e += eslot[s] + erev[s];
    }
    f[3 * (size_t)a] += acc[0];
    f[3 * (size_t)a + 1] += acc[1];
    f[3 * (size_t)a + 2] += acc[2];
    if (epa) epa[a] += 0.5 * e;
    if (wpa)
        for (int q = 0; q < 9; ++q) {
            double w = 0.0;
            for (size_t s = s0; s < s0 + n; ++s) w += wslot[9 * s + q] + wrev[9 * s + q];
            wpa[9 * (size_t)a + q] += 0.5 * w;
        }
}

// -------------------------------------------------------------------------
// screened variants (TersoffScr, KumagaiScr, BrennerScr: SCREENING in
// bop_kernel.f90).  Every list entry i->j (image n) inside the screening
// range gets the screening function of the bond from the third atoms k of
// i's row (loop1, bop_kernel.f90:563-1066), in the reference's row order and
// with its early exit:
//   k_scr_screen<false>  keep flag, pair and bond-order cutoffs (value, d/dr)
//                        and the number of screening records per entry
//   (exclusive scan of the record counts)
//   k_scr_screen<true>   the records (k's entry, d fc_ar / d r_ik,
//                        d fc_ar / d r_jk, and the same for fc_bo), only in
//                        rows that have any
//   k_scr_bonds          kept entries of every row -> the bond list
//   k_bop_scr            the bond-order kernel of k_tersoff with these
//                        cutoffs, plus the screening forces: for the pair
//                        term (:1448-1492) and, with the bond's factor
//                        sum_j g h dE/dzeta_ij, for the bond order
//                        (:1531-1609); the force a record puts on its k is
//                        written per record
//   k_scr_slots          per entry of i's row: the total force i's terms put
//                        on that neighbour (bond slot + records naming it)
//   k_scr_gather         f_a = own + sum over a's row of the slot the
//                        neighbour wrote for a (through the reverse entry)
// No atomics: every sum has one owner and a fixed order.
// -------------------------------------------------------------------------
// per element pair: the symmetric pair index (6 used), or the ordered one
// j + i nel of the Juslin family (9)
struct ScrDev {
    double r1[9], r2[9], or1[9], or2[9], bor1[9], bor2[9];
    double fin[9], fout[9], fbo[9];            // 1 / (r2 - r1) of the three cutoffs
    double Cmin[9], Cmax[9], dC[9], C_dr_cut[9], max_cut_sq[9];
    double ec, ed, efac2, eoff;                // exp_cutoff constants
};

constexpr int SCR_ROW = 256;       // row entries staged per atom
constexpr int SCR_WAVES = 4;       // atoms (wavefronts) per block

// exp_cutoff_f (cutoff.f90:271-293)
__device__ __forceinline__ void exp_cut(const ScrDev &S, double r1, double r2, double fac1, double r, double &val,
                                        double &dval)
{
    if (r <= r1) {
        val = 1.0;
        dval = 0.0;
    } else if (r >= r2) {
        val = 0.0;
        dval = 0.0;
    } else {
        const double x = fac1 * (r - r1), x2 = x * x;
        const double v = exp(-8 * x * x2);
        const double dv = -24 * x2 * v;
        dval = fac1 * S.efac2 * (dv + 3 * S.ec * x2 + 4 * S.ed * x * x2);
        val = S.efac2 * (v + S.ec * x * x2 + S.ed * x2 * x2 - S.eoff);
    }
}

struct ScrShared {
    double x[SCR_ROW], y[SCR_ROW], z[SCR_ROW];
    double r2[SCR_ROW];           // |r_ik|^2 of each entry, tested first
    int2 e[SCR_ROW];
};

// One wavefront per atom, lanes over the row's entries; the row's vectors
// r_j - r_i - A.n are staged in LDS.  FILL = false: keep flag and record
// count; FILL = true: cutoffs and records at roff.
template <bool FILL>
__global__ void __launch_bounds__(256)
k_scr_screen(ScrDev S, int nat_tot, const double *__restrict__ r, const int *__restrict__ el, int nel, int ordered,
             CellArg C,
             const int *__restrict__ offsets, const int2 *__restrict__ nbr, int *__restrict__ keep,
             int *__restrict__ rcnt, const int *__restrict__ roff, double4 *__restrict__ fcv,
             int *__restrict__ rkn, double4 *__restrict__ rco, int *__restrict__ flags)
{
    __shared__ ScrShared sh[SCR_WAVES];
    ScrShared &L = sh[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * SCR_WAVES + (threadIdx.x >> 6);
    if (i >= nat_tot) return;                       // uniform per wave
    const int beg = offsets[i], n = offsets[i + 1] - beg;
    if (n > SCR_ROW) {
        if (lane == 0) atomicOr(flags, DERR_SCR_ROW);
        return;
    }
    // pass 2 only writes records: rows without any have nothing to do
    if (FILL && roff[beg + n] == roff[beg]) return;
    const int eli = el[i];
    for (int t = lane; t < n; t += 64) {
        const int2 e = nbr[beg + t];
        int a, b, c;
        dc_unpack(e.y, a, b, c);
        const double dx = (r[3 * e.x + 0] - r[3 * i + 0]) - image_x(C, 0, a, b, c);
        const double dy = (r[3 * e.x + 1] - r[3 * i + 1]) - image_x(C, 1, a, b, c);
        const double dz = (r[3 * e.x + 2] - r[3 * i + 2]) - image_x(C, 2, a, b, c);
        L.x[t] = dx;
        L.y[t] = dy;
        L.z[t] = dz;
        L.r2[t] = dx * dx + dy * dy + dz * dz;
        L.e[t] = e;
    }
    __builtin_amdgcn_wave_barrier();
    const double thr = -13.815510557964274;        // log(1e-6)
    const double dot_thr = 1e-10;
    for (int t = lane; t < n; t += 64) {
        const int s = beg + t;
        const int2 ej = L.e[t];
        const int elj = el[ej.x];
        int kept = 0, cnt = 0;
        double fcar = 0.0, dfcar = 0.0, fcbo = 0.0, dfcbo = 0.0;
        if (eli >= 0 && elj >= 0) {
            const int typ = ordered ? elj + eli * nel : pair_index(eli, elj, nel);
            const double rij[3] = {L.x[t], L.y[t], L.z[t]};
            const double rl = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
            const double cin_l = S.r1[typ], cin_h = S.r2[typ];
            // records go to [roff[s], roff[s] + rcnt[s]) as counted by pass 1:
            // an entry that ends up dropped or fully screened counted none
            const int q0 = FILL ? roff[s] : 0;
            const int qcap = FILL ? rcnt[s] : 0;
            if (rl < cin_l * cin_l) {
                // region (a): not screened
                kept = 1;
                fcar = fcbo = 1.0;
            } else if (rl < S.max_cut_sq[typ] && S.or1[typ] < S.or2[typ]) {
                bool screened = false, need = false;
                double sij = 0.0, dsijdrij = 0.0;
                const double Cmin = S.Cmin[typ], Cmax = S.Cmax[typ], dC = S.dC[typ];
                const double lim = S.C_dr_cut[typ] * rl;
                for (int u = 0; !(screened || sij < thr) && u < n; ++u) {
                    const double rlik = L.r2[u];
                    if (!(rlik < lim)) continue;
                    const double rik[3] = {L.x[u], L.y[u], L.z[u]};
                    const int2 ek = L.e[u];
                    if (ek.x == ej.x && ek.y == ej.y) continue;
                    const double dot_ij_ik = rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2];
                    const double rjk[3] = {-rij[0] + rik[0], -rij[1] + rik[1], -rij[2] + rik[2]};
                    const double dot_ij_jk = rij[0] * rjk[0] + rij[1] * rjk[1] + rij[2] * rjk[2];
                    const double rljk = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
                    if (!(dot_ij_ik > dot_thr && dot_ij_jk < -dot_thr)) continue;
                    const double xik = rlik / rl, xjk = rljk / rl;
                    const double xm = xik - xjk, xp = xik + xjk;
                    double fac = 1.0 / (1 - xm * xm);
                    const double Cv = (2 * xp - xm * xm - 1) * fac;
                    if (Cv <= Cmin) {
                        screened = true;
                    } else if (Cv < Cmax) {
                        need = true;
                        const double Cmax_C = Cmax - Cv, C_Cmin = Cv - Cmin;
                        sij = sij - (Cmax_C / C_Cmin) * (Cmax_C / C_Cmin);
                        const double dCdrik = 4 * xik * fac * (1 + (Cv - 1) * xm);
                        const double dCdrjk = 4 * xjk * fac * (1 - (Cv - 1) * xm);
                        fac = 2 * Cmax_C * dC / (C_Cmin * C_Cmin * C_Cmin);
                        dsijdrij = dsijdrij + fac * -(dCdrik + dCdrjk);
                        if (FILL && cnt < qcap) {
                            rkn[q0 + cnt] = beg + u;
                            rco[q0 + cnt] = make_double4(fac * dCdrik / rlik, fac * dCdrjk / rljk, 0.0, 0.0);
                        }
                        ++cnt;
                    }
                }
                if ((screened || sij < thr) && rl > cin_h * cin_h) {
                    cnt = 0;              // fully screened: dropped
                } else {
                    kept = 1;
                    const double len = sqrt(rl);
                    double fcin, dfcin;
                    exp_cut(S, cin_l, cin_h, S.fin[typ], len, fcin, dfcin);
                    double fo, dfo, fb, dfb;
                    exp_cut(S, S.or1[typ], S.or2[typ], S.fout[typ], len, fo, dfo);
                    exp_cut(S, S.bor1[typ], S.bor2[typ], S.fbo[typ], len, fb, dfb);
                    if (screened) {
                        fcar = fcbo = fcin;
                        dfcar = dfcbo = dfcin;
                        cnt = 0;
                    } else if (need) {
                        const double sv = exp(sij);
                        fcar = (1.0 - fcin) * sv * fo + fcin;
                        dfcar = (1.0 - fcin) * sv * (dfo + fo * dsijdrij / len) - dfcin * sv * fo + dfcin;
                        fcbo = (1.0 - fcin) * sv * fb + fcin;
                        dfcbo = (1.0 - fcin) * sv * (dfb + fb * dsijdrij / len) - dfcin * sv * fb + dfcin;
                        if (FILL) {
                            const double car = sv * fo * (1.0 - fcin), cbo = sv * fb * (1.0 - fcin);
                            for (int q = 0; q < min(cnt, qcap); ++q) {
                                const double4 v = rco[q0 + q];
                                rco[q0 + q] = make_double4(v.x * car, v.y * car, v.x * cbo, v.y * cbo);
                            }
                        }
                    } else if (len < cin_h) {
                        fcar = (1.0 - fcin) * fo + fcin;
                        dfcar = (1.0 - fcin) * dfo - dfcin * fo + dfcin;
                        fcbo = (1.0 - fcin) * fb + fcin;
                        dfcbo = (1.0 - fcin) * dfb - dfcin * fb + dfcin;
                    } else {
                        fcar = fo;
                        dfcar = dfo;
                        fcbo = fb;
                        dfcbo = dfb;
                    }
                }
            } else if (rl < cin_h * cin_h) {
                // beyond the screening range: the inner cutoff only
                kept = 1;
                exp_cut(S, cin_l, cin_h, S.fin[typ], sqrt(rl), fcar, dfcar);
                fcbo = fcar;
                dfcbo = dfcar;
            }
        }
        if (!FILL) {
            if (kept) fcv[s] = make_double4(fcar, dfcar, fcbo, dfcbo);
            keep[s] = kept;
            rcnt[s] = kept ? cnt : 0;
        }
    }
}

// thread per atom: the kept entries of its row, in row order
__global__ void k_scr_bonds(int nat, const int *__restrict__ offsets, const int *__restrict__ keep,
                            int *__restrict__ bslot, int *__restrict__ bcnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nat) return;
    const int beg = offsets[i], end = offsets[i + 1];
    int nb = 0;
    for (int s = beg; s < end; ++s)
        if (keep[s]) bslot[beg + nb++] = s;
    bcnt[i] = nb;
}

// reverse entry of every list entry: i->j (image n) -> j->i (image -n)
__global__ void k_nl_rev_entries(int nat, const int *__restrict__ offsets, const int2 *__restrict__ nbr,
                                 int *__restrict__ rev, int *__restrict__ flags)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nat) return;
    const int beg = offsets[i], end = offsets[i + 1];
    for (int s = beg + lane; s < end; s += 64) {
        const int2 e = nbr[s];
        int a, b, c;
        dc_unpack(e.y, a, b, c);
        const int want = dc_pack(-a, -b, -c);
        // j's row in chunks of 8 entries, all loads of a chunk in flight
        // together; the first match in row order
        int found = -1;
        const int qe = offsets[e.x + 1];
        for (int q0 = offsets[e.x]; q0 < qe && found < 0; q0 += 8) {
            int2 f[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) f[u] = q0 + u < qe ? nbr[q0 + u] : make_int2(-1, 0);
#pragma unroll
            for (int u = 7; u >= 0; --u)
                if (f[u].x == i && f[u].y == want) found = q0 + u;
        }
        rev[s] = found;
        if (found < 0) atomicOr(flags, DERR_BOND_OVERFLOW);
    }
}

template <typename R, int G, int FAM>
__global__ void __launch_bounds__(256)
k_bop_scr(TersoffDev P, ScrDev S, int nat, const int *__restrict__ order, const double *__restrict__ r,
          const int *__restrict__ el, CellArg C, const int *__restrict__ offsets, const int2 *__restrict__ nbr,
          const int *__restrict__ bslot, const int *__restrict__ bcnt, const int *__restrict__ mask,
          const double4 *__restrict__ fcv, const int *__restrict__ roff, const int *__restrict__ rkn,
          const double4 *__restrict__ rco, double *__restrict__ fslot, double *__restrict__ frec,
          double *__restrict__ fown, double *__restrict__ eslot, double *__restrict__ wslot,
          double *__restrict__ partial, double *__restrict__ epb, double *__restrict__ fpb,
          double *__restrict__ wpb)
{
    constexpr int GPB = 256 / G;
    __shared__ R s_rn[3][256];
    __shared__ R s_rl[256], s_fc[256], s_dfc[256];
    __shared__ int s_typ[256];
    __shared__ double s_red[4][10];

    const int tid = threadIdx.x;
    const int lane = tid % G;
    const int grp = tid / G;
    const int gbase = grp * G;
    const int ai = xcd_block() * GPB + grp;
    const bool valid = ai < nat;
    const int i = valid ? order[ai] : 0;
    const int eli = valid ? el[i] : -1;
    const int nb = (eli >= 0) ? min(bcnt[i], G) : 0;
    const bool act = lane < nb;

    int s = -1, j = 0, typ = 0;
    double rijd[3] = {0.0, 0.0, 0.0};
    R rn[3] = {(R)0, (R)0, (R)0}, rl = (R)1, fc = (R)0, dfc = (R)0;
    bool ok = false, okbo = false;
    R fbo = (R)0, dfbo = (R)0;
    if (act) {
        s = bslot[offsets[i] + lane];
        const int2 e = nbr[s];
        j = e.x;
        int a, b, c;
        dc_unpack(e.y, a, b, c);
        for (int x = 0; x < 3; ++x) rijd[x] = (r[3 * j + x] - r[3 * i + x]) - image_x(C, x, a, b, c);
        const double rld = sqrt(rijd[0] * rijd[0] + rijd[1] * rijd[1] + rijd[2] * rijd[2]);
        typ = FAM == 3 ? el[j] + eli * P.nel : pair_index(eli, el[j], P.nel);
        for (int x = 0; x < 3; ++x) rn[x] = (R)(rijd[x] / rld);
        rl = (R)rld;
        const double4 cv = fcv[s];
        fc = (R)cv.x;
        dfc = (R)cv.y;
        fbo = (R)cv.z;
        dfbo = (R)cv.w;
        ok = rld < S.or2[typ];           // cut_ar_h
        okbo = rld < S.bor2[typ];        // cut_bo_h
    }
    s_rn[0][tid] = rn[0];
    s_rn[1][tid] = rn[1];
    s_rn[2][tid] = rn[2];
    s_rl[tid] = rl;
    s_fc[tid] = fbo;
    s_dfc[tid] = dfbo;
    s_typ[tid] = okbo ? typ : -1;
    __syncthreads();

    int maskfac = 2;
    if (mask && act) {
        int mi = mask[i], mj = mask[j];
        if (mi == 0 && mj == 0) maskfac = 0;
        else if (mi == 0 || mj == 0) maskfac = 1;
    }
    const bool on = act && ok && maskfac > 0;

    const int elc = eli < 0 ? 0 : eli;
    R zij = (R)0;
    int nbw = nb;
    for (int o = 32; o >= 1; o >>= 1) nbw = max(nbw, __shfl_xor(nbw, o, 64));
    for (int k = 0; k < G; ++k) {
        if (k >= nbw) break;
        const int tk = gbase + k;
        const int tpk = s_typ[tk];
        if (on && k != lane && k < nb && tpk >= 0) {
            R rnk[3] = {s_rn[0][tk], s_rn[1][tk], s_rn[2][tk]};
            R costh = rnk[0] * rn[0] + rnk[1] * rn[1] + rnk[2] * rn[2];
            R g, dg;
            g_fn<R, FAM>(P, elc, tpk, costh, g, dg);
            R hv, dh;
            h_fn<R, FAM>(P, typ, tpk, rl - s_rl[tk], hv, dh);
            zij += s_fc[tk] * g * hv;
        }
    }

    R VA = (R)0, dVA = (R)0, VR = (R)0, dVR = (R)0, bij = (R)1, Pz = (R)0;
    if (on) {
        if (FAM >= 2) {
            const R dr0 = rl - (R)P.br0[typ];
            const R ea = rexp<R>(-(R)P.bexpA[typ] * dr0);
            VA = -(R)P.bVAf[typ] * ea;
            dVA = (R)P.bVAf[typ] * (R)P.bexpA[typ] * ea;
            const R er = rexp<R>(-(R)P.bexpR[typ] * dr0);
            VR = (R)P.bVRf[typ] * er;
            dVR = -(R)P.bVRf[typ] * (R)P.bexpR[typ] * er;
        } else {
            R ea = rexp<R>(-(R)P.mu[typ] * rl);
            VA = -(R)P.B[typ] * ea;
            dVA = (R)P.B[typ] * (R)P.mu[typ] * ea;
            R er = rexp<R>(-(R)P.lambda[typ] * rl);
            VR = (R)P.A[typ] * er;
            dVR = -(R)P.A[typ] * (R)P.lambda[typ] * er;
        }
        R mf = (R)0.5 * (R)maskfac;
        VA *= mf;
        dVA *= mf;
        VR *= mf;
        dVR *= mf;
        bo_fn<R, FAM>(P, eli, typ, zij, fc, VA, bij, Pz);
    }

    // derivatives; forces on k and the bond-order screening factor of bond
    // ik (sum_j g h dE/dzeta_ij, bop_kernel.f90:1448-1462) reduced into lane k
    R dbidi[3] = {(R)0, (R)0, (R)0}, dbidj[3] = {(R)0, (R)0, (R)0};
    R wijb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) wijb[q] = (R)0;
    double recv[3] = {0.0, 0.0, 0.0}, sfac = 0.0;
    const R rlijr = (R)1 / rl;
    const R rij[3] = {rl * rn[0], rl * rn[1], rl * rn[2]};
    for (int k = 0; k < G; ++k) {
        if (k >= nbw) break;
        const int tk = gbase + k;
        const int tpk = s_typ[tk];
        R dk[3] = {(R)0, (R)0, (R)0};
        R zf = (R)0;
        if (on && k != lane && k < nb && tpk >= 0) {
            R rnk[3] = {s_rn[0][tk], s_rn[1][tk], s_rn[2][tk]};
            R rlik = s_rl[tk], fcik = s_fc[tk], dfcik = s_dfc[tk];
            R hv, dh;
            h_fn<R, FAM>(P, typ, tpk, rl - rlik, hv, dh);
            R costh = rnk[0] * rn[0] + rnk[1] * rn[1] + rnk[2] * rn[2];
            R g, dg;
            g_fn<R, FAM>(P, elc, tpk, costh, g, dg);
            zf = g * hv * Pz;
            // r_jk = r_k - r_j; its length cancels in d cos / d r_k and d r_j
            R dkc[3];
            for (int x = 0; x < 3; ++x) dkc[x] = rnk[x] * rlik - rn[x] * rl;
            R rlikr = (R)1 / rlik;
            R dcsdij = rlikr - costh * rlijr;
            R dcsdik = rlijr - costh * rlikr;
            R dcsdjk = -rlijr * rlikr;
            R dzfac = fcik * dg * hv;
            R dzdrij = g * fcik * dh;
            R dzdrik = g * (dfcik * hv - fcik * dh);
            R df[3];
            for (int x = 0; x < 3; ++x) {
                R dcsdi = -dcsdij * rn[x] - dcsdik * rnk[x];
                R dcsdj = dcsdij * rn[x] - dcsdjk * dkc[x];
                R dcsdk = dcsdik * rnk[x] + dcsdjk * dkc[x];
                dbidi[x] += -dzdrij * rn[x] - dzdrik * rnk[x] + dzfac * dcsdi;
                df[x] = dzdrij * rn[x] + dzfac * dcsdj;
                dbidj[x] += df[x];
                dk[x] = dzdrik * rnk[x] + dzfac * dcsdk;
            }
            R rik[3] = {rlik * rnk[0], rlik * rnk[1], rlik * rnk[2]};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) wijb[a + 3 * b] -= rij[a] * df[b] + rik[a] * dk[b];
        }
        double c0 = -(double)Pz * (double)dk[0];
        double c1 = -(double)Pz * (double)dk[1];
        double c2v = -(double)Pz * (double)dk[2];
        double c3 = (double)zf;
        c0 = group_sum<G>(c0);
        c1 = group_sum<G>(c1);
        c2v = group_sum<G>(c2v);
        c3 = group_sum<G>(c3);
        if (lane == k) {
            recv[0] = c0;
            recv[1] = c1;
            recv[2] = c2v;
            sfac = c3;
        }
    }

    double fi[3] = {0.0, 0.0, 0.0}, fj[3] = {0.0, 0.0, 0.0}, dfp[3] = {0.0, 0.0, 0.0};
    double e_ij = 0.0;
    double w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    if (on) {
        e_ij = (double)((R)0.5 * fc * (VR + bij * VA));
        R dffac = (R)0.5 * (dVR * fc + bij * dVA * fc + VR * dfc + bij * VA * dfc);
        for (int x = 0; x < 3; ++x) {
            double df = (double)(dffac * rn[x]);
            dfp[x] = df;
            fi[x] = df - (double)Pz * (double)dbidi[x];
            fj[x] = -df - (double)Pz * (double)dbidj[x];
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                w[a + 3 * b] = (double)rij[a] * (double)(dffac * rn[b]) - (double)Pz * (double)wijb[a + 3 * b];
    }
    if (act) {
        // screening forces of this bond's records: the pair term (when on)
        // and the bond order (factor sfac, zero unless the bond was a k-bond)
        const double dff = on ? 0.5 * ((double)VR + (double)bij * (double)VA) : 0.0;
        for (int q = roff[s]; q < roff[s + 1]; ++q) {
            const int kn = rkn[q];
            const int2 ek = nbr[kn];
            int a, b, c;
            dc_unpack(ek.y, a, b, c);
            double rik[3], rjk[3];
            for (int x = 0; x < 3; ++x) {
                rik[x] = (r[3 * ek.x + x] - r[3 * i + x]) - image_x(C, x, a, b, c);
                rjk[x] = -rijd[x] + rik[x];
            }
            const double4 co = rco[q];
            const double cik = dff * co.x + sfac * co.z, cjk = dff * co.y + sfac * co.w;
            double fk[3];
            for (int x = 0; x < 3; ++x) {
                const double d1 = cik * rik[x], d2 = cjk * rjk[x];
                fi[x] += d1;
                fj[x] += d2;
                fk[x] = -d1 - d2;
            }
            for (int a2 = 0; a2 < 3; ++a2)
                for (int b2 = 0; b2 < 3; ++b2) w[a2 + 3 * b2] += rik[a2] * (cik * rik[b2]) + rjk[a2] * (cjk * rjk[b2]);
            frec[3 * (size_t)q + 0] = fk[0];
            frec[3 * (size_t)q + 1] = fk[1];
            frec[3 * (size_t)q + 2] = fk[2];
        }
        fslot[3 * (size_t)s + 0] = fj[0] + recv[0];
        fslot[3 * (size_t)s + 1] = fj[1] + recv[1];
        fslot[3 * (size_t)s + 2] = fj[2] + recv[2];
        eslot[s] = e_ij;
        for (int q = 0; q < 9; ++q) wslot[9 * (size_t)s + q] = w[q];
        // per-bond properties: the pair term of the bond (when on) and the
        // virial of loop 2 plus both screening loops (bop_kernel.f90:1370-1373,
        // 1406-1408, 1511-1512, 1593-1594)
        if (epb) epb[s] += e_ij;
        if (fpb)
            for (int x = 0; x < 3; ++x) fpb[3 * (size_t)s + x] += dfp[x];
        if (wpb)
            for (int q = 0; q < 9; ++q) wpb[9 * (size_t)s + q] += w[q];
    }
    double f0 = group_sum<G>(fi[0]);
    double f1 = group_sum<G>(fi[1]);
    double f2 = group_sum<G>(fi[2]);
    if (valid && lane == 0) {
        fown[3 * (size_t)i + 0] = f0;
        fown[3 * (size_t)i + 1] = f1;
        fown[3 * (size_t)i + 2] = f2;
    }
    double red[10];
    red[0] = e_ij;
#pragma unroll
    for (int q = 0; q < 9; ++q) red[1 + q] = w[q];
#pragma unroll
    for (int q = 0; q < 10; ++q) red[q] = group_sum<64>(red[q]);
    const int wave = tid / 64;
    if ((tid & 63) == 0)
        for (int q = 0; q < 10; ++q) s_red[wave][q] = red[q];
    __syncthreads();
    if (tid < 10)
        partial[(size_t)blockIdx.x * 10 + tid] =
            ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
}

// wavefront per atom, lanes over its row: total force the atom's terms put
// on each neighbour entry (bond slot + records naming the entry, in record
// order)
__global__ void __launch_bounds__(256)
k_scr_slots(int nat, const int *__restrict__ offsets, const int *__restrict__ keep,
            const double *__restrict__ fslot, const int *__restrict__ roff, const int *__restrict__ rkn,
            const double *__restrict__ frec, double *__restrict__ fsl)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nat) return;
    const int beg = offsets[i], end = offsets[i + 1];
    const int q0 = roff[beg], q1 = roff[end];
    for (int s = beg + lane; s < end; s += 64) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (keep[s]) {
            a0 = fslot[3 * (size_t)s];
            a1 = fslot[3 * (size_t)s + 1];
            a2 = fslot[3 * (size_t)s + 2];
        }
        for (int q = q0; q < q1; ++q)
            if (rkn[q] == s) {
                a0 += frec[3 * (size_t)q];
                a1 += frec[3 * (size_t)q + 1];
                a2 += frec[3 * (size_t)q + 2];
            }
        fsl[3 * (size_t)s] = a0;
        fsl[3 * (size_t)s + 1] = a1;
        fsl[3 * (size_t)s + 2] = a2;
    }
}

// 8 lanes per atom over its row: f_a += own + sum of the slots the
// neighbours wrote for a; per-atom energy and virial halves of a's bonds and
// of the bonds pointing at a (pe(i), pe(j) of bop_kernel.f90:1366-1368)
constexpr int SCR_GL = 8;
__global__ void __launch_bounds__(256)
k_scr_gather(int nat, const int *__restrict__ offsets, const int *__restrict__ rev, const int *__restrict__ keep,
             const double *__restrict__ fsl, const double *__restrict__ fown, const double *__restrict__ eslot,
             const double *__restrict__ wslot, double *__restrict__ f, double *__restrict__ epa,
             double *__restrict__ wpa)
{
    const int lane = threadIdx.x % SCR_GL;
    const int a = blockIdx.x * (256 / SCR_GL) + threadIdx.x / SCR_GL;
    const bool ok = a < nat;
    double acc[3] = {0.0, 0.0, 0.0}, e = 0.0, w[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = 0.0;
    if (ok) {
        for (int s = offsets[a] + lane; s < offsets[a + 1]; s += SCR_GL) {
            const int t = rev[s];
            if (t >= 0) {
                acc[0] += fsl[3 * (size_t)t];
                acc[1] += fsl[3 * (size_t)t + 1];
                acc[2] += fsl[3 * (size_t)t + 2];
            }
            if (epa || wpa) {
                if (keep[s]) {
                    if (epa) e += eslot[s];
                    if (wpa)
                        for (int q = 0; q < 9; ++q) w[q] += wslot[9 * (size_t)s + q];
                }
                if (t >= 0 && keep[t]) {
                    if (epa) e += eslot[t];
                    if (wpa)
                        for (int q = 0; q < 9; ++q) w[q] += wslot[9 * (size_t)t + q];
                }
            }
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) acc[x] = group_sum<SCR_GL>(acc[x]);
    if (epa) e = group_sum<SCR_GL>(e);
    if (wpa)
        for (int q = 0; q < 9; ++q) w[q] = group_sum<SCR_GL>(w[q]);
    if (ok && lane == 0) {
        f[3 * (size_t)a] += fown[3 * (size_t)a] + acc[0];
        f[3 * (size_t)a + 1] += fown[3 * (size_t)a + 1] + acc[1];
        f[3 * (size_t)a + 2] += fown[3 * (size_t)a + 2] + acc[2];
        if (epa) epa[a] += 0.5 * e;
        if (wpa)
            for (int q = 0; q < 9; ++q) wpa[9 * (size_t)a + q] += 0.5 * w[q];
    }
}

// -------------------------------------------------------------------------
// host side
// -------------------------------------------------------------------------
static void tersoff_free(void *p) { delete static_cast<TersoffState *>(p); }

static agpu_tersoff_params default_params()
{
    // Tersoff_PRB_39_5566_SiC (tersoff_params.f90:99-138): elements C, Si
    agpu_tersoff_params p;
    std::memset(&p, 0, sizeof p);
    p.nel = 2;
    p.Z[0] = 6;
    p.Z[1] = 14;
    double A[3] = {1.3936e3, std::sqrt(1.3936e3 * 1.8308e3), 1.8308e3};
    double B[3] = {3.4674e2, std::sqrt(3.4674e2 * 4.7118e2), 4.7118e2};
    double xi[3] = {1.0, 0.9776e0, 1.0};
    double la[3] = {3.4879e0, (3.4879e0 + 2.4799e0) / 2, 2.4799e0};
    double mu[3] = {2.2119e0, (2.2119e0 + 1.7322e0) / 2, 1.7322e0};
    double r1[3] = {1.80, std::sqrt(1.80 * 2.70), 2.70};
    double r2[3] = {2.10, std::sqrt(2.10 * 3.00), 3.00};
    for (int k = 0; k < 6; ++k) {
        p.A[k] = k < 3 ? A[k] : 1.0;
        p.B[k] = k < 3 ? B[k] : 1.0;
        p.xi[k] = k < 3 ? xi[k] : 1.0;
        p.lambda[k] = k < 3 ? la[k] : 1.0;
        p.mu[k] = k < 3 ? mu[k] : 1.0;
        p.omega[k] = 1.0;
        p.mubo[k] = 0.0;
        p.m[k] = 1;
        p.r1[k] = k < 3 ? r1[k] : 1.0;
        p.r2[k] = k < 3 ? r2[k] : 2.0;
    }
    double beta[2] = {1.5724e-7, 1.1000e-6}, n[2] = {7.2751e-1, 7.8734e-1},
           c[2] = {3.8049e4, 1.0039e5}, d[2] = {4.3484e0, 1.6217e1}, h[2] = {-5.7058e-1, -5.9825e-1};
    for (int k = 0; k < 3; ++k) {
        p.beta[k] = k < 2 ? beta[k] : 1.0;
        p.n[k] = k < 2 ? n[k] : 1.0;
        p.c[k] = k < 2 ? c[k] : 1.0;
        p.d[k] = k < 2 ? d[k] : 1.0;
        p.h[k] = k < 2 ? h[k] : 1.0;
    }
    return p;
}

int tersoff_bind(agpu_potential *pot)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    agpu_particles *p = pot->p;
    const auto &prm = st->prm;
    for (int z = 0; z < 128; ++z) st->z2db[z] = -1;
    for (int e = 0; e < prm.nel; ++e) {
        if (prm.Z[e] <= 0 || prm.Z[e] >= 128)
            return set_error(AGPU_ERR_ELEMENT, "Tersoff: unknown element in the parameter set");
        st->z2db[prm.Z[e]] = e;
    }
    // request_interaction_range(nl, r2(pair)) for every element pair present
    // (default_bind_to_func.f90:123-144)
    bool present[3] = {false, false, false};
    for (int z : p->h_Z)
        if (z > 0 && z < 128 && st->z2db[z] >= 0) present[st->z2db[z]] = true;
    // screened: x(pair) * sqrt(maxval(max_cut_sq)) with x = sqrt(C_dr_cut)
    // (default_bind_to_func.f90:113-150)
    const bool ordered = st->dev.fam == 3;
    double mc_all = 0.0;
    if (st->scr)
        for (int q = 0; q < (ordered ? prm.nel * prm.nel : prm.nel * (prm.nel + 1) / 2); ++q)
            mc_all = std::max(mc_all, std::max(st->dev.r2[q], std::max(st->sp.or2[q], st->sp.bor2[q])));
    double cut = 0.0;
    for (int a = 0; a < prm.nel; ++a)
        for (int b = 0; b < prm.nel; ++b)
            if (present[a] && present[b]) {
                const int q = ordered ? b + a * prm.nel : pair_index(a, b, prm.nel);
                double r2 = ordered ? st->dev.r2[q] : prm.r2[q];
                if (st->scr) {
                    const double cm = st->sp.Cmax[q];
                    const double cdr = cm > 2.0 ? cm * cm / (4 * (cm - 1)) : 1.0;
                    r2 = std::sqrt(cdr) * mc_all;
                }
                if (r2 > cut) cut = r2;
                AGPU_CHECK(agpu_neighbors_request_interaction_range(pot->nl, r2));
            }
    st->cutoff = cut;
    AGPU_CHECK(map_elements(pot, st->z2db, 128));
    st->Z_rev = p->Z_rev;
    return AGPU_OK;
}

template <typename R, int G, int FAM>
static int launch_tersoff_fam(const char *name, TersoffState *st, const BopBufs &B, int nat, const int *order,
                              const double *r, const int *el, const CellArg &C, agpu_neighbors *nl,
                              const int *mask, double *eslot, double *wslot, const agpu_eval_args *a, int nblk,
                              hipStream_t s)
{
    AGPU_LAUNCH(name, (k_tersoff<R, G, FAM>), dim3(nblk), dim3(256), 0, s, st->dev, nat, order, r, el, C,
                nl->offsets.as<int>(), nl->nbr.as<int2>(), B.bslot, B.bcnt, mask, B.fslot, B.fown, eslot, wslot,
                B.partial, a->epot_per_bond, a->f_per_bond, a->wpot_per_bond);
    return AGPU_OK;
}

template <typename R, int G>
static int launch_tersoff(TersoffState *st, const BopBufs &B, int nat, const int *order, const double *r,
                          const int *el, const CellArg &C, agpu_neighbors *nl, const int *mask, double *eslot,
                          double *wslot, const agpu_eval_args *a, int nblk, hipStream_t s)
{
    const bool f64 = sizeof(R) == 8;
    switch (st->dev.fam) {
    case 3:
        return launch_tersoff_fam<R, G, 3>(f64 ? "juslin_fp64" : "juslin_mixed", st, B, nat, order, r, el, C, nl,
                                           mask, eslot, wslot, a, nblk, s);
    case 1:
        return launch_tersoff_fam<R, G, 1>(f64 ? "kumagai_fp64" : "kumagai_mixed", st, B, nat, order, r, el, C, nl,
                                           mask, eslot, wslot, a, nblk, s);
    case 2:
        return launch_tersoff_fam<R, G, 2>(f64 ? "brenner_fp64" : "brenner_mixed", st, B, nat, order, r, el, C, nl,
                                           mask, eslot, wslot, a, nblk, s);
    default:
        return launch_tersoff_fam<R, G, 0>(f64 ? "tersoff_fp64" : "tersoff_mixed", st, B, nat, order, r, el, C, nl,
                                           mask, eslot, wslot, a, nblk, s);
    }
}

// groups of G lanes, 256 / G atoms per block: the block count grows with G
static int bop_blocks(int nat, int G) { return std::max(1, (nat + 256 / G - 1) / (256 / G)); }

template <typename R>
static int dispatch_tersoff(int G, TersoffState *st, const BopBufs &B, int nat, const int *order, const double *r,
                            const int *el, const CellArg &C, agpu_neighbors *nl, const int *mask,
                            double *eslot, double *wslot, const agpu_eval_args *a, hipStream_t s)
{
    const int nblk = bop_blocks(nat, G);
    switch (G) {
    case 2: return launch_tersoff<R, 2>(st, B, nat, order, r, el, C, nl, mask, eslot, wslot, a, nblk, s);
    case 4: return launch_tersoff<R, 4>(st, B, nat, order, r, el, C, nl, mask, eslot, wslot, a, nblk, s);
    case 8: return launch_tersoff<R, 8>(st, B, nat, order, r, el, C, nl, mask, eslot, wslot, a, nblk, s);
    case 16: return launch_tersoff<R, 16>(st, B, nat, order, r, el, C, nl, mask, eslot, wslot, a, nblk, s);
    case 32: return launch_tersoff<R, 32>(st, B, nat, order, r, el, C, nl, mask, eslot, wslot, a, nblk, s);
    default: return launch_tersoff<R, 64>(st, B, nat, order, r, el, C, nl, mask, eslot, wslot, a, nblk, s);
    }
}

template <int G>
static int launch_gather_g(int nat, agpu_neighbors *nl, const BopBufs &B, double *eslot, double *wslot,
                           const agpu_eval_args *a, hipStream_t s)
{
    const int nblk = (nat + 256 / G - 1) / (256 / G);
    AGPU_LAUNCH("bop_gather", k_bop_gather<G>, dim3(nblk), dim3(256), 0, s, nat, nl->offsets.as<int>(), B.bslot,
                B.bcnt, B.rev, B.fslot, B.fown, eslot, wslot, a->f, a->epot_per_at, a->wpot_per_at);
    return AGPU_OK;
}

static int launch_gather(int G, int nat, agpu_neighbors *nl, const BopBufs &B, double *eslot, double *wslot,
                         const agpu_eval_args *a, hipStream_t s)
{
    switch (G) {
    case 2: return launch_gather_g<2>(nat, nl, B, eslot, wslot, a, s);
    case 4: return launch_gather_g<4>(nat, nl, B, eslot, wslot, a, s);
    case 8: return launch_gather_g<8>(nat, nl, B, eslot, wslot, a, s);
    case 16: return launch_gather_g<16>(nat, nl, B, eslot, wslot, a, s);
    case 32: return launch_gather_g<32>(nat, nl, B, eslot, wslot, a, s);
    default: return launch_gather_g<64>(nat, nl, B, eslot, wslot, a, s);
    }
}

// inner cutoffs from the family's device parameters (the ordered pairs of
// the Juslin family included)
static void derive_scr(ScrDev &S, const TersoffDev &D, const agpu_screening_params &sp)
{
    std::memset(&S, 0, sizeof S);
    for (int q = 0; q < 9; ++q) {
        S.r1[q] = D.r1[q];
        S.r2[q] = D.r2[q];
        S.or1[q] = sp.or1[q];
        S.or2[q] = sp.or2[q];
        S.bor1[q] = sp.bor1[q];
        S.bor2[q] = sp.bor2[q];
        S.fin[q] = S.r2[q] > S.r1[q] ? 1.0 / (S.r2[q] - S.r1[q]) : 0.0;
        S.fout[q] = S.or2[q] > S.or1[q] ? 1.0 / (S.or2[q] - S.or1[q]) : 0.0;
        S.fbo[q] = S.bor2[q] > S.bor1[q] ? 1.0 / (S.bor2[q] - S.bor1[q]) : 0.0;
        S.Cmin[q] = sp.Cmin[q];
        S.Cmax[q] = sp.Cmax[q];
        S.dC[q] = sp.Cmax[q] - sp.Cmin[q];
        // default_bind_to_func.f90:44-70
        S.C_dr_cut[q] = sp.Cmax[q] > 2.0 ? sp.Cmax[q] * sp.Cmax[q] / (4 * (sp.Cmax[q] - 1)) : 1.0;
        double mx = std::max(S.r2[q], std::max(S.or2[q], S.bor2[q]));
        S.max_cut_sq[q] = mx * mx;
    }
    // exp_cutoff_init (cutoff.f90:245-258)
    const double v1 = std::exp(-8.0), dv1 = -24 * v1, ddv1 = -48 * v1 - 24 * dv1;
    S.ec = (-3 * dv1 + ddv1) / 3;
    S.ed = (2 * dv1 - ddv1) / 4;
    S.efac2 = 1.0 / (1 - v1 - S.ec - S.ed);
    S.eoff = v1 + S.ec + S.ed;
}

template <typename R, int G>
static int launch_bop_scr(TersoffState *st, const ScrDev &S, int nat, const int *order, const double *r,
                          const int *el, const CellArg &C, agpu_neighbors *nl, const int *mask,
                          const agpu_eval_args *a, int nblk, hipStream_t s)
{
    const TersoffDev &P = st->dev;
    const int *offs = nl->offsets.as<int>();
    const int2 *nb = nl->nbr.as<int2>();
    const int *bs = st->bslot.as<int>(), *bc = st->bcnt.as<int>();
    const double4 *fcv = st->fcv.as<double4>();
    const int *roff = st->roff.as<int>(), *rkn = st->rkn.as<int>();
    const double4 *rco = st->rco.as<double4>();
    double *fslot = st->fslot.as<double>(), *frec = st->frec.as<double>(), *fown = st->fown.as<double>();
    double *es = st->eslot.as<double>(), *ws = st->wslot.as<double>(), *part = st->partial.as<double>();
    const bool f64 = sizeof(R) == 8;
    if (P.fam == 1)
        AGPU_LAUNCH(f64 ? "kumagai_scr_fp64" : "kumagai_scr_mixed", (k_bop_scr<R, G, 1>), dim3(nblk), dim3(256), 0,
                    s, P, S, nat, order, r, el, C, offs, nb, bs, bc, mask, fcv, roff, rkn, rco, fslot, frec, fown,
                    es, ws, part, a->epot_per_bond, a->f_per_bond, a->wpot_per_bond);
    else if (P.fam == 2)
        AGPU_LAUNCH(f64 ? "brenner_scr_fp64" : "brenner_scr_mixed", (k_bop_scr<R, G, 2>), dim3(nblk), dim3(256), 0,
                    s, P, S, nat, order, r, el, C, offs, nb, bs, bc, mask, fcv, roff, rkn, rco, fslot, frec, fown,
                    es, ws, part, a->epot_per_bond, a->f_per_bond, a->wpot_per_bond);
    else if (P.fam == 3)
        AGPU_LAUNCH(f64 ? "juslin_scr_fp64" : "juslin_scr_mixed", (k_bop_scr<R, G, 3>), dim3(nblk), dim3(256), 0,
                    s, P, S, nat, order, r, el, C, offs, nb, bs, bc, mask, fcv, roff, rkn, rco, fslot, frec, fown,
                    es, ws, part, a->epot_per_bond, a->f_per_bond, a->wpot_per_bond);
    else
        AGPU_LAUNCH(f64 ? "tersoff_scr_fp64" : "tersoff_scr_mixed", (k_bop_scr<R, G, 0>), dim3(nblk), dim3(256), 0,
                    s, P, S, nat, order, r, el, C, offs, nb, bs, bc, mask, fcv, roff, rkn, rco, fslot, frec, fown,
                    es, ws, part, a->epot_per_bond, a->f_per_bond, a->wpot_per_bond);
    return AGPU_OK;
}

template <typename R>
static int dispatch_bop_scr(int G, TersoffState *st, const ScrDev &S, int nat, const int *order, const double *r,
                            const int *el, const CellArg &C, agpu_neighbors *nl, const int *mask,
                            const agpu_eval_args *a, hipStream_t s, int *nblk_out)
{
    const int nblk = std::max(1, (nat + 256 / G - 1) / (256 / G));
    *nblk_out = nblk;
    AGPU_CHECK(st->partial.reserve(sizeof(double) * 10 * nblk));
    switch (G) {
    case 2: return launch_bop_scr<R, 2>(st, S, nat, order, r, el, C, nl, mask, a, nblk, s);
    case 4: return launch_bop_scr<R, 4>(st, S, nat, order, r, el, C, nl, mask, a, nblk, s);
    case 8: return launch_bop_scr<R, 8>(st, S, nat, order, r, el, C, nl, mask, a, nblk, s);
    case 16: return launch_bop_scr<R, 16>(st, S, nat, order, r, el, C, nl, mask, a, nblk, s);
    case 32: return launch_bop_scr<R, 32>(st, S, nat, order, r, el, C, nl, mask, a, nblk, s);
    default: return launch_bop_scr<R, 64>(st, S, nat, order, r, el, C, nl, mask, a, nblk, s);
    }
}

// the screened variants (bop_kernel.f90 with SCREENING); bound and list
// updated by the caller
static int tersoff_scr_compute(agpu_potential *pot, const agpu_eval_args *a)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    if (capturing())
        return set_error(AGPU_ERR_UNSUPPORTED, "screened bond-order potentials are not graph-capturable");
    const int nat = p->nat;
    hipStream_t s = stream();
    const int64_t E = nl->total;
    if (E > 0x7ffffff0ll) return set_error(AGPU_ERR_ARG, "screened bond order: list too long");
    ScrDev S;
    derive_scr(S, st->dev, st->sp);
    AGPU_CHECK(st->keep.reserve(sizeof(int) * (E + 1)));
    AGPU_CHECK(st->rcnt.reserve(sizeof(int) * (E + 1)));
    AGPU_CHECK(st->roff.reserve(sizeof(int) * (E + 1)));
    AGPU_CHECK(st->fcv.reserve(sizeof(double4) * (E + 1)));
    AGPU_CHECK(st->bslot.reserve(sizeof(int) * (E + 1)));
    AGPU_CHECK(st->fslot.reserve(sizeof(double) * 3 * (E + 1)));
    AGPU_CHECK(st->fsl.reserve(sizeof(double) * 3 * (E + 1)));
    AGPU_CHECK(st->eslot.reserve(sizeof(double) * (E + 1)));
    AGPU_CHECK(st->wslot.reserve(sizeof(double) * 9 * (E + 1)));
    AGPU_CHECK(st->nlrev.reserve(sizeof(int) * (E + 1)));
    AGPU_CHECK(st->bcnt.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(st->fown.reserve(sizeof(double) * 3 * (nat + 1)));
    AGPU_CHECK(st->stats.reserve(sizeof(int) * 4));
    AGPU_CHECK(pot->stats.init());
    int *dstats = st->stats.as<int>();
    AGPU_CHECK(zero_dev(dstats, sizeof(int) * 4));
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const int *el = pot->d_el.as<int>();
    const int *offs = nl->offsets.as<int>();
    const int2 *nbr = nl->nbr.as<int2>();
    const dim3 gw((unsigned)((nat + SCR_WAVES - 1) / SCR_WAVES));
    // pass 1: keep flags, cutoffs and record counts; exclusive scan to record
    // offsets
    AGPU_CHECK(zero_dev(st->rcnt.as<int>() + E, sizeof(int)));
    AGPU_LAUNCH("scr_count", k_scr_screen<false>, gw, dim3(256), 0, s, S, nat, p->d_r, el, st->prm.nel,
                (int)(st->dev.fam == 3), C, offs,
                nbr, st->keep.as<int>(), st->rcnt.as<int>(), nullptr, st->fcv.as<double4>(), nullptr, nullptr,
                dstats);
    size_t tmp = 0;
    AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, st->rcnt.as<int>(), st->roff.as<int>(), 0, (size_t)E + 1,
                                     rocprim::plus<int>(), s));
    AGPU_CHECK(st->scan_tmp.reserve(tmp + 16));
    AGPU_HIP(rocprim::exclusive_scan(st->scan_tmp.ptr, tmp, st->rcnt.as<int>(), st->roff.as<int>(), 0,
                                     (size_t)E + 1, rocprim::plus<int>(), s));
    AGPU_HIP(hipMemcpyAsync(pot->stats.h, st->roff.as<int>() + E, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipMemcpyAsync(pot->stats.h + 1, dstats, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (pot->stats.h[1] & DERR_SCR_ROW)
        return set_error(AGPU_ERR_NEB, "screened bond order: more than 256 neighbours within the screening range");
    const int64_t nrec = pot->stats.h[0];
    AGPU_CHECK(st->rkn.reserve(sizeof(int) * (nrec + 1)));
    AGPU_CHECK(st->rco.reserve(sizeof(double4) * (nrec + 1)));
    AGPU_CHECK(st->frec.reserve(sizeof(double) * 3 * (nrec + 1)));
    // pass 2: the records (rows with none return at once)
    AGPU_LAUNCH("scr_fill", k_scr_screen<true>, gw, dim3(256), 0, s, S, nat, p->d_r, el, st->prm.nel,
                (int)(st->dev.fam == 3), C, offs, nbr,
                st->keep.as<int>(), st->rcnt.as<int>(), st->roff.as<int>(), st->fcv.as<double4>(),
                st->rkn.as<int>(), st->rco.as<double4>(), dstats);
    AGPU_LAUNCH("scr_bonds", k_scr_bonds, dim3((nat + 255) / 256), dim3(256), 0, s, nat, offs, st->keep.as<int>(),
                st->bslot.as<int>(), st->bcnt.as<int>());
    // reverse entries of the list, once per list build
    if (st->nlrev_nl != nl || st->nlrev_upd != nl->n_updates) {
        AGPU_LAUNCH("scr_rev", k_nl_rev_entries, gw, dim3(256), 0, s, nat, offs, nbr, st->nlrev.as<int>(), dstats);
        st->nlrev_nl = nl;
        st->nlrev_upd = nl->n_updates;
    }
    AGPU_CHECK(int_max(st->bcnt.as<int>(), nat, dstats + 2));
    AGPU_HIP(hipMemcpyAsync(pot->stats.h, dstats, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (pot->stats.h[1] & DERR_BOND_OVERFLOW) {
        st->nlrev_nl = nullptr;
        return set_error(AGPU_ERR_NEB, "screened bond order: list entry without its reverse entry");
    }
    const int maxb = pot->stats.h[2];
    if (maxb > 64)
        return set_error(AGPU_ERR_NEB, "screened bond order: more than 64 bonds on one atom (nebmax too small)");
    int G = 2;
    while (G < maxb) G <<= 1;
    int nblk = 0;
    const int *order = nl->idx_sorted.as<int>();
    if (pot->precision == AGPU_PREC_MIXED)
        AGPU_CHECK(dispatch_bop_scr<float>(G, st, S, nat, order, p->d_r, el, C, nl, a->mask, a, s, &nblk));
    else
        AGPU_CHECK(dispatch_bop_scr<double>(G, st, S, nat, order, p->d_r, el, C, nl, a->mask, a, s, &nblk));
    AGPU_LAUNCH("scr_slots", k_scr_slots, gw, dim3(256), 0, s, nat, offs, st->keep.as<int>(),
                st->fslot.as<double>(), st->roff.as<int>(), st->rkn.as<int>(), st->frec.as<double>(),
                st->fsl.as<double>());
    AGPU_LAUNCH("scr_gather", k_scr_gather, dim3((unsigned)((nat + 256 / SCR_GL - 1) / (256 / SCR_GL))), dim3(256), 0,
                s, nat, offs, st->nlrev.as<int>(), st->keep.as<int>(), st->fsl.as<double>(), st->fown.as<double>(),
                st->eslot.as<double>(), st->wslot.as<double>(), a->f, a->epot_per_at, a->wpot_per_at);
    AGPU_CHECK(reduce_ev(st->partial.as<double>(), nblk, 10, a->epot, a->wpot, nullptr));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

int bop_variant(agpu_potential *pot, int k)
{
    if (!pot || !pot->impl) return 0;
    switch (pot->kind) {
    case AGPU_KIND_TERSOFF:
    case AGPU_KIND_KUMAGAI:
    case AGPU_KIND_BRENNER:
    case AGPU_KIND_JUSLIN: break;
    default: return 0;
    }
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    if (st->scr) return 0;
    if (k >= 0) st->compact = k == 1;
    return 2;
}

bool tersoff_screened(const agpu_potential *pot)
{
    return pot && pot->impl && static_cast<const TersoffState *>(pot->impl)->scr;
}

// Per-call scratch of the row path: slot forces by list entry (3 per entry
// of the list's entry space), own forces (3 per atom), per-block records
// (10 per block), second-level reduction records (NULL: the library's own),
// 4 flag words and the elements by slot (1 per atom).
struct RowBufs {
    double *frev, *fown, *partial, *red;
    int *stats, *el_s;
};

// the row path's flags cleared and its elements put in slot order (slot p
// holds caller atom order[p])
__global__ void k_bop_prep(int nat, const int *__restrict__ order, const int *__restrict__ el,
                           int *__restrict__ el_s, int *__restrict__ stats)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < 4) stats[p] = 0;
    if (p < nat) el_s[p] = el[order[p]];
}

// the reverse-entry search may read G entries of a row as aligned vectors:
// staged rows (every row has room for S entries) of an even stride S >= G
static int rows_vector_ok(const agpu_neighbors *nl, int G)
{
    const NlRows L = nl_rows(nl);
    return L.S >= G && L.S % 2 == 0 ? 1 : 0;
}

template <typename R, int G, int FAM, bool WB, bool HT = true>
static int launch_rows_k(const char *name, TersoffState *st, const RowBufs &B, agpu_particles *p,
                         agpu_neighbors *nl, const int *el, const agpu_eval_args *a, double *eslot, double *erev,
                         double *wslot, double *wrev, int nblk, hipStream_t s)
{
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    AGPU_LAUNCH(name, (k_bop_rows<R, G, FAM, WB, HT>), dim3(nblk), dim3(256), 0, s, st->dev, p->nat,
                nl->idx_sorted.as<int>(), B.el_s, C, nl_rows(nl), nl->offsets.as<int>(), a->mask, B.frev, B.fown,
                eslot, erev, wslot, wrev, B.partial, a->epot_per_bond, a->f_per_bond, a->wpot_per_bond, B.stats,
                nl->pending ? nl_flags(nl) : nullptr, rows_vector_ok(nl, G));
    return AGPU_OK;
}

// true when some h-term of the parameter set differs from 1: mubo of the
// Tersoff and Brenner forms, alpha of Kumagai; Juslin's omega exp(...) always
static bool has_hterm(const TersoffDev &D)
{
    if (D.fam == 3) return true;
    for (int k = 0; k < 6; ++k)
        if ((D.fam == 1 ? D.kalpha[k] : D.mubo[k]) != 0.0) return true;
    return false;
}

template <typename R, int G, int FAM>
static int launch_rows_wb(const char *name, bool wb, TersoffState *st, const RowBufs &B, agpu_particles *p,
                          agpu_neighbors *nl, const int *el, const agpu_eval_args *a, double *eslot, double *erev,
                          double *wslot, double *wrev, int nblk, hipStream_t s)
{
    // the h-term-free forms of the Tersoff family (FAM 0) without per-bond virials
    if constexpr (FAM == 0)
        if (!wb && !has_hterm(st->dev))
            return launch_rows_k<R, G, FAM, false, false>(name, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk,
                                                          s);
    if (wb) return launch_rows_k<R, G, FAM, true>(name, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
    return launch_rows_k<R, G, FAM, false>(name, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
}

template <typename R, int G>
static int launch_rows_g(bool wb, TersoffState *st, const RowBufs &B, agpu_particles *p, agpu_neighbors *nl,
                         const int *el, const agpu_eval_args *a, double *eslot, double *erev, double *wslot,
                         double *wrev, int nblk, hipStream_t s)
{
    const bool f64 = sizeof(R) == 8;
    switch (st->dev.fam) {
    case 3:
        return launch_rows_wb<R, G, 3>(f64 ? "juslin_fp64" : "juslin_mixed", wb, st, B, p, nl, el, a, eslot, erev,
                                       wslot, wrev, nblk, s);
    case 1:
        return launch_rows_wb<R, G, 1>(f64 ? "kumagai_fp64" : "kumagai_mixed", wb, st, B, p, nl, el, a, eslot,
                                       erev, wslot, wrev, nblk, s);
    case 2:
        return launch_rows_wb<R, G, 2>(f64 ? "brenner_fp64" : "brenner_mixed", wb, st, B, p, nl, el, a, eslot,
                                       erev, wslot, wrev, nblk, s);
    default:
        return launch_rows_wb<R, G, 0>(f64 ? "tersoff_fp64" : "tersoff_mixed", wb, st, B, p, nl, el, a, eslot,
                                       erev, wslot, wrev, nblk, s);
    }
}

template <typename R>
static int dispatch_rows(int G, bool wb, TersoffState *st, const RowBufs &B, agpu_particles *p, agpu_neighbors *nl,
                         const int *el, const agpu_eval_args *a, double *eslot, double *erev, double *wslot,
                         double *wrev, int nblk, hipStream_t s)
{
    switch (G) {
    case 4: return launch_rows_g<R, 4>(wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
    case 8: return launch_rows_g<R, 8>(wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
    case 16: return launch_rows_g<R, 16>(wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
    case 32: return launch_rows_g<R, 32>(wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
    default: return launch_rows_g<R, 64>(wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s);
    }
}

// group width of the row path: the longest row, at least 4 lanes; over an
// unchecked build (whose longest row is not known yet) the longest row of
// the build before, and at least the width a row of an unchecked build
// once outgrew
static int rows_width(const agpu_neighbors *nl, int floor_g = 0)
{
    int G = 4;
    while (G < nl->maxrow) G <<= 1;
    return nl->pending ? std::min(64, std::max(G, floor_g)) : G;
}

// entries the per-entry buffers of the row path must hold: the list's
// total, or over an unchecked build the staging rows' bound (the kernels
// write no entry of a build that overflowed them)
static int64_t rows_entries(const agpu_neighbors *nl)
{
    return nl->pending ? (int64_t)nl->nat * nl->stage_s : nl->total;
}

// The row path of the unscreened bond-order classes over the current list
// (rows of at most 64 entries): the bond-order kernel, the gather and the
// fixed-order energy / virial sum.  Flags are read back unless capturing,
// and with them a deferred list build's (nl_settle): *again = true when the
// call has to be made again (the build was unusable, or a row was longer
// than the lanes the call took).
static int bop_rows_run(agpu_potential *pot, const agpu_eval_args *a, const RowBufs &B, double *eslot,
                        double *erev, double *wslot, double *wrev, int *h_flags, bool *again = nullptr)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    const int nat = p->nat;
    hipStream_t s = stream();
    const int G = rows_width(nl, st->row_g);
    const int nblk = bop_blocks(nat, G);
    const bool wb = wslot || a->wpot_per_bond;
    const int *el = pot->d_el.as<int>();
    const int *order = nl->idx_sorted.as<int>();
    // per-bond arrays are indexed like the reference's list: caller order
    if (a->epot_per_bond || a->f_per_bond || a->wpot_per_bond) AGPU_CHECK(nl_caller_offsets(nl));
    AGPU_LAUNCH("bop_prep", k_bop_prep, dim3((std::max(nat, 4) + 255) / 256), dim3(256), 0, s, nat, order, el,
                B.el_s, B.stats);
    if (pot->precision == AGPU_PREC_MIXED)
        AGPU_CHECK(dispatch_rows<float>(G, wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s));
    else
        AGPU_CHECK(dispatch_rows<double>(G, wb, st, B, p, nl, el, a, eslot, erev, wslot, wrev, nblk, s));
    AGPU_LAUNCH("bop_gather", k_bop_rows_gather, dim3((nat + 255) / 256), dim3(256), 0, s, nat, order,
                nl_rows(nl), B.frev, B.fown, eslot, erev, wslot, wrev, a->f, a->epot_per_at, a->wpot_per_at,
                (const int *)B.stats);
    AGPU_CHECK(reduce_ev(B.partial, nblk, 10, a->epot, a->wpot, nullptr, true, B.red, B.stats));
    if (capturing()) return AGPU_OK;     // the MD driver reads the flags after the replay
    AGPU_HIP(hipMemcpyAsync(h_flags, B.stats, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    const bool deferred = nl->pending;
    bool redo = false;
    AGPU_CHECK(nl_settle(nl, &redo));
    if (redo || (h_flags[0] & DERR_REDO)) {
        if (!deferred || !again)
            return set_error(AGPU_ERR_STATE, "Tersoff: a row longer than the lanes of its group");
        if (!redo) st->row_g = std::max(st->row_g, rows_width(nl));
        *again = true;
        return AGPU_OK;
    }
    if (h_flags[0] & DERR_REV_MISSING)
        return set_error(AGPU_ERR_NEB, "Tersoff: bond without its reverse bond (inconsistent list)");
    return AGPU_OK;
}

static int bop_rows_compute(agpu_potential *pot, const agpu_eval_args *a, bool *again)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    const int nat = pot->p->nat;
    const int64_t total = rows_entries(pot->nl);
    AGPU_CHECK(st->fslot.reserve(sizeof(double) * 3 * (total + 1)));
    AGPU_CHECK(st->fown.reserve(sizeof(double) * 3 * (nat + 1)));
    AGPU_CHECK(st->el_s.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(st->stats.reserve(sizeof(int) * 4));
    AGPU_CHECK(st->partial.reserve(sizeof(double) * 10 * bop_blocks(nat, rows_width(pot->nl, st->row_g))));
    AGPU_CHECK(pot->stats.init());
    double *eslot = nullptr, *erev = nullptr, *wslot = nullptr, *wrev = nullptr;
    if (a->epot_per_at) {
        AGPU_CHECK(st->eslot.reserve(sizeof(double) * 2 * (total + 1)));
        eslot = st->eslot.as<double>();
        erev = eslot + total + 1;
    }
    if (a->wpot_per_at) {
        AGPU_CHECK(st->wslot.reserve(sizeof(double) * 18 * (total + 1)));
        wslot = st->wslot.as<double>();
        wrev = wslot + 9 * (total + 1);
    }
    RowBufs B{st->fslot.as<double>(), st->fown.as<double>(), st->partial.as<double>(), nullptr, st->stats.as<int>(),
              st->el_s.as<int>()};
    return bop_rows_run(pot, a, B, eslot, erev, wslot, wrev, pot->stats.h, again);
}

// bonds of every row (k_bop_bonds) into B.bslot / B.bcnt
static int bop_bonds_launch(agpu_potential *pot, const BopBufs &B)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const int nat = p->nat;
    AGPU_LAUNCH("bop_bonds", k_bop_bonds, dim3((nat + 255) / 256), dim3(256), 0, stream(), nat, p->d_r,
                pot->d_el.as<int>(), C, st->dev, nl->offsets.as<int>(), nl->nbr.as<int2>(), B.bslot, B.bcnt,
                B.stats);
    return AGPU_OK;
}

// reverse slots, the bond-order kernel (G lanes per atom, G >= every bond
// count), the owner-computes gather and the fixed-order energy / virial sum
static int bop_plain_tail(agpu_potential *pot, const agpu_eval_args *a, const BopBufs &B, int G, double *eslot,
                          double *wslot)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    const int nat = p->nat;
    hipStream_t s = stream();
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const int *el = pot->d_el.as<int>();
    AGPU_LAUNCH("bop_rev", k_bop_rev, dim3((unsigned)(((long long)nat * G + 255) / 256)), dim3(256), 0, s, nat, G,
                nl->offsets.as<int>(), nl->nbr.as<int2>(), B.bslot, B.bcnt, B.rev, B.stats);
    const int *order = nl->idx_sorted.as<int>();
    if (pot->precision == AGPU_PREC_MIXED)
        AGPU_CHECK(dispatch_tersoff<float>(G, st, B, nat, order, p->d_r, el, C, nl, a->mask, eslot, wslot, a, s));
    else
        AGPU_CHECK(dispatch_tersoff<double>(G, st, B, nat, order, p->d_r, el, C, nl, a->mask, eslot, wslot, a, s));
    AGPU_CHECK(launch_gather(G, nat, nl, B, eslot, wslot, a, s));
    // per-block records [e, W(9)] -> epot, wpot
    return reduce_ev(B.partial, bop_blocks(nat, G), 10, a->epot, a->wpot, nullptr, false, B.red);
}

int tersoff_compute(agpu_potential *pot, const agpu_eval_args *a)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    if (st->Z_rev != p->Z_rev) AGPU_CHECK(tersoff_bind(pot));
    // the row path reads the rows where the builder left them.  When the
    // last build's rows fit it (and no per-bond array is asked for, which
    // the bond-order kernel writes itself), a new build's flags are checked
    // with this call's own: one host round trip per call
    const bool rows = !st->compact && !st->scr;
    const bool defer = rows && nl->maxrow > 0 && nl->maxrow <= 64 && !a->epot_per_bond && !a->f_per_bond &&
                       !a->wpot_per_bond;
    AGPU_CHECK(nl_update(nl, p, 0, false, defer));
    const int nat = p->nat;
    hipStream_t s = stream();
    if (nat == 0) return zero_dev(a->epot, sizeof(double));
    // the row path sets epot from its block records; the others add to it
    if (rows && (nl->pending || nl->maxrow <= 64)) {
        bool again = false;
        AGPU_CHECK(bop_rows_compute(pot, a, &again));
        if (!again) return AGPU_OK;
        // the build is redone synchronously when it was unusable (radix
        // sort, a wider staging stride), then the call with its rows known
        if (!nl_current(nl, p)) AGPU_CHECK(nl_update(nl, p, 1, false));
        return tersoff_compute(pot, a);
    }
    AGPU_CHECK(zero_dev(a->epot, sizeof(double)));
    if (st->scr) {
        AGPU_CHECK(nl_update(nl, p, 0));
        return tersoff_scr_compute(pot, a);
    }
    // longer rows: compact them to their bonds (at most 64) over CSR rows
    AGPU_CHECK(nl_update(nl, p, 0));
    const int64_t total = nl->total;
    AGPU_CHECK(st->bslot.reserve(sizeof(int) * (total + 1)));
    AGPU_CHECK(st->rev.reserve(sizeof(int) * (total + 1)));
    AGPU_CHECK(st->fslot.reserve(sizeof(double) * 3 * (total + 1)));
    AGPU_CHECK(st->bcnt.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(st->fown.reserve(sizeof(double) * 3 * (nat + 1)));
    AGPU_CHECK(st->stats.reserve(sizeof(int) * 4));
    // records for the widest group (64 lanes per atom: the most blocks)
    AGPU_CHECK(st->partial.reserve(sizeof(double) * 10 * bop_blocks(nat, 64)));
    AGPU_CHECK(pot->stats.init());
    double *eslot = nullptr, *wslot = nullptr;
    if (a->epot_per_at) {
        AGPU_CHECK(st->eslot.reserve(sizeof(double) * (total + 1)));
        eslot = st->eslot.as<double>();
    }
    if (a->wpot_per_at) {
        AGPU_CHECK(st->wslot.reserve(sizeof(double) * 9 * (total + 1)));
        wslot = st->wslot.as<double>();
    }
    BopBufs B{st->bslot.as<int>(), st->bcnt.as<int>(), st->rev.as<int>(), st->stats.as<int>(),
              st->fslot.as<double>(), st->fown.as<double>(), st->partial.as<double>(), nullptr};
    int *dstats = B.stats;
    AGPU_CHECK(zero_dev(dstats, sizeof(int) * 4));
    AGPU_CHECK(bop_bonds_launch(pot, B));
    // group width: the row maximum bounds the bond count without a host
    // round trip when it is small; otherwise read the exact bond maximum
    int maxb = nl->maxrow;
    if (maxb > 64 && capturing())
        return set_error(AGPU_ERR_UNSUPPORTED, "bond-order step capture needs at most 64 neighbours per atom");
    if (maxb > 64) {
        AGPU_CHECK(int_max(B.bcnt, nat, dstats));
        AGPU_HIP(hipMemcpyAsync(pot->stats.h, dstats, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        AGPU_HIP(hipStreamSynchronize(s));
        maxb = pot->stats.h[0];
        if (maxb > 64)
            return set_error(AGPU_ERR_NEB, "Tersoff: more than 64 bonds on one atom (nebmax too small)");
    }
    int G = 2;
    while (G < maxb) G <<= 1;
    AGPU_CHECK(bop_plain_tail(pot, a, B, G, eslot, wslot));
    if (capturing()) return AGPU_OK;     // the MD driver reads the flags after the replay
    AGPU_HIP(hipMemcpyAsync(pot->stats.h, dstats, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (pot->stats.h[1] & DERR_BOND_OVERFLOW)
        return set_error(AGPU_ERR_NEB, "Tersoff: bond without its reverse bond (inconsistent list)");
    return AGPU_OK;
}

// -------------------------------------------------------------------------
// Allocation-free energy_and_forces of the unscreened bond-order classes.
// The per-call scratch (bond slots, reverse slots, slot and own forces,
// per-block records, flags) lives in a file-scope device array of fixed
// size instead of library buffers, for callers that cannot allocate device
// memory.  Every device holds its own copy of the array, at its own
// address, until the process ends; host code uses the address
// hipGetSymbolAddress returns on that device, never the array's name.
// All calls on one device share its array: calls that overlap on one
// device (other streams, other host threads) race.
// -------------------------------------------------------------------------
constexpr size_t NOALLOC_BYTES = 64ull << 20;
constexpr int NOALLOC_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[NOALLOC_DEVICES] = {};

// the address of this device's copy of scratch, looked up once per device
static int scratch_base(unsigned char **out)
{
    int dev = 0;
    AGPU_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_DEVICES) return set_error(AGPU_ERR_ARG, "noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(AGPU_ERR_HIP, std::string("noalloc: hipGetSymbolAddress(scratch): ") +
                                               hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return AGPU_OK;
}

// 256-byte aligned sub-buffers of the scratch array, in order; returns the
// bytes the layout needs (compared with NOALLOC_BYTES before any use)
static size_t noalloc_layout(unsigned char *base, int64_t total, int nat, int G, RowBufs *B)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base ? base + off : nullptr;
        off += (bytes + 255) & ~size_t(255);
        return p;
    };
    const int nblk = bop_blocks(nat, G);
    const int nred = (nblk + 255) / 256;
    B->frev = reinterpret_cast<double *>(take(sizeof(double) * 3 * (size_t)(total + 1)));
    B->fown = reinterpret_cast<double *>(take(sizeof(double) * 3 * (size_t)(nat + 1)));
    B->partial = reinterpret_cast<double *>(take(sizeof(double) * 10 * (size_t)nblk));
    B->red = reinterpret_cast<double *>(take(sizeof(double) * 10 * (size_t)nred));
    B->stats = reinterpret_cast<int *>(take(sizeof(int) * 4));
    B->el_s = reinterpret_cast<int *>(take(sizeof(int) * (size_t)(nat + 1)));
    return off;
}

// The row path with its scratch in the array above
static int tersoff_compute_noalloc(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, double *epot,
                                   double *f, double *wpot, const int *mask)
{
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    if (st->scr) return set_error(AGPU_ERR_UNSUPPORTED, "noalloc: the screened variants are not covered");
    if (st->Z_rev != p->Z_rev)
        return set_error(AGPU_ERR_STATE, "noalloc: atomic numbers changed since bind_to");
    if (!nl_current(nl, p))
        return set_error(AGPU_ERR_STATE, "noalloc: the neighbour list is not current (agpu_neighbors_update first)");
    const int nat = p->nat;
    if (nl->maxrow > 64)
        return set_error(AGPU_ERR_UNSUPPORTED, "noalloc: rows of more than 64 neighbours");
    RowBufs B;
    const int G = rows_width(nl);
    const size_t need = noalloc_layout(nullptr, nl->total, nat, G, &B);
    if (need > NOALLOC_BYTES)
        return set_error(AGPU_ERR_ARG, "noalloc: " + std::to_string(nat) + " atoms and " +
                                           std::to_string(nl->total) + " list entries need " +
                                           std::to_string(need) + " bytes of scratch; the maximum is " +
                                           std::to_string(NOALLOC_BYTES) +
                                           " (24 per list entry + about 30 per atom)");
    unsigned char *base = nullptr;
    AGPU_CHECK(scratch_base(&base));
    noalloc_layout(base, nl->total, nat, G, &B);
    // slot positions of a held list follow the atoms (no allocation)
    AGPU_CHECK(nl_slots_current(nl, p));
    AGPU_CHECK(zero_dev(epot, sizeof(double)));
    if (nat == 0) return AGPU_OK;
    agpu_eval_args a{};
    a.epot = epot;
    a.f = f;
    a.wpot = wpot;
    a.mask = mask;
    int h[4] = {0, 0, 0, 0};
    // the flags start at zero (in bop_rows_run); every other buffer is
    // written before it is read
    return bop_rows_run(pot, &a, B, nullptr, nullptr, nullptr, nullptr, h);
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_potential_energy_and_forces_noalloc(agpu_potential *pot, agpu_particles *p,
                                                        agpu_neighbors *nl, double *epot, double *f, double *wpot,
                                                        const int *mask)
{
    if (!pot || !p || !nl || !epot || !f || !wpot) return set_error(AGPU_ERR_ARG, "noalloc: bad arguments");
    if (!pot->bound || pot->p != p || pot->nl != nl)
        return set_error(AGPU_ERR_STATE, "noalloc: potential not bound to these objects");
    switch (pot->kind) {
    case AGPU_KIND_TERSOFF:
    case AGPU_KIND_KUMAGAI:
    case AGPU_KIND_BRENNER:
    case AGPU_KIND_JUSLIN: return tersoff_compute_noalloc(pot, p, nl, epot, f, wpot, mask);
    default: return set_error(AGPU_ERR_UNSUPPORTED, "noalloc: only the unscreened bond-order classes");
    }
}

extern "C" int agpu_tersoff_create(agpu_potential **pot, const agpu_tersoff_params *params)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    TersoffState *st = new TersoffState();
    st->prm = params ? *params : default_params();
    if (st->prm.nel < 1 || st->prm.nel > 3) {
        delete st;
        return set_error(AGPU_ERR_ARG, "Tersoff: 1 to 3 elements supported");
    }
    derive(st->dev, st->prm);
    agpu_potential *p = new agpu_potential();
    p->kind = AGPU_KIND_TERSOFF;
    for (int k = 0; k < st->prm.nel * (st->prm.nel + 1) / 2; ++k)
        if (st->prm.r2[k] > p->cutoff) p->cutoff = st->prm.r2[k];
    p->impl = st;
    p->free_impl = tersoff_free;
    *pot = p;
    return AGPU_OK;
}

extern "C" int agpu_tersoff_set_params(agpu_potential *pot, const agpu_tersoff_params *params)
{
    if (!pot || pot->kind != AGPU_KIND_TERSOFF || !params) return set_error(AGPU_ERR_ARG, "bad arguments");
    TersoffState *st = static_cast<TersoffState *>(pot->impl);
    st->prm = *params;
    derive(st->dev, st->prm);
    pot->cutoff = 0.0;
    for (int k = 0; k < st->prm.nel * (st->prm.nel + 1) / 2; ++k)
        if (st->prm.r2[k] > pot->cutoff) pot->cutoff = st->prm.r2[k];
    if (pot->bound) return tersoff_bind(pot);
    return AGPU_OK;
}

// Kumagai_CompMaterSci_39_457_Si (kumagai_params.f90:97-132)
static agpu_kumagai_params kumagai_default()
{
    agpu_kumagai_params k;
    k.Z = 14;
    k.A = 3281.5905;
    k.B = 121.00047;
    k.lambda1 = 3.2300135;
    k.lambda2 = 1.3457970;
    k.eta = 1.0;
    k.delta = 0.53298909;
    k.alpha = 2.3890327;
    k.beta = 1;
    k.c1 = 0.20173476;
    k.c2 = 730418.72;
    k.c3 = 1000000.0;
    k.c4 = 1.0;
    k.c5 = 26.0;
    k.h = -0.36500000;
    k.r1 = 2.70;
    k.r2 = 3.30;
    return k;
}

// Kumagai (INIT_FUNC, kumagai_module.f90:30-78): one element, the generic bond-
// order kernel with the Kumagai g, h and bond-order forms.
extern "C" int agpu_kumagai_create(agpu_potential **pot, const agpu_kumagai_params *params)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    const agpu_kumagai_params k = params ? *params : kumagai_default();
    if (!(k.r2 > k.r1 && k.r1 > 0.0) || k.Z <= 0 || k.Z >= 128)
        return set_error(AGPU_ERR_ARG, "Kumagai: need 0 < r1 < r2 and a valid element");
    TersoffState *st = new TersoffState();
    agpu_tersoff_params &p = st->prm;
    std::memset(&p, 0, sizeof p);
    p.nel = 1;
    p.Z[0] = k.Z;
    for (int q = 0; q < 6; ++q) {
        p.A[q] = k.A;
        p.B[q] = k.B;
        p.lambda[q] = k.lambda1;
        p.mu[q] = k.lambda2;
        p.xi[q] = 1.0;
        p.omega[q] = 1.0;
        p.mubo[q] = 0.0;
        p.m[q] = 1;
        p.r1[q] = k.r1;
        p.r2[q] = k.r2;
    }
    for (int q = 0; q < 3; ++q) {
        p.beta[q] = p.n[q] = p.c[q] = p.d[q] = 1.0;
        p.h[q] = k.h;
    }
    derive(st->dev, p);
    TersoffDev &D = st->dev;
    D.fam = 1;
    for (int q = 0; q < 3; ++q) {
        D.kc1[q] = k.c1;
        D.kc2[q] = k.c2;
        D.kc3[q] = k.c3;
        D.kc4[q] = k.c4;
        D.kc5[q] = k.c5;
        D.keta[q] = k.eta;
        D.kdelta[q] = k.delta;
    }
    for (int q = 0; q < 6; ++q) {
        D.kalpha[q] = k.alpha;
        D.kbeta[q] = k.beta;
    }
    agpu_potential *P = new agpu_potential();
    P->kind = AGPU_KIND_KUMAGAI;
    P->cutoff = k.r2;
    P->impl = st;
    P->free_impl = tersoff_free;
    *pot = P;
    return AGPU_OK;
}

// Erhart_PRB_71_035211_SiC without screening (brenner_params.f90:78-113)
static agpu_brenner_params brenner_default()
{
    agpu_brenner_params b;
    std::memset(&b, 0, sizeof b);
    b.nel = 2;
    b.Z[0] = 6;
    b.Z[1] = 14;
    const double D0[3] = {6.00, 4.36, 3.24}, r0[3] = {1.4276, 1.79, 2.232}, S[3] = {2.167, 1.847, 1.842},
                 beta[3] = {2.0099, 1.6991, 1.4761}, gam[3] = {0.11233, 0.011877, 0.114354},
                 c[3] = {181.910, 273987.0, 2.00494}, d[3] = {6.28433, 180.314, 0.81472}, h[3] = {0.5556, 0.68, 0.259},
                 r1[3] = {1.85, 2.20, 2.68}, r2[3] = {2.15, 2.60, 2.96};
    for (int q = 0; q < 3; ++q) {
        b.D0[q] = D0[q];
        b.r0[q] = r0[q];
        b.S[q] = S[q];
        b.beta[q] = beta[q];
        b.gamma[q] = gam[q];
        b.c[q] = c[q];
        b.d[q] = d[q];
        b.h[q] = h[q];
        b.mu[q] = 0.0;
        b.n[q] = 1.0;
        b.m[q] = 1;
        b.r1[q] = r1[q];
        b.r2[q] = r2[q];
    }
    return b;
}

// Brenner (INIT_FUNC and the constants BIND_TO_FUNC derives,
// brenner_module.f90:27-152, 267-296): up to three elements, every
// parameter per element pair, the generic bond-order kernel with the Brenner
// forms.
extern "C" int agpu_brenner_create(agpu_potential **pot, const agpu_brenner_params *params)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    const agpu_brenner_params b = params ? *params : brenner_default();
    if (b.nel < 1 || b.nel > 3) return set_error(AGPU_ERR_ARG, "Brenner: 1 to 3 elements supported");
    const int npairs = b.nel * (b.nel + 1) / 2;
    for (int q = 0; q < npairs; ++q) {
        if (b.d[q] == 0.0)
            return set_error(AGPU_ERR_ARG, "d = 0! This leads to problems computing c**2/d**2. Please specify d != 0.");
        if (b.S[q] <= 1.0)
            return set_error(AGPU_ERR_ARG, "S <= 1! This leads to problems computing (S-1)**(-1). Please specify S > 1.");
        if (!(b.r2[q] > b.r1[q])) return set_error(AGPU_ERR_ARG, "Brenner: need r1 < r2 for every pair");
    }
    TersoffState *st = new TersoffState();
    agpu_tersoff_params &p = st->prm;
    std::memset(&p, 0, sizeof p);
    p.nel = b.nel;
    for (int e = 0; e < 3; ++e) p.Z[e] = b.Z[e];
    for (int q = 0; q < 6; ++q) {
        const bool in = q < npairs;
        p.A[q] = p.B[q] = p.lambda[q] = p.mu[q] = 1.0;
        p.xi[q] = p.omega[q] = 1.0;
        p.mubo[q] = in ? b.mu[q] : 0.0;
        p.m[q] = in ? b.m[q] : 1;
        p.r1[q] = in ? b.r1[q] : 1.0;
        p.r2[q] = in ? b.r2[q] : 2.0;
    }
    for (int q = 0; q < 3; ++q) p.beta[q] = p.n[q] = p.c[q] = p.d[q] = p.h[q] = 1.0;
    derive(st->dev, p);
    TersoffDev &D = st->dev;
    D.fam = 2;
    for (int q = 0; q < 6; ++q) {
        const bool in = q < npairs;
        const double S = in ? b.S[q] : 2.0, n = in ? b.n[q] : 1.0, beta = in ? b.beta[q] : 1.0;
        D.br0[q] = in ? b.r0[q] : 0.0;
        D.bVRf[q] = in ? b.D0[q] / (S - 1.0) : 0.0;
        D.bVAf[q] = in ? S * b.D0[q] / (S - 1.0) : 0.0;
        D.bexpR[q] = beta * std::sqrt(2.0 * S);
        D.bexpA[q] = beta * std::sqrt(2.0 / S);
        D.bgam[q] = in ? b.gamma[q] : 0.0;
        D.bc2[q] = in ? b.c[q] * b.c[q] : 0.0;
        D.bd2[q] = in ? b.d[q] * b.d[q] : 1.0;
        D.bcd2[q] = D.bd2[q] != 0.0 ? D.bc2[q] / D.bd2[q] : 0.0;
        D.bh[q] = in ? b.h[q] : 0.0;
        D.bnn[q] = n;
        D.bboe[q] = -0.5 / n;
    }
    agpu_potential *P = new agpu_potential();
    P->kind = AGPU_KIND_BRENNER;
    for (int q = 0; q < npairs; ++q)
        if (b.r2[q] > P->cutoff) P->cutoff = b.r2[q];
    P->impl = st;
    P->free_impl = tersoff_free;
    *pot = P;
    return AGPU_OK;
}

// -------------------------------------------------------------------------
// screened variants: TersoffScr, KumagaiScr, BrennerScr
// -------------------------------------------------------------------------
static agpu_screening_params scr_uniform(int npairs, const double *or1, const double *or2)
{
    agpu_screening_params sp;
    std::memset(&sp, 0, sizeof sp);
    for (int q = 0; q < npairs; ++q) {
        sp.or1[q] = sp.bor1[q] = or1[q];
        sp.or2[q] = sp.bor2[q] = or2[q];
        sp.Cmin[q] = 1.0;
        sp.Cmax[q] = 3.0;
    }
    return sp;
}

static int attach_screening(agpu_potential *P, const agpu_screening_params &sp, int kind)
{
    TersoffState *st = static_cast<TersoffState *>(P->impl);
    const int nel = st->prm.nel;
    const bool ordered = st->dev.fam == 3;
    const int npairs = ordered ? nel * nel : nel * (nel + 1) / 2;
    double mc = 0.0, x = 1.0;
    for (int q = 0; q < npairs; ++q) {
        // ordered pairs without parameters (r0 < 0 in the table) take their
        // transposed pair, as the unscreened Juslin family does
        if (ordered && !(st->dev.r2[q] > 0.0)) continue;
        if (!(sp.Cmax[q] > sp.Cmin[q])) {
            agpu_potential_destroy(P);
            return set_error(AGPU_ERR_ARG, "screening: need Cmin < Cmax for every element pair");
        }
        mc = std::max(mc, std::max(st->dev.r2[q], std::max(sp.or2[q], sp.bor2[q])));
        const double cm = sp.Cmax[q];
        x = std::max(x, cm > 2.0 ? std::sqrt(cm * cm / (4 * (cm - 1))) : 1.0);
    }
    st->scr = true;
    st->sp = sp;
    P->kind = kind;
    P->cutoff = x * mc;
    return AGPU_OK;
}

extern "C" int agpu_tersoff_scr_create(agpu_potential **pot, const agpu_tersoff_params *params,
                                       const agpu_screening_params *scr)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    agpu_tersoff_params p = params ? *params : default_params();
    agpu_screening_params sp;
    if (!params) {
        // Tersoff_PRB_39_5566_SiC with SCREENING (tersoff_params.f90:111-133)
        const double mubo[3] = {0.69103023078057590, 0.56580821386164815, 0.43569872294774004};
        const double r1[3] = {2.00, std::sqrt(2.00 * 2.50), 2.50};
        const double r2[3] = {2.00 * 1.2, std::sqrt(2.00 * 2.50) * 1.2, 2.50 * 1.2};
        for (int q = 0; q < 3; ++q) {
            p.mubo[q] = mubo[q];
            p.m[q] = 3;
            p.r1[q] = r1[q];
            p.r2[q] = r2[q];
        }
    }
    if (scr) {
        sp = *scr;
    } else {
        const double o1[3] = {2.00, std::sqrt(2.00 * 3.00), 3.00};
        const double o2[3] = {2.00 * 2.0, std::sqrt(2.00 * 3.00) * 2.0, 3.00 * 2.0};
        sp = scr_uniform(3, o1, o2);
    }
    AGPU_CHECK(agpu_tersoff_create(pot, &p));
    return attach_screening(*pot, sp, AGPU_KIND_TERSOFF_SCR);
}

extern "C" int agpu_kumagai_scr_create(agpu_potential **pot, const agpu_kumagai_params *params,
                                       const agpu_screening_params *scr)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    agpu_kumagai_params k = params ? *params : kumagai_default();
    if (!params) {
        // Kumagai_CompMaterSci_39_457_Si with SCREENING (kumagai_params.f90:119-127)
        k.r1 = 2.50;
        k.r2 = 2.50 * 1.2;
    }
    agpu_screening_params sp;
    if (scr) {
        sp = *scr;
    } else {
        const double o1[1] = {3.00}, o2[1] = {3.00 * 2.0};
        sp = scr_uniform(1, o1, o2);
    }
    AGPU_CHECK(agpu_kumagai_create(pot, &k));
    return attach_screening(*pot, sp, AGPU_KIND_KUMAGAI_SCR);
}

extern "C" int agpu_brenner_scr_create(agpu_potential **pot, const agpu_brenner_params *params,
                                       const agpu_screening_params *scr)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    agpu_brenner_params b = params ? *params : brenner_default();
    if (!params) {
        // Erhart_PRB_71_035211_SiC with SCREENING (brenner_params.f90:94-105)
        const double mu[3] = {1.0 / 1.4276, 1.0 / 1.79, 1.0 / 1.842};
        const double r1[3] = {2.00, std::sqrt(2.00 * 2.50), 2.50};
        const double r2[3] = {2.00 * 1.2, std::sqrt(2.00 * 2.50) * 1.2, 2.50 * 1.2};
        for (int q = 0; q < 3; ++q) {
            b.mu[q] = mu[q];
            b.n[q] = 1.0;
            b.m[q] = 3;
            b.r1[q] = r1[q];
            b.r2[q] = r2[q];
        }
    }
    agpu_screening_params sp;
    if (scr) {
        sp = *scr;
    } else {
        const double o1[3] = {2.00, std::sqrt(2.00 * 3.00), 3.00};
        const double o2[3] = {2.00 * 2.0, std::sqrt(2.00 * 3.00) * 2.0, 3.00 * 2.0};
        sp = scr_uniform(3, o1, o2);
    }
    AGPU_CHECK(agpu_brenner_create(pot, &b));
    return attach_screening(*pot, sp, AGPU_KIND_BRENNER_SCR);
}

// -------------------------------------------------------------------------
// Juslin (juslin.f90): ordered element pairs and per-triplet h-terms
// -------------------------------------------------------------------------
static agpu_juslin_params juslin_default()
{
    // Juslin_J_Appl_Phys_98_123520_WCH (juslin_params.f90:74-112, without
    // SCREENING)
    agpu_juslin_params j;
    std::memset(&j, 0, sizeof j);
    j.nel = 3;
    j.Z[0] = 74;
    j.Z[1] = 6;
    j.Z[2] = 1;
    const double D0[9] = {5.41861, 6.64, 2.748, 0.0, 6.0, 3.6422, 0.0, 3.642, 4.7509};
    const double r0[9] = {2.34095, 1.90547, 1.727, -1.0, 1.39, 1.1199, -1.0, 1.1199, 0.74144};
    const double S[9] = {1.92708, 2.96149, 1.2489, 0.0, 1.22, 1.69077, 0.0, 1.69077, 2.3432};
    const double beta[9] = {1.38528, 1.80370, 1.52328, 0.0, 2.1, 1.9583, 0.0, 1.9583, 1.9436};
    const double gam[9] = {0.00188227, 0.072855, 0.0054, 0.0, 0.00020813, 0.00020813, 0.0, 12.33, 12.33};
    const double c[9] = {2.14969, 1.10304, 1.788, 0.0, 330.0, 330.0, 0.0, 0.0, 0.0};
    const double d[9] = {0.17126, 0.33018, 0.8255, 0.0, 3.5, 3.5, 0.0, 1.0, 1.0};
    const double h[9] = {-0.27780, 0.75107, 0.38912, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0};
    const double n[9] = {1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0};
    const double r1[9] = {3.20, 2.60, 1.95, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10};
    const double r2[9] = {3.80, 3.00, 2.35, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70};
    for (int q = 0; q < 9; ++q) {
        j.D0[q] = D0[q];
        j.r0[q] = r0[q];
        j.S[q] = S[q];
        j.beta[q] = beta[q];
        j.gamma[q] = gam[q];
        j.c[q] = c[q];
        j.d[q] = d[q];
        j.h[q] = h[q];
        j.n[q] = n[q];
        j.r1[q] = r1[q];
        j.r2[q] = r2[q];
    }
    for (int t = 0; t < 27; ++t) {
        j.alpha[t] = 0.0;
        j.omega[t] = 1.0;
        j.m[t] = 1;
    }
    j.alpha[0] = j.alpha[3] = j.alpha[6] = 0.45876;
    const int a4[] = {14, 16, 17, 22, 23, 25, 26};
    for (int t : a4) j.alpha[t] = 4.0;
    j.omega[16] = 2.94586;
    j.omega[17] = 4.54415;
    j.omega[22] = 0.33946;
    j.omega[23] = 0.22006;
    return j;
}

extern "C" int agpu_juslin_create(agpu_potential **pot, const agpu_juslin_params *params)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    agpu_juslin_params j = params ? *params : juslin_default();
    const int nel = j.nel;
    if (nel < 1 || nel > 3) return set_error(AGPU_ERR_ARG, "Juslin: 1 to 3 elements supported");
    // entries with r0 < 0 take the reversed pair (juslin_module.f90:273-305)
    for (int a = 0; a < nel; ++a)
        for (int b = 0; b < nel; ++b) {
            const int q = b + a * nel, t = a + b * nel;
            if (j.r0[q] < 0.0) {
                j.D0[q] = j.D0[t];
                j.r0[q] = j.r0[t];
                j.S[q] = j.S[t];
                j.beta[q] = j.beta[t];
                j.gamma[q] = j.gamma[t];
                j.c[q] = j.c[t];
                j.d[q] = j.d[t];
                j.h[q] = j.h[t];
                j.n[q] = j.n[t];
                j.r1[q] = j.r1[t];
                j.r2[q] = j.r2[t];
            }
        }
    for (int q = 0; q < nel * nel; ++q) {
        if (j.d[q] == 0.0)
            return set_error(AGPU_ERR_ARG, "d = 0! This leads to problems computing c**2/d**2. Please specify d != 0.");
        if (j.S[q] <= 1.0)
            return set_error(AGPU_ERR_ARG, "S <= 1! This leads to problems computing (S-1)**(-1). Please specify S > 1.");
        if (!(j.r2[q] > j.r1[q])) return set_error(AGPU_ERR_ARG, "Juslin: need r1 < r2 for every pair");
    }
    TersoffState *st = new TersoffState();
    agpu_tersoff_params &p = st->prm;
    std::memset(&p, 0, sizeof p);
    p.nel = nel;
    for (int e = 0; e < 3; ++e) p.Z[e] = j.Z[e];
    for (int q = 0; q < 6; ++q) {
        p.A[q] = p.B[q] = p.lambda[q] = p.mu[q] = p.xi[q] = p.omega[q] = 1.0;
        p.m[q] = 1;
        p.r1[q] = 1.0;
        p.r2[q] = 2.0;
    }
    for (int q = 0; q < 3; ++q) p.beta[q] = p.n[q] = p.c[q] = p.d[q] = p.h[q] = 1.0;
    derive(st->dev, p);
    TersoffDev &D = st->dev;
    D.fam = 3;
    for (int q = 0; q < 9; ++q) {
        const bool in = q < nel * nel;
        const double S = in ? j.S[q] : 2.0, nn = in ? j.n[q] : 1.0, beta = in ? j.beta[q] : 1.0;
        D.r1[q] = in ? j.r1[q] : 1.0;
        D.r2[q] = in ? j.r2[q] : 2.0;
        D.fac[q] = M_PI / (D.r2[q] - D.r1[q]);
        D.br0[q] = in ? j.r0[q] : 0.0;
        D.bVRf[q] = in ? j.D0[q] / (S - 1.0) : 0.0;
        D.bVAf[q] = in ? S * j.D0[q] / (S - 1.0) : 0.0;
        D.bexpR[q] = beta * std::sqrt(2.0 * S);
        D.bexpA[q] = beta * std::sqrt(2.0 / S);
        D.bgam[q] = in ? j.gamma[q] : 0.0;
        D.bc2[q] = in ? j.c[q] * j.c[q] : 0.0;
        D.bd2[q] = in ? j.d[q] * j.d[q] : 1.0;
        D.bcd2[q] = D.bd2[q] != 0.0 ? D.bc2[q] / D.bd2[q] : 0.0;
        D.bh[q] = in ? j.h[q] : 0.0;
        D.bnn[q] = nn;
        D.bboe[q] = -0.5 / nn;
    }
    for (int t = 0; t < 27; ++t) {
        const bool in = t < nel * nel * nel;
        D.jal[t] = in ? j.alpha[t] : 0.0;
        D.jom[t] = in ? j.omega[t] : 1.0;
        D.jm[t] = in ? j.m[t] : 1;
    }
    agpu_potential *P = new agpu_potential();
    P->kind = AGPU_KIND_JUSLIN;
    for (int q = 0; q < nel * nel; ++q)
        if (D.r2[q] > P->cutoff) P->cutoff = D.r2[q];
    P->impl = st;
    P->free_impl = tersoff_free;
    *pot = P;
    return AGPU_OK;
}

// JuslinScr (juslin_scr.f90): the Juslin family with the screening of the
// shared kernel (SCREENING with EXP_BOP).  Screening parameters are per
// ordered pair; entries whose r0 < 0 take those of the transposed pair, as
// the other parameters do (juslin_module.f90:273-305).  Defaults: the
// screened W-C-H set of juslin_params.f90 (W-H cutoffs 2.68-2.96 A; outer and
// bond-order cutoffs equal to the inner ones; Cmin 1, Cmax 3).
extern "C" int agpu_juslin_scr_create(agpu_potential **pot, const agpu_juslin_params *params,
                                      const agpu_screening_params *scr)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pot is NULL");
    agpu_juslin_params j = params ? *params : juslin_default();
    if (!params) {
        j.r1[2] = 2.68;
        j.r2[2] = 2.96;
    }
    const int nel = j.nel;
    if (nel < 1 || nel > 3) return set_error(AGPU_ERR_ARG, "Juslin: 1 to 3 elements supported");
    agpu_screening_params sp;
    if (scr) {
        sp = *scr;
    } else {
        std::memset(&sp, 0, sizeof sp);
        for (int q = 0; q < nel * nel; ++q) {
            if (j.r0[q] < 0.0) continue;
            sp.or1[q] = sp.bor1[q] = j.r1[q];
            sp.or2[q] = sp.bor2[q] = j.r2[q];
            sp.Cmin[q] = 1.0;
            sp.Cmax[q] = 3.0;
        }
    }
    for (int a = 0; a < nel; ++a)
        for (int b = 0; b < nel; ++b) {
            const int q = b + a * nel, t = a + b * nel;
            if (j.r0[q] < 0.0) {
                sp.or1[q] = sp.or1[t];
                sp.or2[q] = sp.or2[t];
                sp.bor1[q] = sp.bor1[t];
                sp.bor2[q] = sp.bor2[t];
                sp.Cmin[q] = sp.Cmin[t];
                sp.Cmax[q] = sp.Cmax[t];
            }
        }
    AGPU_CHECK(agpu_juslin_create(pot, &j));
    return attach_screening(*pot, sp, AGPU_KIND_JUSLIN_SCR);
}
