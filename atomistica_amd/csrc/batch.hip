// Batched single-element Tersoff for parameter fitting: thousands of small
// periodic structures of uneven size evaluated together over one
// concatenated neighbour list, with a hand-written backward pass for the 11
// parameters theta = (A, B, lambda, mu, beta, n, c, d, h, R, S).
//
// The functional form is the reference's Tersoff bond-order kernel for one
// element with xi = omega = 1 and no h-term (mubo = 0)
// (bop_kernel.f90:1075-1529, tersoff_func.f90:32-211):
//   E = 1/2 sum_{i != j} fc(r_ij) [A e^{-lambda r_ij} - b_ij B e^{-mu r_ij}]
//   b_ij = (1 + (beta zeta_ij)^n)^{-1/(2n)},  zeta_ij = sum_k fc(r_ik) g(cos)
//   g = 1 + c^2/d^2 - c^2/(d^2 + (h - cos)^2)
//   fc = trig_off on [R, S] (cutoff.f90:171-196).
//
// Kernels:
//   k_bnl_count / k_bnl_fill  thread per atom: every periodic image of every
//       atom of the same structure with |r_i - r_j + A n|^2 < S^2 (the exact
//       image range per axis from the reciprocal vectors), counted, scanned,
//       then written in (j ascending, n lexicographic) order.
//   k_tb_forward<R>  thread per atom i: every bond ij, zeta_ij over k, b_ij,
//       energy of the directed bond into e_atom[i], forces on i, j, k
//       through 64-bit fixed-point accumulators (order independent, so
//       bitwise reproducible).  R = double, or float for FP32 pair/triplet
//       arithmetic with FP64 displacements and FP64/fixed-point accumulation.
//   k_tb_backward  thread per atom i in FP64 dual numbers: dE_i/dtheta by
//       differentiating the bond energy by hand, with every position carrying
//       the tangent w_i = dL/dF_i.  The tangent part is then
//       sum_i w_i . d(dE/dtheta)/dr_i = -sum_i w_i . dF_i/dtheta, so one pass
//       gives sum_s gE_s dE_s/dtheta + sum_i gF_i . dF_i/dtheta with nothing
//       stored per triplet (zeta and g are recomputed).
// Reductions over atoms and structures are fixed-order.
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "profile.hpp"

struct agpu_batch {
    int nstruct = 0;
    int ntot = 0;
    std::vector<int> h_off;        // atom offsets (nstruct + 1)
    std::vector<double> h_cell;    // 9 per structure: cell[9*s + 3*k + x]
    agpu::DevBuf off, sof, cell, rec, r, counts, offsets, nbr, scan_tmp, facc, eatom,
        partial, gacc, flags;
    agpu::HostStats stats;
    double nl_cut = -1.0;          // cutoff of the current list
    uint64_t pos_rev = 0, nl_rev = ~0ull;
    int64_t total = 0;
};

namespace agpu {

struct TbParams {
    double A, B, lam, mu, beta, n, c, d, h, R, S;
    // derived
    double c2, d2, c2d2, S2, fac, e;   // e = -1/(2n)
};

static TbParams tb_params(const double *t)
{
    TbParams P;
    P.A = t[0]; P.B = t[1]; P.lam = t[2]; P.mu = t[3]; P.beta = t[4]; P.n = t[5];
    P.c = t[6]; P.d = t[7]; P.h = t[8]; P.R = t[9]; P.S = t[10];
    P.c2 = P.c * P.c;
    P.d2 = P.d * P.d;
    P.c2d2 = P.c2 / P.d2;
    P.S2 = P.S * P.S;
    P.fac = M_PI / (P.S - P.R);
    P.e = -0.5 / P.n;
    return P;
}

// ---------------------------------------------------------------------------
// batched neighbour list
// ---------------------------------------------------------------------------
struct StructGeom {
    const int *sof;        // structure of each atom
    const int *off;        // atom offsets per structure
    const double *cell;    // 9 per structure, cell[3k+x] = a_k[x]
    const double *rec;     // 9 per structure, rec[3k+x] = b_k[x] with a_k . b_l = delta_kl
};

template <bool FILL>
__global__ void __launch_bounds__(256)
k_bnl(int ntot, const double *__restrict__ r, StructGeom G, double rc, const int *__restrict__ offsets,
      int *__restrict__ counts, int2 *__restrict__ nbr, int *__restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    const int s = G.sof[i];
    const int a0 = G.off[s], a1 = G.off[s + 1];
    CellArg C;
    double b[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        C.A[q] = G.cell[9 * (size_t)s + q];
        b[q] = G.rec[9 * (size_t)s + q];
    }
    const double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    double t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        t[k] = rc * sqrt(b[3 * k] * b[3 * k] + b[3 * k + 1] * b[3 * k + 1] + b[3 * k + 2] * b[3 * k + 2]);
    const double rc2 = rc * rc;
    int cnt = 0, pos = FILL ? offsets[i] : 0;
    int bad = 0;
    for (int j = a0; j < a1; ++j) {
        const double rj[3] = {r[3 * j], r[3 * j + 1], r[3 * j + 2]};
        // fractional difference s_i - s_j along each reciprocal vector
        int lo[3], hi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double dsk = b[3 * k] * (ri[0] - rj[0]) + b[3 * k + 1] * (ri[1] - rj[1]) +
                               b[3 * k + 2] * (ri[2] - rj[2]);
            lo[k] = (int)ceil(-dsk - t[k]) - 1;
            hi[k] = (int)floor(-dsk + t[k]) + 1;
        }
        for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
            for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
                for (int n2 = lo[2]; n2 <= hi[2]; ++n2) {
                    if (j == i && n0 == 0 && n1 == 0 && n2 == 0) continue;
                    double d[3];
                    delta_r(C, ri, rj, n0, n1, n2, d);
                    if (dot3_nc(d) < rc2) {
                        if (!(dc_pack_ok(n0) && dc_pack_ok(n1) && dc_pack_ok(n2))) {
                            bad = 1;
                            continue;
                        }
                        if (FILL) nbr[pos + cnt] = make_int2(j, dc_pack(n0, n1, n2));
                        ++cnt;
                    }
                }
    }
    if (!FILL) counts[i] = cnt;
    if (bad) atomicOr(flags, DERR_IMAGE_RANGE);
}

__global__ void k_bnl_tail(int n, const int *__restrict__ counts, int *__restrict__ offsets)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) offsets[n] = n > 0 ? offsets[n - 1] + counts[n - 1] : 0;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
constexpr double TB_FX = 1099511627776.0;      // 2^40
constexpr double TB_FX_INV = 1.0 / 1099511627776.0;

__device__ __forceinline__ void tb_add(unsigned long long *a, double v)
{
    atomicAdd(a, (unsigned long long)__double2ll_rn(v * TB_FX));
}

template <typename R>
__device__ __forceinline__ void tb_fc(const TbParams &P, R r, R &fc, R &dfc)
{
    if (r <= (R)P.R) { fc = 1; dfc = 0; }
    else if (r >= (R)P.S) { fc = 0; dfc = 0; }
    else {
        const R x = (R)P.fac * (r - (R)P.R);
        fc = (R)0.5 * ((R)1 + cos(x));
        dfc = (R)-0.5 * (R)P.fac * sin(x);
    }
}

// bond i -> (j, image): unit vector (from i towards j) and length
template <typename R>
__device__ __forceinline__ void tb_bond(const CellArg &C, const double *ri, const double *__restrict__ r,
                                        int2 e, R *n, R &len)
{
    int a, b, c;
    dc_unpack(e.y, a, b, c);
    const double rj[3] = {r[3 * e.x], r[3 * e.x + 1], r[3 * e.x + 2]};
    double d[3];
    delta_r(C, ri, rj, a, b, c, d);
    const double l = sqrt(dot3_nc(d));
    len = (R)l;
    // r_ij = r_j - r_i - A.dc = -d
    n[0] = (R)(-d[0] / l);
    n[1] = (R)(-d[1] / l);
    n[2] = (R)(-d[2] / l);
}

template <typename R>
__global__ void __launch_bounds__(256)
k_tb_forward(int ntot, const double *__restrict__ r, const int *__restrict__ sof,
             const double *__restrict__ cells, const int *__restrict__ offsets, const int2 *__restrict__ nbr,
             TbParams P, double *__restrict__ eatom, unsigned long long *__restrict__ facc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    CellArg C;
    const int s = sof[i];
#pragma unroll
    for (int q = 0; q < 9; ++q) C.A[q] = cells[9 * (size_t)s + q];
    const double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    const int beg = offsets[i], end = offsets[i + 1];
    double e = 0.0, fi[3] = {0, 0, 0};
    for (int jb = beg; jb < end; ++jb) {
        const int2 ej = nbr[jb];
        R nij[3], rij;
        tb_bond<R>(C, ri, r, ej, nij, rij);
        if (!(rij < (R)P.S)) continue;
        R fcij, dfcij;
        tb_fc<R>(P, rij, fcij, dfcij);
        const R exa = exp(-(R)P.mu * rij), exr = exp(-(R)P.lam * rij);
        const R VA = -(R)P.B * exa, dVA = (R)P.B * (R)P.mu * exa;
        const R VR = (R)P.A * exr, dVR = -(R)P.A * (R)P.lam * exr;
        // zeta
        R z = 0;
        for (int kb = beg; kb < end; ++kb) {
            if (kb == jb) continue;
            R nik[3], rik;
            tb_bond<R>(C, ri, r, nbr[kb], nik, rik);
            if (!(rik < (R)P.S)) continue;
            R fcik, dfcik;
            tb_fc<R>(P, rik, fcik, dfcik);
            const R cs = nij[0] * nik[0] + nij[1] * nik[1] + nij[2] * nik[2];
            const R t = (R)P.h - cs;
            const R g = (R)1 + (R)P.c2 * t * t / ((R)P.d2 * ((R)P.d2 + t * t));
            z += fcik * g;
        }
        R b = 1, dbdz = 0;
        if (z > (R)0) {
            const R u = pow((R)P.beta * z, (R)P.n);
            b = pow((R)1 + u, (R)P.e);
            dbdz = (R)P.e * (R)P.n * b * u / (z * ((R)1 + u));
        }
        // E_ij = 1/2 fc (VR + b VA)
        e += 0.5 * (double)(fcij * (VR + b * VA));
        const R dEdr = (R)0.5 * (dfcij * (VR + b * VA) + fcij * (dVR + b * dVA));
        const R pre = (R)0.5 * fcij * VA * dbdz;     // dE/dzeta
        // radial part: force on i is +dEdr * n_ij, on j the opposite
        double fj[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            fi[x] += (double)(dEdr * nij[x]);
            fj[x] = -(double)(dEdr * nij[x]);
        }
        if (pre != (R)0) {
            for (int kb = beg; kb < end; ++kb) {
                if (kb == jb) continue;
                const int2 ek = nbr[kb];
                R nik[3], rik;
                tb_bond<R>(C, ri, r, ek, nik, rik);
                if (!(rik < (R)P.S)) continue;
                R fcik, dfcik;
                tb_fc<R>(P, rik, fcik, dfcik);
                const R cs = nij[0] * nik[0] + nij[1] * nik[1] + nij[2] * nik[2];
                const R t = (R)P.h - cs;
                const R den = (R)P.d2 + t * t;
                const R g = (R)1 + (R)P.c2 * t * t / ((R)P.d2 * den);
                // dg/dcos = -2 c^2 t / den^2
                const R dg = (R)-2 * (R)P.c2 * t / (den * den);
                // d cos / d r_j = (n_ik - cos n_ij) / r_ij, d cos / d r_k = (n_ij - cos n_ik) / r_ik
                double fk[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const R dcj = (nik[x] - cs * nij[x]) / rij;
                    const R dck = (nij[x] - cs * nik[x]) / rik;
                    const R gj = fcik * dg * dcj;
                    const R gk = fcik * dg * dck + dfcik * g * nik[x];
                    // F = -pre * d zeta / dr
                    fj[x] += -(double)(pre * gj);
                    fk[x] = -(double)(pre * gk);
                    fi[x] += (double)(pre * (gj + gk));
                }
                tb_add(facc + 3 * (size_t)ek.x + 0, fk[0]);
                tb_add(facc + 3 * (size_t)ek.x + 1, fk[1]);
                tb_add(facc + 3 * (size_t)ek.x + 2, fk[2]);
            }
        }
        tb_add(facc + 3 * (size_t)ej.x + 0, fj[0]);
        tb_add(facc + 3 * (size_t)ej.x + 1, fj[1]);
        tb_add(facc + 3 * (size_t)ej.x + 2, fj[2]);
    }
    eatom[i] = e;
    tb_add(facc + 3 * (size_t)i + 0, fi[0]);
    tb_add(facc + 3 * (size_t)i + 1, fi[1]);
    tb_add(facc + 3 * (size_t)i + 2, fi[2]);
}

__global__ void k_tb_finish(int ntot, unsigned long long *__restrict__ facc, double *__restrict__ f)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * ntot) return;
    f[q] += (double)(long long)facc[q] * TB_FX_INV;
    facc[q] = 0ull;
}

__global__ void k_tb_struct_sum(int nstruct, const int *__restrict__ off, const double *__restrict__ eatom,
                                double *__restrict__ es)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstruct) return;
    double e = 0.0;
    for (int a = off[s]; a < off[s + 1]; ++a) e += eatom[a];
    es[s] = e;
}

__global__ void k_tb_zero(unsigned long long *__restrict__ a, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        a[q] = 0ull;
}

// ---------------------------------------------------------------------------
// backward: dual numbers (value, tangent along the force cotangent)
// ---------------------------------------------------------------------------
struct Dl {
    double v, t;
};
__device__ __forceinline__ Dl dl(double v, double t = 0.0) { return Dl{v, t}; }
__device__ __forceinline__ Dl operator+(Dl a, Dl b) { return Dl{a.v + b.v, a.t + b.t}; }
__device__ __forceinline__ Dl operator-(Dl a, Dl b) { return Dl{a.v - b.v, a.t - b.t}; }
__device__ __forceinline__ Dl operator-(Dl a) { return Dl{-a.v, -a.t}; }
__device__ __forceinline__ Dl operator*(Dl a, Dl b) { return Dl{a.v * b.v, a.t * b.v + a.v * b.t}; }
__device__ __forceinline__ Dl operator*(double s, Dl a) { return Dl{s * a.v, s * a.t}; }
__device__ __forceinline__ Dl operator/(Dl a, Dl b)
{
    const double q = a.v / b.v;
    return Dl{q, (a.t - q * b.t) / b.v};
}
__device__ __forceinline__ Dl operator*(Dl a, double s) { return Dl{s * a.v, s * a.t}; }
__device__ __forceinline__ Dl operator/(double s, Dl a)
{
    const double q = s / a.v;
    return Dl{q, -q * a.t / a.v};
}
__device__ __forceinline__ Dl operator+(Dl a, double s) { return Dl{a.v + s, a.t}; }
__device__ __forceinline__ Dl operator+(double s, Dl a) { return Dl{a.v + s, a.t}; }
__device__ __forceinline__ Dl operator-(double s, Dl a) { return Dl{s - a.v, -a.t}; }
__device__ __forceinline__ Dl operator-(Dl a, double s) { return Dl{a.v - s, a.t}; }
__device__ __forceinline__ Dl dexp(Dl a)
{
    const double e = exp(a.v);
    return Dl{e, e * a.t};
}
__device__ __forceinline__ Dl dlog(Dl a) { return Dl{log(a.v), a.t / a.v}; }
__device__ __forceinline__ Dl dsqrt(Dl a)
{
    const double s = sqrt(a.v);
    return Dl{s, 0.5 * a.t / s};
}
__device__ __forceinline__ Dl dpow(Dl a, double p)
{
    const double v = pow(a.v, p);
    return Dl{v, p * pow(a.v, p - 1.0) * a.t};
}
__device__ __forceinline__ void dsincos(Dl a, Dl &s, Dl &c)
{
    double sv, cv;
    sincos(a.v, &sv, &cv);
    s = Dl{sv, cv * a.t};
    c = Dl{cv, -sv * a.t};
}

// bond vector r_ij = r_j - r_i - A.dc with tangent w_j - w_i
__device__ __forceinline__ void tb_dbond(const CellArg &C, const double *ri, const double *wi,
                                         const double *__restrict__ r, const double *__restrict__ w, int2 e,
                                         Dl *v, Dl &len)
{
    int a, b, c;
    dc_unpack(e.y, a, b, c);
    const double rj[3] = {r[3 * e.x], r[3 * e.x + 1], r[3 * e.x + 2]};
    double d[3];
    delta_r(C, ri, rj, a, b, c, d);
#pragma unroll
    for (int x = 0; x < 3; ++x) v[x] = Dl{-d[x], w ? w[3 * e.x + x] - wi[x] : 0.0};
    len = dsqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// fc and its derivatives with respect to R and S (value and tangent)
__device__ __forceinline__ void tb_dfc(const TbParams &P, Dl r, Dl &fc, Dl &fcR, Dl &fcS)
{
    if (r.v <= P.R) { fc = dl(1.0); fcR = dl(0.0); fcS = dl(0.0); }
    else if (r.v >= P.S) { fc = dl(0.0); fcR = dl(0.0); fcS = dl(0.0); }
    else {
        const double L = P.S - P.R;
        Dl sx, cx;
        dsincos(P.fac * (r - P.R), sx, cx);
        fc = 0.5 * (1.0 + cx);
        // x = pi (r - R) / (S - R): dx/dR = -pi (S - r) / L^2, dx/dS = -pi (r - R) / L^2
        fcR = (0.5 * M_PI / (L * L)) * (sx * (P.S - r));
        fcS = (0.5 * M_PI / (L * L)) * (sx * (r - P.R));
    }
}

constexpr int TB_NP = 11;

__global__ void __launch_bounds__(256)
k_tb_backward(int ntot, const double *__restrict__ r, const double *__restrict__ w,
              const double *__restrict__ gE, const int *__restrict__ sof, const double *__restrict__ cells,
              const int *__restrict__ offsets, const int2 *__restrict__ nbr, TbParams P,
              double *__restrict__ partial)
{
    __shared__ double s_red[4][TB_NP];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[TB_NP];
#pragma unroll
    for (int q = 0; q < TB_NP; ++q) acc[q] = 0.0;
    if (i < ntot) {
        CellArg C;
        const int s = sof[i];
#pragma unroll
        for (int q = 0; q < 9; ++q) C.A[q] = cells[9 * (size_t)s + q];
        const double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
        const double wi[3] = {w ? w[3 * i] : 0.0, w ? w[3 * i + 1] : 0.0, w ? w[3 * i + 2] : 0.0};
        const double ge = gE ? gE[s] : 0.0;
        const int beg = offsets[i], end = offsets[i + 1];
        Dl g[TB_NP];
#pragma unroll
        for (int q = 0; q < TB_NP; ++q) g[q] = dl(0.0);
        for (int jb = beg; jb < end; ++jb) {
            Dl vij[3], rij;
            tb_dbond(C, ri, wi, r, w, nbr[jb], vij, rij);
            if (!(rij.v < P.S)) continue;
            Dl fc, fcR, fcS;
            tb_dfc(P, rij, fc, fcR, fcS);
            const Dl exa = dexp(-P.mu * rij), exr = dexp(-P.lam * rij);
            const Dl VA = -P.B * exa, VR = P.A * exr;
            Dl z = dl(0.0), zc = dl(0.0), zd = dl(0.0), zh = dl(0.0), zR = dl(0.0), zS = dl(0.0);
            for (int kb = beg; kb < end; ++kb) {
                if (kb == jb) continue;
                Dl vik[3], rik;
                tb_dbond(C, ri, wi, r, w, nbr[kb], vik, rik);
                if (!(rik.v < P.S)) continue;
                Dl fck, fckR, fckS;
                tb_dfc(P, rik, fck, fckR, fckS);
                const Dl cs = (vij[0] * vik[0] + vij[1] * vik[1] + vij[2] * vik[2]) / (rij * rik);
                const Dl t = P.h - cs;
                const Dl den = P.d2 + t * t;
                const Dl gq = 1.0 + P.c2d2 - P.c2 / den;
                // dg/dc = 2c/d^2 - 2c/den; dg/dd = -2c^2/d^3 + 2c^2 d/den^2; dg/dh = 2c^2 t/den^2
                const Dl gc = (2.0 * P.c / P.d2) - (2.0 * P.c) / den;
                const Dl gd = (-2.0 * P.c2 / (P.d2 * P.d)) + (2.0 * P.c2 * P.d) / (den * den);
                const Dl gh = (2.0 * P.c2) * (t / (den * den));
                z = z + fck * gq;
                zc = zc + fck * gc;
                zd = zd + fck * gd;
                zh = zh + fck * gh;
                zR = zR + fckR * gq;
                zS = zS + fckS * gq;
            }
            Dl b = dl(1.0), bb = dl(0.0), bn = dl(0.0), bz = dl(0.0);
            if (z.v > 0.0) {
                const Dl bz_ = P.beta * z;
                const Dl u = dpow(bz_, P.n);
                const Dl opu = 1.0 + u;
                b = dpow(opu, P.e);
                // d ln b: beta -> -u/(2 beta (1+u)), zeta -> -u/(2 zeta (1+u)),
                // n -> ln(1+u)/(2 n^2) - u ln(beta zeta)/(2 n (1+u))
                bb = b * (-1.0 / (2.0 * P.beta)) * (u / opu);
                bz = b * (-0.5 * u / (z * opu));
                bn = b * ((0.5 / (P.n * P.n)) * dlog(opu) - (0.5 / P.n) * (u * dlog(bz_) / opu));
            }
            const Dl hf = 0.5 * fc;
            const Dl pair = VR + b * VA;
            const Dl pz = hf * VA * bz;        // dE/dzeta
            g[0] = g[0] + hf * exr;                           // A
            g[1] = g[1] - hf * b * exa;                       // B
            g[2] = g[2] - hf * rij * VR;                      // lambda
            g[3] = g[3] - hf * b * rij * VA;                  // mu
            g[4] = g[4] + hf * VA * bb;                       // beta
            g[5] = g[5] + hf * VA * bn;                       // n
            g[6] = g[6] + pz * zc;                            // c
            g[7] = g[7] + pz * zd;                            // d
            g[8] = g[8] + pz * zh;                            // h
            g[9] = g[9] + 0.5 * fcR * pair + pz * zR;         // R
            g[10] = g[10] + 0.5 * fcS * pair + pz * zS;       // S
        }
        // value weighted by dL/dE_s; tangent = -sum w . dF/dtheta
#pragma unroll
        for (int q = 0; q < TB_NP; ++q) acc[q] = ge * g[q].v - g[q].t;
    }
#pragma unroll
    for (int q = 0; q < TB_NP; ++q) acc[q] = group_sum<64>(acc[q]);
    const int tid = threadIdx.x, wave = tid / 64;
    if ((tid & 63) == 0)
        for (int q = 0; q < TB_NP; ++q) s_red[wave][q] = acc[q];
    __syncthreads();
    if (tid < TB_NP)
        partial[(size_t)blockIdx.x * TB_NP + tid] =
            ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
static int batch_nl(agpu_batch *b, double rc)
{
    if (b->nl_rev == b->pos_rev && b->nl_cut == rc) return AGPU_OK;
    hipStream_t s = stream();
    const int n = b->ntot;
    AGPU_CHECK(b->counts.reserve(sizeof(int) * (n + 1)));
    AGPU_CHECK(b->offsets.reserve(sizeof(int) * (n + 1)));
    AGPU_CHECK(b->flags.reserve(sizeof(int) * 4));
    AGPU_CHECK(b->stats.init());
    int *flags = b->flags.as<int>();
    AGPU_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 4, s));
    StructGeom G{b->sof.as<int>(), b->off.as<int>(), b->cell.as<double>(), b->rec.as<double>()};
    const int nb = (n + 255) / 256;
    AGPU_LAUNCH("batch_nl_count", k_bnl<false>, dim3(nb), dim3(256), 0, s, n, b->r.as<double>(), G, rc,
                nullptr, b->counts.as<int>(), nullptr, flags);
    size_t tmp = 0;
    AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, b->counts.as<int>(), b->offsets.as<int>(), 0, (size_t)n,
                                     rocprim::plus<int>(), s));
    AGPU_CHECK(b->scan_tmp.reserve(tmp + 16));
    AGPU_HIP(rocprim::exclusive_scan(b->scan_tmp.ptr, tmp, b->counts.as<int>(), b->offsets.as<int>(), 0,
                                     (size_t)n, rocprim::plus<int>(), s));
    AGPU_LAUNCH("batch_nl_tail", k_bnl_tail, dim3(1), dim3(64), 0, s, n, b->counts.as<int>(),
                b->offsets.as<int>());
    int total = 0;
    AGPU_HIP(hipMemcpyAsync(&total, b->offsets.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (total < 0) return set_error(AGPU_ERR_NL_OVERFLOW, "batch: neighbour list exceeds 2^31 entries");
    AGPU_CHECK(b->nbr.reserve(sizeof(int2) * ((size_t)total + 1)));
    AGPU_LAUNCH("batch_nl_fill", k_bnl<true>, dim3(nb), dim3(256), 0, s, n, b->r.as<double>(), G, rc,
                b->offsets.as<int>(), b->counts.as<int>(), b->nbr.as<int2>(), flags);
    AGPU_HIP(hipMemcpyAsync(b->stats.h, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (b->stats.h[0] & DERR_IMAGE_RANGE)
        return set_error(AGPU_ERR_DOMAIN, "batch: image shift beyond +-511 cells");
    b->total = total;
    b->nl_cut = rc;
    b->nl_rev = b->pos_rev;
    return AGPU_OK;
}

static int check_theta(const double *theta)
{
    if (!theta) return set_error(AGPU_ERR_ARG, "batch: theta is NULL");
    const TbParams P = tb_params(theta);
    if (!(P.S > P.R && P.R > 0.0 && P.n > 0.0 && P.beta > 0.0 && P.d != 0.0))
        return set_error(AGPU_ERR_ARG, "batch: need 0 < R < S, n > 0, beta > 0, d != 0");
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_batch_create(agpu_batch **b)
{
    if (!b) return set_error(AGPU_ERR_ARG, "b is NULL");
    *b = new agpu_batch();
    return AGPU_OK;
}

extern "C" int agpu_batch_destroy(agpu_batch *b)
{
    delete b;
    return AGPU_OK;
}

extern "C" int agpu_batch_set_structures(agpu_batch *b, int nstruct, const int *atom_offsets,
                                         const double *cells)
{
    if (!b || nstruct < 0 || (nstruct > 0 && (!atom_offsets || !cells)))
        return set_error(AGPU_ERR_ARG, "batch_set_structures: bad arguments");
    if (nstruct > 0 && atom_offsets[0] != 0) return set_error(AGPU_ERR_ARG, "batch: atom_offsets[0] must be 0");
    for (int s = 0; s < nstruct; ++s)
        if (atom_offsets[s + 1] < atom_offsets[s])
            return set_error(AGPU_ERR_ARG, "batch: atom_offsets must be non-decreasing");
    const int ntot = nstruct > 0 ? atom_offsets[nstruct] : 0;
    std::vector<double> rec(9 * (size_t)nstruct);
    std::vector<int> sof(ntot);
    for (int s = 0; s < nstruct; ++s) {
        // reciprocal vectors b_k (rows of A^-T): a_k . b_l = delta_kl
        double B[9];
        if (compute_inverse(cells + 9 * (size_t)s, B) != AGPU_OK)
            return set_error(AGPU_ERR_ARG, "batch: singular cell");
        // B[3*k + x] = Bbox(x, k) = (A^-1)(x, k) with Abox(x, k) = cell[3k + x];
        // b_k[x] = (A^-1)(k, x) = B[3*x + k]
        for (int k = 0; k < 3; ++k)
            for (int x = 0; x < 3; ++x) rec[9 * (size_t)s + 3 * k + x] = B[3 * x + k];
        for (int a = atom_offsets[s]; a < atom_offsets[s + 1]; ++a) sof[a] = s;
    }
    b->nstruct = nstruct;
    b->ntot = ntot;
    b->h_off.assign(atom_offsets, atom_offsets + nstruct + 1);
    b->h_cell.assign(cells, cells + 9 * (size_t)nstruct);
    AGPU_CHECK(b->off.reserve(sizeof(int) * (nstruct + 1)));
    AGPU_CHECK(b->sof.reserve(sizeof(int) * (ntot + 1)));
    AGPU_CHECK(b->cell.reserve(sizeof(double) * 9 * (nstruct + 1)));
    AGPU_CHECK(b->rec.reserve(sizeof(double) * 9 * (nstruct + 1)));
    AGPU_CHECK(b->r.reserve(sizeof(double) * 3 * (ntot + 1)));
    hipStream_t st = stream();
    AGPU_HIP(hipMemcpyAsync(b->off.ptr, atom_offsets, sizeof(int) * (nstruct + 1), hipMemcpyHostToDevice, st));
    if (ntot) AGPU_HIP(hipMemcpyAsync(b->sof.ptr, sof.data(), sizeof(int) * ntot, hipMemcpyHostToDevice, st));
    if (nstruct) {
        AGPU_HIP(hipMemcpyAsync(b->cell.ptr, cells, sizeof(double) * 9 * nstruct, hipMemcpyHostToDevice, st));
        AGPU_HIP(hipMemcpyAsync(b->rec.ptr, rec.data(), sizeof(double) * 9 * nstruct, hipMemcpyHostToDevice, st));
    }
    AGPU_HIP(hipStreamSynchronize(st));
    b->pos_rev++;
    b->nl_rev = ~0ull;
    return AGPU_OK;
}

extern "C" int agpu_batch_set_positions(agpu_batch *b, const double *r, int flags)
{
    if (!b || (!r && b->ntot > 0)) return set_error(AGPU_ERR_ARG, "batch_set_positions: bad arguments");
    if (b->ntot > 0) {
        hipStream_t st = stream();
        AGPU_HIP(hipMemcpyAsync(b->r.ptr, r, sizeof(double) * 3 * b->ntot,
                                (flags & AGPU_HOST) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, st));
        AGPU_HIP(hipStreamSynchronize(st));
    }
    b->pos_rev++;
    return AGPU_OK;
}

extern "C" int agpu_batch_get_size(agpu_batch *b, int *nstruct, int *ntot, int64_t *npairs)
{
    if (!b) return set_error(AGPU_ERR_ARG, "batch is NULL");
    if (nstruct) *nstruct = b->nstruct;
    if (ntot) *ntot = b->ntot;
    if (npairs) *npairs = b->total;
    return AGPU_OK;
}

extern "C" int agpu_tersoff_batch_forward(agpu_batch *b, const double *theta, int precision, double *e_struct,
                                          double *f, int flags)
{
    if (!b || !e_struct || !f) return set_error(AGPU_ERR_ARG, "tersoff_batch_forward: bad arguments");
    if (precision != AGPU_PREC_FP64 && precision != AGPU_PREC_MIXED)
        return set_error(AGPU_ERR_ARG, "tersoff_batch_forward: unknown precision");
    AGPU_CHECK(check_theta(theta));
    const TbParams P = tb_params(theta);
    const int n = b->ntot, ns = b->nstruct;
    hipStream_t s = stream();
    if (n == 0) {
        if (ns && !(flags & AGPU_HOST)) AGPU_HIP(hipMemsetAsync(e_struct, 0, sizeof(double) * ns, s));
        if (ns && (flags & AGPU_HOST)) std::memset(e_struct, 0, sizeof(double) * ns);
        return AGPU_OK;
    }
    AGPU_CHECK(batch_nl(b, P.S));
    AGPU_CHECK(b->facc.reserve(sizeof(unsigned long long) * 3 * (n + 1)));
    AGPU_CHECK(b->eatom.reserve(sizeof(double) * (n + 1) + sizeof(double) * (ns + 1) + sizeof(double) * 3 * n));
    double *eatom = b->eatom.as<double>();
    double *es = eatom + n + 1;
    double *fd = es + ns + 1;
    AGPU_LAUNCH("batch_zero", k_tb_zero, dim3(256), dim3(256), 0, s, b->facc.as<unsigned long long>(), (size_t)3 * n);
    double *fout = f;
    if (flags & AGPU_HOST) {
        AGPU_HIP(hipMemcpyAsync(fd, f, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        fout = fd;
    }
    const int nb = (n + 255) / 256;
    if (precision == AGPU_PREC_MIXED)
        AGPU_LAUNCH("tersoff_batch_mixed", k_tb_forward<float>, dim3(nb), dim3(256), 0, s, n, b->r.as<double>(),
                    b->sof.as<int>(), b->cell.as<double>(), b->offsets.as<int>(), b->nbr.as<int2>(), P, eatom,
                    b->facc.as<unsigned long long>());
    else
        AGPU_LAUNCH("tersoff_batch_fp64", k_tb_forward<double>, dim3(nb), dim3(256), 0, s, n, b->r.as<double>(),
                    b->sof.as<int>(), b->cell.as<double>(), b->offsets.as<int>(), b->nbr.as<int2>(), P, eatom,
                    b->facc.as<unsigned long long>());
    AGPU_LAUNCH("batch_finish", k_tb_finish, dim3((3 * n + 255) / 256), dim3(256), 0, s, n,
                b->facc.as<unsigned long long>(), fout);
    double *eout = (flags & AGPU_HOST) ? es : e_struct;
    AGPU_LAUNCH("batch_struct_sum", k_tb_struct_sum, dim3((ns + 255) / 256), dim3(256), 0, s, ns,
                b->off.as<int>(), eatom, eout);
    if (flags & AGPU_HOST) {
        AGPU_HIP(hipMemcpyAsync(f, fd, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        AGPU_HIP(hipMemcpyAsync(e_struct, es, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
    }
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

extern "C" int agpu_tersoff_batch_backward(agpu_batch *b, const double *theta, const double *grad_e,
                                           const double *grad_f, double *grad_theta, int flags)
{
    if (!b || !grad_theta) return set_error(AGPU_ERR_ARG, "tersoff_batch_backward: bad arguments");
    AGPU_CHECK(check_theta(theta));
    const TbParams P = tb_params(theta);
    const int n = b->ntot, ns = b->nstruct;
    hipStream_t s = stream();
    double zero[TB_NP] = {0};
    if (n == 0) {
        if (flags & AGPU_HOST) std::memcpy(grad_theta, zero, sizeof zero);
        else AGPU_HIP(hipMemsetAsync(grad_theta, 0, sizeof zero, s));
        return AGPU_OK;
    }
    AGPU_CHECK(batch_nl(b, P.S));
    const int nb = (n + 255) / 256;
    AGPU_CHECK(b->partial.reserve(sizeof(double) * TB_NP * nb));
    AGPU_CHECK(b->gacc.reserve(sizeof(double) * (TB_NP + 3 * n + ns + 2)));
    double *gout = b->gacc.as<double>();
    double *gf = gout + TB_NP;
    double *ge = gf + 3 * n;
    const double *gfp = grad_f, *gep = grad_e;
    if (flags & AGPU_HOST) {
        if (grad_f) AGPU_HIP(hipMemcpyAsync(gf, grad_f, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        if (grad_e) AGPU_HIP(hipMemcpyAsync(ge, grad_e, sizeof(double) * ns, hipMemcpyHostToDevice, s));
        gfp = grad_f ? gf : nullptr;
        gep = grad_e ? ge : nullptr;
    }
    AGPU_LAUNCH("tersoff_batch_backward", k_tb_backward, dim3(nb), dim3(256), 0, s, n, b->r.as<double>(), gfp,
                gep, b->sof.as<int>(), b->cell.as<double>(), b->offsets.as<int>(), b->nbr.as<int2>(), P,
                b->partial.as<double>());
    AGPU_HIP(hipMemsetAsync(gout, 0, sizeof(double) * TB_NP, s));
    // 11 columns: column 0 -> gout[0], 1..9 -> gout[1..9], 10 -> gout[10]
    AGPU_CHECK(reduce_ev(b->partial.as<double>(), nb, TB_NP, gout, gout + 1, gout + 10));
    if (flags & AGPU_HOST)
        AGPU_HIP(hipMemcpyAsync(grad_theta, gout, sizeof(double) * TB_NP, hipMemcpyDeviceToHost, s));
    else
        AGPU_HIP(hipMemcpyAsync(grad_theta, gout, sizeof(double) * TB_NP, hipMemcpyDeviceToDevice, s));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}
