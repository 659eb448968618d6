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
//       then written in (j ascending, n lexicographic) order together with
//       the bond record {unit vector, length} (FP64 displacement, then an
//       FP64 and an FP32 copy).  k_bnl_rev finds the reverse entry of each.
//       The list is kept while positions do not change and the cut-off S
//       stays within the list's cut-off (built with a 2 % skin).
//   k_tb_fc / k_tb_bond / k_tb_force<R>  thread per list entry (directed
//       bond ij): fc; zeta_ij, b_ij and the bond energy; then the total
//       force of E_i on j, written to the entry (owner computes).
//       k_tb_gather adds, for each atom, the reverse entries of its row minus
//       its own entries, in row order.  No atomics, so the result is bitwise
//       reproducible.  R = double, or float for FP32 pair/triplet arithmetic
//       with FP64 displacements and FP64 accumulation.
//   k_tb_backward  thread per list entry in FP64 dual numbers: dE_ij/dtheta
//       by differentiating the bond energy by hand, with every position
//       carrying the tangent w_i = dL/dF_i.  The tangent part is then
//       sum_i w_i . d(dE/dtheta)/dr_i = -sum_i w_i . dF_i/dtheta, so one pass
//       gives sum_s gE_s dE_s/dtheta + sum_i gF_i . dF_i/dtheta with nothing
//       stored per triplet (zeta and g are recomputed).
// Reductions over atoms and structures are fixed-order.
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "batch_internal.hpp"
#include "profile.hpp"

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
      int *__restrict__ counts, int2 *__restrict__ nbr, int *__restrict__ own, double4 *__restrict__ bnd64,
      float4 *__restrict__ bnd32, int *__restrict__ flags)
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
                        if (FILL) {
                            // bond i -> image of j: r_ij = -d, unit vector -d / |d|
                            const double l = sqrt(dot3_nc(d));
                            const double u0 = -d[0] / l, u1 = -d[1] / l, u2 = -d[2] / l;
                            nbr[pos + cnt] = make_int2(j, dc_pack(n0, n1, n2));
                            own[pos + cnt] = i;
                            bnd64[pos + cnt] = make_double4(u0, u1, u2, l);
                            bnd32[pos + cnt] = make_float4((float)u0, (float)u1, (float)u2, (float)l);
                        }
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

// reverse entry of every list entry: (i -> j, n) has (j -> i, -n) in j's row
__global__ void __launch_bounds__(256)
k_bnl_rev(int ntot, const int *__restrict__ offsets, const int2 *__restrict__ nbr, int *__restrict__ rev)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    for (int p = offsets[i]; p < offsets[i + 1]; ++p) {
        const int2 e = nbr[p];
        int a, b, c;
        dc_unpack(e.y, a, b, c);
        const int back = (a == -512 || b == -512 || c == -512) ? -1 : dc_pack(-a, -b, -c);
        int q = -1;
        for (int t = offsets[e.x]; t < offsets[e.x + 1]; ++t) {
            const int2 f = nbr[t];
            if (f.x == i && f.y == back) {
                q = t;
                break;
            }
        }
        rev[p] = q;
    }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <typename R> struct TbVec;
template <> struct TbVec<double> {
    using v4 = double4;
    using v2 = double2;
};
template <> struct TbVec<float> {
    using v4 = float4;
    using v2 = float2;
};

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

// Thread per list entry (directed bond), in three launches; the bond records
// are cached with the list, so there are no position loads and no
// divisions per triplet, and every lane does O(row length) work:
//   k_tb_fc     fc, dfc of the bond;
//   k_tb_bond   zeta_ij over the other bonds k of i, b_ij, the bond energy
//               1/2 fc (V_R + b V_A), dE/dzeta_ij and the radial dE/dr_ij;
//   k_tb_force  the total force of E_i on the bond's neighbour m,
//        -dE/dr_im n_im
//        - sum_{q != m} [ (dE/dzeta_im fc_iq + dE/dzeta_iq fc_im) dg
//                         (n_iq - cos n_im) / r_im + dE/dzeta_iq dfc_im g n_im ]
//               (m as the j of triplet (i, m, q) and as the k of triplet
//               (i, q, m)), written to the entry.  By translation invariance
//               the force of E_i on i is minus the sum over its row.
// R = double, or float for FP32 pair/triplet arithmetic; forces and energies
// are accumulated in FP64.
template <typename R>
__global__ void __launch_bounds__(256)
k_tb_fc(int np, const typename TbVec<R>::v4 *__restrict__ bnd, TbParams P, typename TbVec<R>::v2 *__restrict__ fcd)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    R fc, dfc;
    tb_fc<R>(P, bnd[p].w, fc, dfc);
    fcd[p] = typename TbVec<R>::v2{fc, dfc};
}

template <typename R>
__global__ void __launch_bounds__(256)
k_tb_bond(int np, const int *__restrict__ own, const int *__restrict__ offsets,
          const typename TbVec<R>::v4 *__restrict__ bnd, const typename TbVec<R>::v2 *__restrict__ fcd, TbParams P,
          typename TbVec<R>::v2 *__restrict__ pp, double *__restrict__ ebond)
{
    using V4 = typename TbVec<R>::v4;
    using V2 = typename TbVec<R>::v2;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const R S = (R)P.S, h = (R)P.h, c2 = (R)P.c2, d2 = (R)P.d2;
    const V4 bp = bnd[p];
    if (!(bp.w < S)) {
        pp[p] = V2{0, 0};
        ebond[p] = 0.0;
        return;
    }
    const int i = own[p];
    const int beg = offsets[i], end = offsets[i + 1];
    R z = 0;
    for (int q = beg; q < end; ++q) {
        if (q == p) continue;
        const V4 bq = bnd[q];
        if (!(bq.w < S)) continue;
        const R cs = bp.x * bq.x + bp.y * bq.y + bp.z * bq.z;
        const R t = h - cs;
        z += fcd[q].x * ((R)1 + c2 * t * t / (d2 * (d2 + t * t)));
    }
    R b = 1, dbdz = 0;
    if (z > (R)0) {
        const R u = pow((R)P.beta * z, (R)P.n);
        b = pow((R)1 + u, (R)P.e);
        dbdz = (R)P.e * (R)P.n * b * u / (z * ((R)1 + u));
    }
    const V2 fp = fcd[p];
    const R rij = bp.w;
    const R exa = exp(-(R)P.mu * rij), exr = exp(-(R)P.lam * rij);
    const R VA = -(R)P.B * exa, dVA = (R)P.B * (R)P.mu * exa;
    const R VR = (R)P.A * exr, dVR = -(R)P.A * (R)P.lam * exr;
    ebond[p] = 0.5 * (double)(fp.x * (VR + b * VA));
    const R dEdr = (R)0.5 * (fp.y * (VR + b * VA) + fp.x * (dVR + b * dVA));
    const R pre = (R)0.5 * fp.x * VA * dbdz;
    pp[p] = V2{pre, dEdr};
}

template <typename R>
__global__ void __launch_bounds__(256)
k_tb_force(int np, const int *__restrict__ own, const int *__restrict__ offsets,
           const typename TbVec<R>::v4 *__restrict__ bnd, const typename TbVec<R>::v2 *__restrict__ fcd,
           const typename TbVec<R>::v2 *__restrict__ pp, TbParams P, double *__restrict__ fslot)
{
    using V4 = typename TbVec<R>::v4;
    using V2 = typename TbVec<R>::v2;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= np) return;
    const R S = (R)P.S, h = (R)P.h, c2 = (R)P.c2, d2 = (R)P.d2;
    const V4 bm = bnd[m];
    double F[3] = {0.0, 0.0, 0.0};
    if (bm.w < S) {
        const int i = own[m];
        const int beg = offsets[i], end = offsets[i + 1];
        const V2 fm = fcd[m], pm = pp[m];
        F[0] = -(double)(pm.y * bm.x);
        F[1] = -(double)(pm.y * bm.y);
        F[2] = -(double)(pm.y * bm.z);
        const R inv_r = (R)1 / bm.w;
        for (int q = beg; q < end; ++q) {
            if (q == m) continue;
            const V4 bq = bnd[q];
            if (!(bq.w < S)) continue;
            const R fq = fcd[q].x, preq = pp[q].x;
            const R cs = bm.x * bq.x + bm.y * bq.y + bm.z * bq.z;
            const R t = h - cs;
            const R den = d2 + t * t;
            const R g = (R)1 + c2 * t * t / (d2 * den);
            const R dg = (R)-2 * c2 * t / (den * den);
            const R a = (pm.x * fq + preq * fm.x) * dg * inv_r;
            const R bc = preq * fm.y * g;
            F[0] -= (double)(a * (bq.x - cs * bm.x) + bc * bm.x);
            F[1] -= (double)(a * (bq.y - cs * bm.y) + bc * bm.y);
            F[2] -= (double)(a * (bq.z - cs * bm.z) + bc * bm.z);
        }
    }
    fslot[3 * (size_t)m + 0] = F[0];
    fslot[3 * (size_t)m + 1] = F[1];
    fslot[3 * (size_t)m + 2] = F[2];
}

// Fused forward over a chunk of whole rows (at most TB_CHUNK list entries,
// one lane per entry), staged in LDS: fc of every bond, then zeta_ij, b_ij,
// the bond energy, dE/dzeta and dE/dr, then the force of E_i on each
// neighbour (the same expressions as k_tb_fc / k_tb_bond / k_tb_force), and
// finally, per atom of the chunk, its own force (minus the sum of its row's
// entry forces, in row order) and its energy.  k_tb_gather then only adds the
// reverse entries.  Used when every row fits a chunk.
constexpr int TB_CHUNK = 256;

template <typename R>
__global__ void __launch_bounds__(TB_CHUNK)
k_tb_fused(const int *__restrict__ chunk, const int *__restrict__ own, const int *__restrict__ offsets,
           const typename TbVec<R>::v4 *__restrict__ bnd, TbParams P, double *__restrict__ fslot,
           double *__restrict__ fown, double *__restrict__ eatom)
{
    using V4 = typename TbVec<R>::v4;
    __shared__ V4 s_b[TB_CHUNK];
    __shared__ R s_fc[TB_CHUNK], s_dfc[TB_CHUNK], s_pre[TB_CHUNK], s_dEdr[TB_CHUNK];
    __shared__ double s_F[3][TB_CHUNK], s_e[TB_CHUNK];
    const int a0 = chunk[blockIdx.x], a1 = chunk[blockIdx.x + 1];
    const int p0 = offsets[a0];
    const int ne = offsets[a1] - p0;
    const int t = threadIdx.x;
    const bool act = t < ne;
    const R S = (R)P.S, h = (R)P.h, c2 = (R)P.c2, d2 = (R)P.d2;
    V4 b{0, 0, 0, 0};
    int beg = 0, end = 0;
    if (act) {
        b = bnd[p0 + t];
        const int i = own[p0 + t];
        beg = offsets[i] - p0;
        end = offsets[i + 1] - p0;
        R fc, dfc;
        tb_fc<R>(P, b.w, fc, dfc);
        s_b[t] = b;
        s_fc[t] = fc;
        s_dfc[t] = dfc;
    }
    __syncthreads();
    const bool live = act && b.w < S;
    double e = 0.0;
    R pre = 0, dEdr = 0;
    if (live) {
        R z = 0;
        for (int q = beg; q < end; ++q) {
            if (q == t) continue;
            const V4 bq = s_b[q];
            if (!(bq.w < S)) continue;
            const R cs = b.x * bq.x + b.y * bq.y + b.z * bq.z;
            const R tt = h - cs;
            z += s_fc[q] * ((R)1 + c2 * tt * tt / (d2 * (d2 + tt * tt)));
        }
        R bo = 1, dbdz = 0;
        if (z > (R)0) {
            const R u = pow((R)P.beta * z, (R)P.n);
            bo = pow((R)1 + u, (R)P.e);
            dbdz = (R)P.e * (R)P.n * bo * u / (z * ((R)1 + u));
        }
        const R fc = s_fc[t], dfc = s_dfc[t];
        const R exa = exp(-(R)P.mu * b.w), exr = exp(-(R)P.lam * b.w);
        const R VA = -(R)P.B * exa, dVA = (R)P.B * (R)P.mu * exa;
        const R VR = (R)P.A * exr, dVR = -(R)P.A * (R)P.lam * exr;
        e = 0.5 * (double)(fc * (VR + bo * VA));
        dEdr = (R)0.5 * (dfc * (VR + bo * VA) + fc * (dVR + bo * dVA));
        pre = (R)0.5 * fc * VA * dbdz;
    }
    if (act) {
        s_pre[t] = pre;
        s_dEdr[t] = dEdr;
        s_e[t] = e;
    }
    __syncthreads();
    double F[3] = {0.0, 0.0, 0.0};
    if (live) {
        const R fm = s_fc[t], dfm = s_dfc[t];
        F[0] = -(double)(dEdr * b.x);
        F[1] = -(double)(dEdr * b.y);
        F[2] = -(double)(dEdr * b.z);
        const R inv_r = (R)1 / b.w;
        for (int q = beg; q < end; ++q) {
            if (q == t) continue;
            const V4 bq = s_b[q];
            if (!(bq.w < S)) continue;
            const R fq = s_fc[q], preq = s_pre[q];
            const R cs = b.x * bq.x + b.y * bq.y + b.z * bq.z;
            const R tt = h - cs;
            const R den = d2 + tt * tt;
            const R g = (R)1 + c2 * tt * tt / (d2 * den);
            const R dg = (R)-2 * c2 * tt / (den * den);
            const R a = (pre * fq + preq * fm) * dg * inv_r;
            const R bc = preq * dfm * g;
            F[0] -= (double)(a * (bq.x - cs * b.x) + bc * b.x);
            F[1] -= (double)(a * (bq.y - cs * b.y) + bc * b.y);
            F[2] -= (double)(a * (bq.z - cs * b.z) + bc * b.z);
        }
    }
    if (act) {
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            fslot[3 * (size_t)(p0 + t) + x] = F[x];
            s_F[x][t] = F[x];
        }
    }
    __syncthreads();
    for (int a = a0 + t; a < a1; a += TB_CHUNK) {
        const int q0 = offsets[a] - p0, q1 = offsets[a + 1] - p0;
        double fi[3] = {0.0, 0.0, 0.0}, ea = 0.0;
        for (int q = q0; q < q1; ++q) {
            fi[0] -= s_F[0][q];
            fi[1] -= s_F[1][q];
            fi[2] -= s_F[2][q];
            ea += s_e[q];
        }
#pragma unroll
        for (int x = 0; x < 3; ++x) fown[3 * (size_t)a + x] = fi[x];
        eatom[a] = ea;
    }
}

// owner-computes gather: f_a += sum over a's row of the force the
// neighbour's energy put on a (its reverse entry), plus a's own force: from
// fown (OWN, the fused kernel) or as minus the sum of a's own entries.  8
// lanes per atom take every 8th entry of the row, then a fixed butterfly adds
// the 8 partial sums: no atomics, fixed order.
constexpr int TB_GL = 8;

template <bool OWN>
__global__ void __launch_bounds__(256)
k_tb_gather(int ntot, const int *__restrict__ offsets, const int *__restrict__ rev,
            const double *__restrict__ fslot, const double *__restrict__ fown, double *__restrict__ f)
{
    const int a = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / TB_GL);
    const int l = threadIdx.x % TB_GL;
    if (a >= ntot) return;        // uniform over the lane group
    double F[3] = {0.0, 0.0, 0.0};
    for (int p = offsets[a] + l; p < offsets[a + 1]; p += TB_GL) {
        const int q = rev[p];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double in = q >= 0 ? fslot[3 * (size_t)q + x] : 0.0;
            F[x] += OWN ? in : in - fslot[3 * (size_t)p + x];
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) F[x] = group_sum<TB_GL>(F[x]);
    if (l == 0) {
#pragma unroll
        for (int x = 0; x < 3; ++x) f[3 * (size_t)a + x] += OWN ? fown[3 * (size_t)a + x] + F[x] : F[x];
    }
}

// energy per structure: one wavefront per structure over its atoms' energies
// (ATOMS) or its atoms' bond energies, lane-strided partial sums, then a
// fixed butterfly (every lane ends with the same value)
template <bool ATOMS>
__global__ void __launch_bounds__(256)
k_tb_struct_sum(int nstruct, const int *__restrict__ off, const int *__restrict__ offsets,
                const double *__restrict__ e_in, double *__restrict__ es)
{
    const int s = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x & 63;
    if (s >= nstruct) return;
    const int p0 = ATOMS ? off[s] : offsets[off[s]], p1 = ATOMS ? off[s + 1] : offsets[off[s + 1]];
    double e = 0.0;
    for (int p = p0 + lane; p < p1; p += 64) e += e_in[p];
    e = group_sum<64>(e);
    if (lane == 0) es[s] = e;
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

// unit vector and length of the cached bond i -> j (r_ij = r_j - r_i - A.dc)
// with tangents along w: d r_ij = w_j - w_i, d|r| = n.dr, dn = (dr - n (n.dr)) / |r|
__device__ __forceinline__ void tb_dunit(const double4 &bv, int j, const double *wi, const double *__restrict__ w,
                                         Dl *n, Dl &len)
{
    double t[3] = {0.0, 0.0, 0.0};
    if (w) {
        t[0] = w[3 * j] - wi[0];
        t[1] = w[3 * j + 1] - wi[1];
        t[2] = w[3 * j + 2] - wi[2];
    }
    const double nt = bv.x * t[0] + bv.y * t[1] + bv.z * t[2];
    len = Dl{bv.w, nt};
    n[0] = Dl{bv.x, (t[0] - bv.x * nt) / bv.w};
    n[1] = Dl{bv.y, (t[1] - bv.y * nt) / bv.w};
    n[2] = Dl{bv.z, (t[2] - bv.z * nt) / bv.w};
}

// fc and its derivatives with respect to R and S, each with its slope d/dr
// (the triplet loop forms their tangents without a sincos per triplet)
struct alignas(16) TbCut {
    double fc, dfc, fcR, dfcR, fcS, dfcS;
};

__device__ __forceinline__ TbCut tb_cut(const TbParams &P, double r)
{
    TbCut c{r <= P.R ? 1.0 : 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (r > P.R && r < P.S) {
        const double L = P.S - P.R, k = 0.5 * M_PI / (L * L);
        double sx, cx;
        sincos(P.fac * (r - P.R), &sx, &cx);
        c.fc = 0.5 * (1.0 + cx);
        c.dfc = -0.5 * P.fac * sx;
        // x = pi (r - R) / (S - R): dx/dR = -pi (S - r) / L^2, dx/dS = -pi (r - R) / L^2
        c.fcR = k * sx * (P.S - r);
        c.dfcR = k * (cx * P.fac * (P.S - r) - sx);
        c.fcS = k * sx * (r - P.R);
        c.dfcS = k * (cx * P.fac * (r - P.R) + sx);
    }
    return c;
}

__global__ void __launch_bounds__(256)
k_tb_cut(int np, const double4 *__restrict__ bnd, TbParams P, TbCut *__restrict__ cut)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    cut[p] = tb_cut(P, bnd[p].w);
}

// one bond in dual numbers: v[0..2] unit vector, v[3] length, v[4] fc,
// v[5] dfc/dR, v[6] dfc/dS
struct TbDual {
    Dl v[7];
};
constexpr int TD_LEN = 3, TD_FC = 4, TD_FCR = 5, TD_FCS = 6;

__device__ __forceinline__ TbDual tb_dual(const double4 &bv, const TbCut &c, int j, const double *wi,
                                          const double *__restrict__ w)
{
    TbDual d;
    tb_dunit(bv, j, wi, w, d.v, d.v[TD_LEN]);
    const double lt = d.v[TD_LEN].t;
    d.v[TD_FC] = Dl{c.fc, c.dfc * lt};
    d.v[TD_FCR] = Dl{c.fcR, c.dfcR * lt};
    d.v[TD_FCS] = Dl{c.fcS, c.dfcS * lt};
    return d;
}

constexpr int TB_NP = 11;

// dE_ij/dtheta of one bond (value and tangent) from its dual data and the
// other bonds k of the row, read through get(k, bond) (false: k is beyond S)
template <class Get>
__device__ __forceinline__ void tb_bond_grad(const TbParams &P, const TbDual &bij, int p, int beg, int end,
                                             Get get, Dl *g)
{
    const Dl rij = bij.v[TD_LEN];
    const Dl exa = dexp(-P.mu * rij), exr = dexp(-P.lam * rij);
    const Dl VA = -P.B * exa, VR = P.A * exr;
    Dl z = dl(0.0), zc = dl(0.0), zd = dl(0.0), zh = dl(0.0), zR = dl(0.0), zS = dl(0.0);
    for (int kb = beg; kb < end; ++kb) {
        if (kb == p) continue;
        TbDual bk;
        if (!get(kb, bk)) continue;
        const Dl cs = bij.v[0] * bk.v[0] + bij.v[1] * bk.v[1] + bij.v[2] * bk.v[2];
        const Dl t = P.h - cs;
        const Dl den = P.d2 + t * t;
        const Dl inv = 1.0 / den, inv2 = inv * inv;
        const Dl gq = 1.0 + P.c2d2 - P.c2 * inv;
        // dg/dc = 2c/d^2 - 2c/den; dg/dd = -2c^2/d^3 + 2c^2 d/den^2; dg/dh = 2c^2 t/den^2
        const Dl gc = (2.0 * P.c / P.d2) - (2.0 * P.c) * inv;
        const Dl gd = (-2.0 * P.c2 / (P.d2 * P.d)) + (2.0 * P.c2 * P.d) * inv2;
        const Dl gh = (2.0 * P.c2) * (t * inv2);
        const Dl fck = bk.v[TD_FC];
        z = z + fck * gq;
        zc = zc + fck * gc;
        zd = zd + fck * gd;
        zh = zh + fck * gh;
        zR = zR + bk.v[TD_FCR] * gq;
        zS = zS + bk.v[TD_FCS] * gq;
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
    const Dl hf = 0.5 * bij.v[TD_FC];
    const Dl pair = VR + b * VA;
    const Dl pz = hf * VA * bz;        // dE/dzeta
    g[0] = hf * exr;                           // A
    g[1] = -(hf * b * exa);                    // B
    g[2] = -(hf * rij * VR);                   // lambda
    g[3] = -(hf * b * rij * VA);               // mu
    g[4] = hf * VA * bb;                       // beta
    g[5] = hf * VA * bn;                       // n
    g[6] = pz * zc;                            // c
    g[7] = pz * zd;                            // d
    g[8] = pz * zh;                            // h
    g[9] = 0.5 * bij.v[TD_FCR] * pair + pz * zR;     // R
    g[10] = 0.5 * bij.v[TD_FCS] * pair + pz * zS;    // S
}

// block sum of the 11 per-lane values in a fixed order -> partial[blockIdx.x]
__device__ __forceinline__ void tb_block_partial(double *acc, double (*s_red)[TB_NP], double *__restrict__ partial)
{
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

// thread per list entry, the other bonds' dual data from cut[] and bnd[]
__global__ void __launch_bounds__(256)
k_tb_backward(int np, const int *__restrict__ own, const int2 *__restrict__ nbr, const int *__restrict__ offsets,
              const double4 *__restrict__ bnd, const TbCut *__restrict__ cut, const double *__restrict__ w,
              const double *__restrict__ gE, const int *__restrict__ sof, TbParams P, double *__restrict__ partial)
{
    __shared__ double s_red[4][TB_NP];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[TB_NP];
#pragma unroll
    for (int q = 0; q < TB_NP; ++q) acc[q] = 0.0;
    const double4 bp = p < np ? bnd[p] : make_double4(0.0, 0.0, 0.0, P.S);
    if (bp.w < P.S) {
        const int i = own[p];
        const double wi[3] = {w ? w[3 * i] : 0.0, w ? w[3 * i + 1] : 0.0, w ? w[3 * i + 2] : 0.0};
        const double ge = gE ? gE[sof[i]] : 0.0;
        const TbDual bij = tb_dual(bp, cut[p], nbr[p].x, wi, w);
        Dl g[TB_NP];
        tb_bond_grad(P, bij, p, offsets[i], offsets[i + 1],
                     [&](int k, TbDual &o) {
                         const double4 bk = bnd[k];
                         if (!(bk.w < P.S)) return false;
                         o = tb_dual(bk, cut[k], nbr[k].x, wi, w);
                         return true;
                     },
                     g);
        // value weighted by dL/dE_s; tangent = -sum w . dF/dtheta
#pragma unroll
        for (int q = 0; q < TB_NP; ++q) acc[q] = ge * g[q].v - g[q].t;
    }
    tb_block_partial(acc, s_red, partial);
}

// a chunk of whole rows per block: every entry's dual data computed once and
// staged in LDS, then one lane per entry
__global__ void __launch_bounds__(TB_CHUNK)
k_tb_backward_fused(const int *__restrict__ chunk, const int *__restrict__ own, const int2 *__restrict__ nbr,
                    const int *__restrict__ offsets, const double4 *__restrict__ bnd, const double *__restrict__ w,
                    const double *__restrict__ gE, const int *__restrict__ sof, TbParams P,
                    double *__restrict__ partial)
{
    __shared__ double s_d[14][TB_CHUNK];
    __shared__ double s_red[4][TB_NP];
    const int a0 = chunk[blockIdx.x], a1 = chunk[blockIdx.x + 1];
    const int p0 = offsets[a0];
    const int ne = offsets[a1] - p0;
    const int t = threadIdx.x;
    const bool act = t < ne;
    TbDual me;
    int i = 0;
    double r = P.S;
    if (act) {
        const double4 bv = bnd[p0 + t];
        i = own[p0 + t];
        r = bv.w;
        const double wi[3] = {w ? w[3 * i] : 0.0, w ? w[3 * i + 1] : 0.0, w ? w[3 * i + 2] : 0.0};
        me = tb_dual(bv, tb_cut(P, bv.w), nbr[p0 + t].x, wi, w);
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            s_d[2 * q][t] = me.v[q].v;
            s_d[2 * q + 1][t] = me.v[q].t;
        }
    }
    __syncthreads();
    double acc[TB_NP];
#pragma unroll
    for (int q = 0; q < TB_NP; ++q) acc[q] = 0.0;
    if (r < P.S) {
        const double ge = gE ? gE[sof[i]] : 0.0;
        Dl g[TB_NP];
        tb_bond_grad(P, me, t, offsets[i] - p0, offsets[i + 1] - p0,
                     [&](int k, TbDual &o) {
                         if (!(s_d[2 * TD_LEN][k] < P.S)) return false;
#pragma unroll
                         for (int q = 0; q < 7; ++q) o.v[q] = Dl{s_d[2 * q][k], s_d[2 * q + 1][k]};
                         return true;
                     },
                     g);
#pragma unroll
        for (int q = 0; q < TB_NP; ++q) acc[q] = ge * g[q].v - g[q].t;
    }
    tb_block_partial(acc, s_red, partial);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
// The list is kept while positions are unchanged and rc stays within its
// cut-off; it is built at rc plus a 2 % skin so that a cut-off that moves
// during fitting does not rebuild it every step (the kernels test r < S).
constexpr double TB_SKIN = 1.02;

static int batch_nl(agpu_batch *b, double rc_req)
{
    if (b->nl_rev == b->pos_rev && rc_req <= b->nl_cut) return AGPU_OK;
    const double rc = TB_SKIN * rc_req;
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
                nullptr, b->counts.as<int>(), nullptr, nullptr, nullptr, nullptr, flags);
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
    AGPU_CHECK(b->bnd64.reserve(sizeof(double4) * ((size_t)total + 1)));
    AGPU_CHECK(b->bnd32.reserve(sizeof(float4) * ((size_t)total + 1)));
    AGPU_CHECK(b->rev.reserve(sizeof(int) * ((size_t)total + 1)));
    AGPU_CHECK(b->own.reserve(sizeof(int) * ((size_t)total + 1)));
    AGPU_LAUNCH("batch_nl_fill", k_bnl<true>, dim3(nb), dim3(256), 0, s, n, b->r.as<double>(), G, rc,
                b->offsets.as<int>(), b->counts.as<int>(), b->nbr.as<int2>(), b->own.as<int>(), b->bnd64.as<double4>(),
                b->bnd32.as<float4>(), flags);
    AGPU_LAUNCH("batch_nl_rev", k_bnl_rev, dim3(nb), dim3(256), 0, s, n, b->offsets.as<int>(),
                b->nbr.as<int2>(), b->rev.as<int>());
    AGPU_HIP(hipMemcpyAsync(b->stats.h, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (b->stats.h[0] & DERR_IMAGE_RANGE)
        return set_error(AGPU_ERR_DOMAIN, "batch: image shift beyond +-511 cells");
    // chunks of whole rows, at most TB_CHUNK entries each (greedy, in atom order)
    std::vector<int> hoff(n + 1), ch;
    AGPU_HIP(hipMemcpyAsync(hoff.data(), b->offsets.ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    ch.push_back(0);
    bool fits = true;
    for (int a = 0; a < n; ++a) {
        if (hoff[a + 1] - hoff[a] > TB_CHUNK) {
            fits = false;
            break;
        }
        if (hoff[a + 1] - hoff[ch.back()] > TB_CHUNK) ch.push_back(a);
    }
    ch.push_back(n);
    b->nchunk = 0;
    if (fits) {
        AGPU_CHECK(b->chunk.reserve(sizeof(int) * ch.size()));
        AGPU_HIP(hipMemcpyAsync(b->chunk.ptr, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice, s));
        AGPU_HIP(hipStreamSynchronize(s));
        b->nchunk = (int)ch.size() - 1;
    }
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
    b->struct_rev++;
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
    const int np = (int)b->total;
    const bool mixed = precision == AGPU_PREC_MIXED;
    AGPU_CHECK(b->fslot.reserve(sizeof(double) * 3 * (np + 1)));
    // [E per structure | staged host forces | E per atom]
    AGPU_CHECK(b->eatom.reserve(sizeof(double) * ((ns + 1) + 3 * (size_t)n + (n + 1))));
    double *es = b->eatom.as<double>();
    double *fd = es + ns + 1;
    double *fout = f;
    if (flags & AGPU_HOST) {
        AGPU_HIP(hipMemcpyAsync(fd, f, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        fout = fd;
    }
    const int *own = b->own.as<int>(), *offs = b->offsets.as<int>();
    const int nbg = (int)(((size_t)n * TB_GL + 255) / 256);
    double *eout = (flags & AGPU_HOST) ? es : e_struct;
    if (b->nchunk > 0) {
        // every row fits a chunk: one fused launch, then the reverse gather
        AGPU_CHECK(b->fown.reserve(sizeof(double) * 3 * (n + 1)));
        double *eat = fd + 3 * (size_t)n;
        if (mixed)
            AGPU_LAUNCH("tersoff_batch_mixed", k_tb_fused<float>, dim3(b->nchunk), dim3(TB_CHUNK), 0, s,
                        b->chunk.as<int>(), own, offs, b->bnd32.as<float4>(), P, b->fslot.as<double>(),
                        b->fown.as<double>(), eat);
        else
            AGPU_LAUNCH("tersoff_batch_fp64", k_tb_fused<double>, dim3(b->nchunk), dim3(TB_CHUNK), 0, s,
                        b->chunk.as<int>(), own, offs, b->bnd64.as<double4>(), P, b->fslot.as<double>(),
                        b->fown.as<double>(), eat);
        AGPU_LAUNCH("batch_gather", k_tb_gather<true>, dim3(nbg), dim3(256), 0, s, n, offs, b->rev.as<int>(),
                    b->fslot.as<double>(), b->fown.as<double>(), fout);
        AGPU_LAUNCH("batch_struct_sum", k_tb_struct_sum<true>, dim3((ns + 3) / 4), dim3(256), 0, s, ns,
                    b->off.as<int>(), offs, eat, eout);
    } else {
        // a row longer than a chunk: one launch per pass over the entries
        const size_t rec2 = mixed ? sizeof(float2) : sizeof(double2);
        AGPU_CHECK(b->fcd.reserve(rec2 * (np + 1)));
        AGPU_CHECK(b->pp.reserve(rec2 * (np + 1)));
        AGPU_CHECK(b->ebond.reserve(sizeof(double) * (np + 1)));
        if (np > 0) {
            const int nbp = (np + 255) / 256;
            if (mixed) {
                const float4 *bnd = b->bnd32.as<float4>();
                AGPU_LAUNCH("batch_fc", k_tb_fc<float>, dim3(nbp), dim3(256), 0, s, np, bnd, P,
                            b->fcd.as<float2>());
                AGPU_LAUNCH("tersoff_batch_mixed", k_tb_bond<float>, dim3(nbp), dim3(256), 0, s, np, own, offs,
                            bnd, b->fcd.as<float2>(), P, b->pp.as<float2>(), b->ebond.as<double>());
                AGPU_LAUNCH("tersoff_batch_force_mixed", k_tb_force<float>, dim3(nbp), dim3(256), 0, s, np, own,
                            offs, bnd, b->fcd.as<float2>(), b->pp.as<float2>(), P, b->fslot.as<double>());
            } else {
                const double4 *bnd = b->bnd64.as<double4>();
                AGPU_LAUNCH("batch_fc", k_tb_fc<double>, dim3(nbp), dim3(256), 0, s, np, bnd, P,
                            b->fcd.as<double2>());
                AGPU_LAUNCH("tersoff_batch_fp64", k_tb_bond<double>, dim3(nbp), dim3(256), 0, s, np, own, offs,
                            bnd, b->fcd.as<double2>(), P, b->pp.as<double2>(), b->ebond.as<double>());
                AGPU_LAUNCH("tersoff_batch_force_fp64", k_tb_force<double>, dim3(nbp), dim3(256), 0, s, np, own,
                            offs, bnd, b->fcd.as<double2>(), b->pp.as<double2>(), P, b->fslot.as<double>());
            }
        }
        AGPU_LAUNCH("batch_gather", k_tb_gather<false>, dim3(nbg), dim3(256), 0, s, n, offs, b->rev.as<int>(),
                    b->fslot.as<double>(), nullptr, fout);
        AGPU_LAUNCH("batch_struct_sum", k_tb_struct_sum<false>, dim3((ns + 3) / 4), dim3(256), 0, s, ns,
                    b->off.as<int>(), offs, b->ebond.as<double>(), eout);
    }
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
    const int np = (int)b->total;
    const int nb = np > 0 ? (np + 255) / 256 : 1;
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
    int nparts = nb;
    if (b->nchunk > 0) {
        nparts = b->nchunk;
        AGPU_CHECK(b->partial.reserve(sizeof(double) * TB_NP * nparts));
        AGPU_LAUNCH("tersoff_batch_backward", k_tb_backward_fused, dim3(b->nchunk), dim3(TB_CHUNK), 0, s,
                    b->chunk.as<int>(), b->own.as<int>(), b->nbr.as<int2>(), b->offsets.as<int>(),
                    b->bnd64.as<double4>(), gfp, gep, b->sof.as<int>(), P, b->partial.as<double>());
    } else {
        AGPU_CHECK(b->cut.reserve(sizeof(TbCut) * (np + 1)));
        if (np > 0)
            AGPU_LAUNCH("batch_cut", k_tb_cut, dim3(nb), dim3(256), 0, s, np, b->bnd64.as<double4>(), P,
                        b->cut.as<TbCut>());
        AGPU_LAUNCH("tersoff_batch_backward", k_tb_backward, dim3(nb), dim3(256), 0, s, np, b->own.as<int>(),
                    b->nbr.as<int2>(), b->offsets.as<int>(), b->bnd64.as<double4>(), b->cut.as<TbCut>(), gfp,
                    gep, b->sof.as<int>(), P, b->partial.as<double>());
    }
    AGPU_HIP(hipMemsetAsync(gout, 0, sizeof(double) * TB_NP, s));
    // 11 columns: column 0 -> gout[0], 1..9 -> gout[1..9], 10 -> gout[10]
    AGPU_CHECK(reduce_ev(b->partial.as<double>(), nparts, TB_NP, gout, gout + 1, gout + 10));
    if (flags & AGPU_HOST)
        AGPU_HIP(hipMemcpyAsync(grad_theta, gout, sizeof(double) * TB_NP, hipMemcpyDeviceToHost, s));
    else
        AGPU_HIP(hipMemcpyAsync(grad_theta, gout, sizeof(double) * TB_NP, hipMemcpyDeviceToDevice, s));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}
