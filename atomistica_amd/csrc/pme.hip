// Smooth particle-mesh Ewald reciprocal sum on gfx950: the reciprocal part of
// the reference's PME class (pme.f90:443-478 calling the Darden routines of
// pme_kernel.f90:251-752).  Real space and self term are the Ewald class's
// kernels (ewald.hip); this file replaces its direct structure-factor sum.
//
//   k_pme_prep     thread per ion: scaled fractionals fr = K (w - anint(w) +
//                  1/2) with w from the columns of Bbox (as the reference
//                  passes it), B-spline weights and derivatives of order p in
//                  each dimension (fill_bspline), grid cell of int(fr)
//   (radix sort of the ions by cell, stable, so each cell lists its ions in
//   index order; cell starts by count + scan)
//   k_pme_pack     sorted copy of cell x index, q and theta per ion
//   k_pme_spread   thread per grid point: the charge grid as a gather -- the
//                  ions whose p^3 stencil covers the point sit in the p^3
//                  cells at +0 .. +(p-1), p^2 ranges of the sorted order;
//                  fixed order, no atomics
//   hipFFT Z2Z backward (FFTW_BACKWARD of the reference), then
//   k_pme_energy   thread per grid point: exp(-pi^2 m^2/beta^2)/(B(m) m^2),
//                  energy and virial records, Q *= eterm (m = 0 untouched)
//   hipFFT Z2Z forward, then
//   k_pme_gather   thread per ion: potential and field from the p^3 stencil;
//                  f += q E, phi += phi_rec.
// Grid layout: index g1 + K1 (g2 + K2 g3), the memory order of the
// reference's Q(nfft1, nfft2, nfft3).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hipfft/hipfft.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

constexpr int PME_MAXORD = 12;

struct PmeState {
    int order = 8;
    int K[3] = {64, 64, 64};
    bool have_plan = false;
    hipfftHandle plan;
    hipStream_t plan_stream = nullptr;
    DevBuf bsp, th, base, key, key_s, idx, idx_s, cnt, start, sort_tmp, grid, partial, b1s, ths;
    ~PmeState()
    {
        if (have_plan) (void)hipfftDestroy(plan);
    }
};

PmeState *pme_state_new(int order, const int *grid)
{
    PmeState *pm = new PmeState();
    pm->order = order;
    for (int d = 0; d < 3; ++d) pm->K[d] = grid[d];
    return pm;
}

void pme_state_free(PmeState *pm) { delete pm; }

// bsp_one_pass (pme_kernel.f90), 0-based
__host__ __device__ __forceinline__ void bsp_one_pass(double *c, double x, int k)
{
    const double div = 1.0 / (k - 1);
    c[k - 1] = div * x * c[k - 2];
    for (int j = 1; j <= k - 2; ++j) c[k - j - 1] = div * ((x + j) * c[k - j - 2] + (k - j - x) * c[k - j - 1]);
    c[0] = div * (1 - x) * c[0];
}

// fill_bspline (pme_kernel.f90): weights c and derivatives d at offset w
__host__ __device__ __forceinline__ void fill_bspline(double w, int order, double *c, double *d)
{
    c[order - 1] = 0.0;
    c[1] = w;
    c[0] = 1.0 - w;
    for (int k = 3; k <= order - 1; ++k) bsp_one_pass(c, w, k);
    d[0] = -c[0];
    for (int j = 1; j < order; ++j) d[j] = c[j - 1] - c[j];
    bsp_one_pass(c, w, order);
}

__device__ __forceinline__ int wrapi(int x, int K) { return ((x % K) + K) % K; }

// per ion: th[n][6][order] = theta(d = 0..2) then dtheta(d = 0..2)
__global__ void k_pme_prep(int nat, const double *__restrict__ r, CellArg B, int K1, int K2, int K3, int order,
                           double *__restrict__ th, int4 *__restrict__ base, int *__restrict__ key,
                           int *__restrict__ idx)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nat) return;
    const int K[3] = {K1, K2, K3};
    int ib[3];
    double c[PME_MAXORD], d[PME_MAXORD];
    double *t = th + (size_t)n * 6 * order;
    for (int e = 0; e < 3; ++e) {
        // w = x recip(1,e) + y recip(2,e) + z recip(3,e)  (get_scaled_fractionals)
        const double w = r[3 * n] * B.A[0 + 3 * e] + r[3 * n + 1] * B.A[1 + 3 * e] + r[3 * n + 2] * B.A[2 + 3 * e];
        const double fr = K[e] * (w - round(w) + 0.5);
        ib[e] = (int)fr;
        fill_bspline(fr - ib[e], order, c, d);
        for (int q = 0; q < order; ++q) {
            t[e * order + q] = c[q];
            t[(3 + e) * order + q] = d[q];
        }
    }
    base[n] = make_int4(ib[0], ib[1], ib[2], 0);
    key[n] = wrapi(ib[0], K1) + K1 * (wrapi(ib[1], K2) + K2 * wrapi(ib[2], K3));
    idx[n] = n;
}

__global__ void k_pme_count(int nat, const int *__restrict__ key, int *__restrict__ cnt)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < nat) atomicAdd(&cnt[key[n]], 1);
}

// Sorted copy of the spreading data: per sorted slot a, the ion's x cell
// and q with the theta of the three dimensions
template <int ORD>
__global__ void k_pme_pack(int nat, int K1, const int *__restrict__ idx_s, const int4 *__restrict__ base,
                           const double *__restrict__ th, const double *__restrict__ q, int *__restrict__ b1s,
                           double *__restrict__ ths)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nat) return;
    const int n = idx_s[a];
    b1s[a] = wrapi(base[n].x, K1);
    const double *t = th + (size_t)n * 6 * ORD;
    double *o = ths + (size_t)a * (3 * ORD + 1);
#pragma unroll
    for (int k = 0; k < 3 * ORD; ++k) o[k] = t[k];
    o[3 * ORD] = q[n];
}

// Q(g) = sum over the ions n whose cell is g + (p-1-t) (per dimension, mod
// K) of q_n theta1(t1) theta2(t2) theta3(t3)  (fill_charge_grid).  The ions
// are sorted by cell with x fastest, so for each (t2, t3) the p cells along
// x form one range of the sorted order (two when it wraps).
template <int ORD>
__global__ void __launch_bounds__(256)
k_pme_spread(int K1, int K2, int K3, const int *__restrict__ start, const int *__restrict__ b1s,
             const double *__restrict__ ths, double2 *__restrict__ grid)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ng = (long long)K1 * K2 * K3;
    if (g >= ng) return;
    const int g1 = (int)(g % K1), g2 = (int)((g / K1) % K2), g3 = (int)(g / ((long long)K1 * K2));
    int c2[ORD], c3[ORD];
#pragma unroll
    for (int t = 0; t < ORD; ++t) {
        int b = g2 + ORD - 1 - t, c = g3 + ORD - 1 - t;
        c2[t] = K1 * (b >= K2 ? b - K2 : b);
        c3[t] = K1 * K2 * (c >= K3 ? c - K3 : c);
    }
    // x cells g1 .. g1 + ORD - 1: [g1, hi) without wrap, then [0, g1 + ORD - K1)
    const int hi = min(g1 + ORD, K1), wrap = g1 + ORD - K1;
    constexpr int W = 3 * ORD + 1;
    double acc = 0.0;
    for (int t3 = 0; t3 < ORD; ++t3) {
        for (int t2 = 0; t2 < ORD; ++t2) {
            const int row = c3[t3] + c2[t2];
            const int a0 = start[row + g1], a1 = start[row + hi];
            for (int a = a0; a < a1; ++a) {
                const double *t = ths + (size_t)a * W;
                const int t1 = g1 + ORD - 1 - b1s[a];
                acc += t[t1] * (t[ORD + t2] * t[2 * ORD + t3] * t[3 * ORD]);
            }
            if (wrap > 0) {
                const int w0 = start[row], w1 = start[row + wrap];
                for (int a = w0; a < w1; ++a) {
                    const double *t = ths + (size_t)a * W;
                    const int t1 = g1 + ORD - 1 - (b1s[a] + K1);
                    acc += t[t1] * (t[ORD + t2] * t[2 * ORD + t3] * t[3 * ORD]);
                }
            }
        }
    }
    grid[g] = make_double2(acc, 0.0);
}

// energy_and_virial_sum (pme_kernel.f90): records [e, W column-major] of
// the already scaled (1 / (2 pi V)) terms; Q(m) *= eterm except m = 0
__global__ void __launch_bounds__(256)
k_pme_energy(int K1, int K2, int K3, CellArg B, const double *__restrict__ bsp, double fac, double pre,
             double2 *__restrict__ grid, double *__restrict__ partial)
{
    __shared__ double s_red[4][7];
    const long long ind = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ng = (long long)K1 * K2 * K3;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (ind > 0 && ind < ng) {
        const int k1 = (int)(ind % K1), k2 = (int)((ind / K1) % K2), k3 = (int)(ind / ((long long)K1 * K2));
        const int nf1 = (K1 + 1) / 2, nf2 = (K2 + 1) / 2, nf3 = (K3 + 1) / 2;
        const double m1 = k1 < nf1 ? k1 : k1 - K1, m2 = k2 < nf2 ? k2 : k2 - K2, m3 = k3 < nf3 ? k3 : k3 - K3;
        // mhat = recip . m
        const double mh1 = B.A[0] * m1 + B.A[3] * m2 + B.A[6] * m3;
        const double mh2 = B.A[1] * m1 + B.A[4] * m2 + B.A[7] * m3;
        const double mh3 = B.A[2] * m1 + B.A[5] * m2 + B.A[8] * m3;
        const double msq = mh1 * mh1 + mh2 * mh2 + mh3 * mh3;
        const double denom = bsp[k1] * bsp[K1 + k2] * bsp[K1 + K2 + k3] * msq;
        const double eterm = exp(-fac * msq) / denom;
        const double vterm = 2.0 * (fac * msq + 1.0) / msq;
        const double2 Q = grid[ind];
        const double es = eterm * (Q.x * Q.x + Q.y * Q.y) * pre;
        v[0] = es;
        v[1] = es * (vterm * mh1 * mh1 - 1.0);
        v[2] = es * (vterm * mh2 * mh2 - 1.0);
        v[3] = es * (vterm * mh3 * mh3 - 1.0);
        v[4] = es * (vterm * mh1 * mh2);
        v[5] = es * (vterm * mh1 * mh3);
        v[6] = es * (vterm * mh2 * mh3);
        grid[ind] = make_double2(eterm * Q.x, eterm * Q.y);
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) v[q] = group_sum<64>(v[q]);
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 7; ++q) s_red[wave][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 7) {
        const int q = threadIdx.x;
        const double t = ((s_red[0][q] + s_red[1][q]) + s_red[2][q]) + s_red[3][q];
        double *rec = partial + (size_t)blockIdx.x * 10;
        if (q == 0) rec[0] = t;
        else if (q == 1) rec[1 + 0] = t;
        else if (q == 2) rec[1 + 4] = t;
        else if (q == 3) rec[1 + 8] = t;
        else if (q == 4) { rec[1 + 1] = t; rec[1 + 3] = t; }
        else if (q == 5) { rec[1 + 2] = t; rec[1 + 6] = t; }
        else { rec[1 + 5] = t; rec[1 + 7] = t; }
    }
}

// potential_and_field_sum (pme_kernel.f90): f += q E, phi += phi_rec
template <int ORD>
__global__ void __launch_bounds__(256)
k_pme_gather(int nat, int K1, int K2, int K3, CellArg B, const int4 *__restrict__ base,
             const double *__restrict__ th, const double2 *__restrict__ grid, const double *__restrict__ q,
             double fac, double *__restrict__ f, double *__restrict__ phi)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nat) return;
    const int4 b = base[n];
    const double *t = th + (size_t)n * 6 * ORD;
    double t1[ORD], t2[ORD], t3[ORD], d1[ORD], d2[ORD], d3[ORD];
    int ii[ORD];
    size_t jj[ORD], kk[ORD];
#pragma unroll
    for (int q0 = 0; q0 < ORD; ++q0) {
        t1[q0] = t[q0];
        t2[q0] = t[ORD + q0];
        t3[q0] = t[2 * ORD + q0];
        d1[q0] = t[3 * ORD + q0];
        d2[q0] = t[4 * ORD + q0];
        d3[q0] = t[5 * ORD + q0];
        ii[q0] = wrapi(b.x - ORD + 1 + q0, K1);
        jj[q0] = (size_t)K1 * wrapi(b.y - ORD + 1 + q0, K2);
        kk[q0] = (size_t)K1 * K2 * wrapi(b.z - ORD + 1 + q0, K3);
    }
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int i3 = 0; i3 < ORD; ++i3) {
        for (int i2 = 0; i2 < ORD; ++i2) {
            const size_t row = kk[i3] + jj[i2];
#pragma unroll
            for (int i1 = 0; i1 < ORD; ++i1) {
                const double term = grid[row + ii[i1]].x;
                f0 = f0 + term * t1[i1] * t2[i2] * t3[i3];
                f1 = f1 - K1 * term * d1[i1] * t2[i2] * t3[i3];
                f2 = f2 - K2 * term * t1[i1] * d2[i2] * t3[i3];
                f3 = f3 - K3 * term * t1[i1] * t2[i2] * d3[i3];
            }
        }
    }
    if (phi) phi[n] += fac * f0;
    if (f) {
        const double qn = q[n];
        f[3 * (size_t)n + 0] += qn * (fac * (B.A[0] * f1 + B.A[3] * f2 + B.A[6] * f3));
        f[3 * (size_t)n + 1] += qn * (fac * (B.A[1] * f1 + B.A[4] * f2 + B.A[7] * f3));
        f[3 * (size_t)n + 2] += qn * (fac * (B.A[2] * f1 + B.A[5] * f2 + B.A[8] * f3));
    }
}

template <int ORD>
static int pme_spread_gather(bool gather, int nat, int K1, int K2, int K3, const PmeState *pm, const double *q,
                             const CellArg &B, double fac, double *f, double *phi, hipStream_t s)
{
    const long long ng = (long long)K1 * K2 * K3;
    if (!gather) {
        AGPU_LAUNCH("pme_pack", k_pme_pack<ORD>, dim3((nat + 255) / 256), dim3(256), 0, s, nat, K1,
                    pm->idx_s.as<int>(), pm->base.as<int4>(), pm->th.as<double>(), q, pm->b1s.as<int>(),
                    pm->ths.as<double>());
        AGPU_LAUNCH("pme_spread", k_pme_spread<ORD>, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, K1, K2, K3,
                    pm->start.as<int>(), pm->b1s.as<int>(), pm->ths.as<double>(), pm->grid.as<double2>());
    } else
        AGPU_LAUNCH("pme_gather", k_pme_gather<ORD>, dim3((nat + 255) / 256), dim3(256), 0, s, nat, K1, K2, K3, B,
                    pm->base.as<int4>(), pm->th.as<double>(), pm->grid.as<double2>(), q, fac, f, phi);
    return AGPU_OK;
}

static int pme_order_dispatch(int order, bool gather, int nat, int K1, int K2, int K3, const PmeState *pm,
                              const double *q, const CellArg &B, double fac, double *f, double *phi, hipStream_t s)
{
    switch (order) {
    case 3: return pme_spread_gather<3>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 4: return pme_spread_gather<4>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 5: return pme_spread_gather<5>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 6: return pme_spread_gather<6>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 7: return pme_spread_gather<7>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 8: return pme_spread_gather<8>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 9: return pme_spread_gather<9>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 10: return pme_spread_gather<10>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 11: return pme_spread_gather<11>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    case 12: return pme_spread_gather<12>(gather, nat, K1, K2, K3, pm, q, B, fac, f, phi, s);
    default: return set_error(AGPU_ERR_ARG, "PME: order must be 3 .. 12");
    }
}

// load_bsp_moduli / dftmod (pme_kernel.f90) on the host
static void bsp_moduli(int K, int order, std::vector<double> &mod)
{
    double arr[PME_MAXORD], darr[PME_MAXORD];
    fill_bspline(0.0, order, arr, darr);
    std::vector<double> bsp(K, 0.0);
    for (int i = 1; i <= order && i < K; ++i) bsp[i] = arr[i - 1];
    mod.assign(K, 0.0);
    const double twopi = 2.0 * M_PI;
    for (int k = 0; k < K; ++k) {
        double s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < K; ++j) {
            const double arg = twopi * k * j / K;
            s1 = s1 + bsp[j] * std::cos(arg);
            s2 = s2 + bsp[j] * std::sin(arg);
        }
        mod[k] = s1 * s1 + s2 * s2;
    }
    for (int k = 0; k < K; ++k)
        if (mod[k] < 1e-7) mod[k] = 0.5 * (mod[(k + K - 1) % K] + mod[(k + 1) % K]);
}

int pme_recip(PmeState *pm, const agpu_particles *p, const double *q, double beta, double *rec10, double *f,
              double *phi)
{
    const int nat = p->nat;
    const int K1 = pm->K[0], K2 = pm->K[1], K3 = pm->K[2], order = pm->order;
    const long long ng = (long long)K1 * K2 * K3;
    hipStream_t s = stream();
    if (!pm->have_plan || pm->plan_stream != s) {
        if (pm->have_plan) (void)hipfftDestroy(pm->plan);
        pm->have_plan = false;
        if (hipfftPlan3d(&pm->plan, K3, K2, K1, HIPFFT_Z2Z) != HIPFFT_SUCCESS)
            return set_error(AGPU_ERR_HIP, "PME: hipfftPlan3d failed");
        if (hipfftSetStream(pm->plan, s) != HIPFFT_SUCCESS)
            return set_error(AGPU_ERR_HIP, "PME: hipfftSetStream failed");
        pm->have_plan = true;
        pm->plan_stream = s;
        std::vector<double> all, m;
        for (int d = 0; d < 3; ++d) {
            bsp_moduli(pm->K[d], order, m);
            all.insert(all.end(), m.begin(), m.end());
        }
        AGPU_CHECK(pm->bsp.reserve(sizeof(double) * all.size()));
        AGPU_HIP(hipMemcpyAsync(pm->bsp.ptr, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice, s));
    }
    AGPU_CHECK(pm->grid.reserve(sizeof(double2) * (size_t)ng));
    AGPU_CHECK(pm->start.reserve(sizeof(int) * (size_t)(ng + 1)));
    AGPU_CHECK(pm->cnt.reserve(sizeof(int) * (size_t)(ng + 1)));
    const int nblk_g = (int)((ng + 255) / 256);
    AGPU_CHECK(pm->partial.reserve(sizeof(double) * 10 * (size_t)nblk_g));
    CellArg B;
    std::memcpy(B.A, p->cell.B, sizeof B.A);
    const double *A = p->cell.A;
    const double V = std::fabs(A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                               A[2] * (A[3] * A[7] - A[4] * A[6]));
    AGPU_CHECK(zero_dev(rec10, sizeof(double) * 10));
    if (nat > 0) {
        AGPU_CHECK(pm->th.reserve(sizeof(double) * 6 * order * (size_t)nat));
        AGPU_CHECK(pm->base.reserve(sizeof(int4) * (size_t)nat));
        AGPU_CHECK(pm->key.reserve(sizeof(int) * (size_t)nat));
        AGPU_CHECK(pm->key_s.reserve(sizeof(int) * (size_t)nat));
        AGPU_CHECK(pm->idx.reserve(sizeof(int) * (size_t)nat));
        AGPU_CHECK(pm->idx_s.reserve(sizeof(int) * (size_t)nat));
        AGPU_CHECK(pm->b1s.reserve(sizeof(int) * (size_t)nat));
        AGPU_CHECK(pm->ths.reserve(sizeof(double) * (3 * order + 1) * (size_t)nat));
        const dim3 ga((nat + 255) / 256);
        AGPU_LAUNCH("pme_prep", k_pme_prep, ga, dim3(256), 0, s, nat, p->d_r, B, K1, K2, K3, order,
                    pm->th.as<double>(), pm->base.as<int4>(), pm->key.as<int>(), pm->idx.as<int>());
        int bits = 1;
        while ((1ll << bits) < ng) ++bits;
        size_t tmp = 0;
        AGPU_HIP(rocprim::radix_sort_pairs(nullptr, tmp, pm->key.as<int>(), pm->key_s.as<int>(), pm->idx.as<int>(),
                                           pm->idx_s.as<int>(), (size_t)nat, 0, bits, s));
        size_t tmp2 = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp2, pm->cnt.as<int>(), pm->start.as<int>(), 0, (size_t)ng + 1,
                                         rocprim::plus<int>(), s));
        AGPU_CHECK(pm->sort_tmp.reserve(std::max(tmp, tmp2) + 16));
        AGPU_HIP(rocprim::radix_sort_pairs(pm->sort_tmp.ptr, tmp, pm->key.as<int>(), pm->key_s.as<int>(),
                                           pm->idx.as<int>(), pm->idx_s.as<int>(), (size_t)nat, 0, bits, s));
        AGPU_CHECK(zero_dev(pm->cnt.ptr, sizeof(int) * (size_t)(ng + 1)));
        AGPU_LAUNCH("pme_count", k_pme_count, ga, dim3(256), 0, s, nat, pm->key.as<int>(), pm->cnt.as<int>());
        AGPU_HIP(rocprim::exclusive_scan(pm->sort_tmp.ptr, tmp2, pm->cnt.as<int>(), pm->start.as<int>(), 0,
                                         (size_t)ng + 1, rocprim::plus<int>(), s));
        AGPU_CHECK(pme_order_dispatch(order, false, nat, K1, K2, K3, pm, q, B, 0.0, nullptr, nullptr, s));
    } else {
        AGPU_CHECK(zero_dev(pm->grid.ptr, sizeof(double2) * (size_t)ng));
    }
    hipfftDoubleComplex *G = reinterpret_cast<hipfftDoubleComplex *>(pm->grid.ptr);
    if (hipfftExecZ2Z(pm->plan, G, G, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
        return set_error(AGPU_ERR_HIP, "PME: backward FFT failed");
    AGPU_LAUNCH("pme_energy", k_pme_energy, dim3(nblk_g), dim3(256), 0, s, K1, K2, K3, B, pm->bsp.as<double>(),
                M_PI * M_PI / (beta * beta), 1.0 / (2.0 * M_PI * V), pm->grid.as<double2>(),
                pm->partial.as<double>());
    if (hipfftExecZ2Z(pm->plan, G, G, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return set_error(AGPU_ERR_HIP, "PME: forward FFT failed");
    if (nat > 0 && (f || phi))
        AGPU_CHECK(pme_order_dispatch(order, true, nat, K1, K2, K3, pm, q, B, 1.0 / (M_PI * V), f, phi, s));
    AGPU_CHECK(reduce_ev(pm->partial.as<double>(), nblk_g, 10, rec10, rec10 + 1, nullptr));
    return AGPU_OK;
}

}  // namespace agpu
