// Ewald Coulomb sum (e^2/Angstrom) on gfx950.
//
// Real space and self term follow pme_energy_and_forces / pme_potential
// (src/potentials/coulomb/pme.f90:215-321, 386-437, 467-472): erfc(beta r)/r
// over every neighbour image inside the cutoff, charges with |q| <= 1e-10
// skipped, self energy -beta/sqrt(pi) sum q^2.  Here each atom owns its full
// row (one wavefront per atom), so the reference's i<j scatter becomes a
// gather with the 1/2 of the directed double count -- no atomics.
//
// Reciprocal space (Ewald class): direct structure-factor sum over the half space of
// k = 2 pi Bbox^T n with |k| <= kmax (the reference's PME class uses smooth PME,
// pme_kernel.f90:251-752, built in pme.hip).  S(k) is accumulated in FP64:
//   k_ewald_sk     thread = run of up to L consecutive n3 in one (n1, n2)
//                  column; ions staged through LDS in chunks; the phase of
//                  the first k comes from sincospi, the rest of the run by
//                  complex multiplication with exp(2 pi i s3).
//   k_ewald_kenergy thread = k: c(k) = exp(-k^2/4beta^2)/k^2, energy and
//                  virial records, and T(k) = c(k) S(k) for the gather.
//   k_ewald_kforce thread = ion: gathers sum_k T(k) exp(-i k.r) over k staged
//                  through LDS; force and electrostatic potential per ion.
// The PME class (agpu_pme_create) shares the real space and self term and
// takes its reciprocal sum from pme.hip.
#include <cmath>
#include <cstring>
#include <vector>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

constexpr int EW_L = 16;       // k per run in k_ewald_sk
constexpr int EW_CHUNK = 256;  // ions per LDS chunk

struct EwItem {
    int n1, n2, n3s, cnt, kofs, pad0, pad1, pad2;
};

struct EwaldState {
    double cutoff = 10.0, beta = 0.0, kmax = 0.0;
    double beta_req = 0.0, kmax_req = 0.0;
    double cell_key[9];
    bool have_k = false;
    int K = 0;
    int nitems = 0;
    std::vector<EwItem> items;
    std::vector<int> kpack;
    DevBuf d_items, d_kpack, d_S, d_T, d_frac, partial, phi_tmp, q4;
    // smooth PME reciprocal sum instead of the direct S(k) sum (PME class)
    PmeState *pme = nullptr;
    ~EwaldState()
    {
        if (pme) pme_state_free(pme);
    }
};

static void ewald_free(void *p) { delete static_cast<EwaldState *>(p); }

__device__ __forceinline__ int kp_pack(int a, int b, int c) { return dc_pack(a, b, c); }

// ---------------------------------------------------------------------------
// real space
// ---------------------------------------------------------------------------
template <typename R> __device__ __forceinline__ R rerfc(R x);
template <> __device__ __forceinline__ double rerfc<double>(double x) { return erfc(x); }
template <> __device__ __forceinline__ float rerfc<float>(float x) { return erfcf(x); }
template <typename R> __device__ __forceinline__ R rexp_(R x);
template <> __device__ __forceinline__ double rexp_<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float rexp_<float>(float x) { return expf(x); }

template <typename T> __device__ __forceinline__ T wsum(T v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// {x, y, z, q} per slot of the list's bin order (caller atom order[p])
__global__ void k_ewald_pack(int nat, const SortedAtom *__restrict__ atoms, const int *__restrict__ order,
                             const double *__restrict__ q, double4 *__restrict__ q4)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    const SortedAtom a = atoms[p];
    q4[p] = make_double4(a.x, a.y, a.z, q[order[p]]);
}

template <typename R>
__global__ void __launch_bounds__(256)
k_ewald_real(int nat, const int *__restrict__ order, const double4 *__restrict__ q4, CellArg C, NlRows L,
             double cutoff_sq, double beta, double sqrt_alpha_pi, double *__restrict__ f, double *__restrict__ phi,
             double *__restrict__ partial)
{
    __shared__ double s_red[4][10];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ai = blockIdx.x * 4 + wave;
    double e = 0.0, acc[3] = {0, 0, 0}, ph = 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};
    int i = -1;
    double qi = 0.0;
    if (ai < nat) {
        i = order[ai];
        const double4 hi = q4[ai];
        qi = hi.w;
        if (fabs(qi) > 1e-10) {
            const double ri[3] = {hi.x, hi.y, hi.z};
            int n;
            const int2 *row = nl_row(L, ai, n);
            const R b = (R)beta, sap = (R)sqrt_alpha_pi, al = (R)(beta * beta);
            for (int s = lane; s < n; s += 64) {
                int2 en = row[s];
                const double4 qa = q4[en.x];
                const double qj = qa.w;
                if (!(fabs(qj) > 1e-10)) continue;
                int a, bb, c;
                dc_unpack(en.y, a, bb, c);
                const double rj[3] = {qa.x, qa.y, qa.z};
                double d[3];
                delta_r(C, ri, rj, a, bb, c, d);
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (!(r2 < cutoff_sq)) continue;
                double ar = sqrt(r2);
                R arr = (R)ar;
                R ec = rerfc<R>(b * arr);
                R rinv = (R)1 / arr;
                R hlp1 = ec * rinv;
                R fr = rinv * rinv * rinv * (ec + (R)2 * sap * arr * rexp_<R>(-al * (R)r2));
                double qq = qi * qj;
                e += 0.5 * qq * (double)hlp1;
                ph += qj * (double)hlp1;
                double fs = qq * (double)fr;
                double h[3] = {fs * d[0], fs * d[1], fs * d[2]};
                acc[0] += h[0];
                acc[1] += h[1];
                acc[2] += h[2];
                w[0] -= 0.5 * d[0] * h[0];
                w[1] -= 0.5 * d[1] * h[1];
                w[2] -= 0.5 * d[2] * h[2];
                w[3] -= 0.5 * d[1] * h[2];
                w[4] -= 0.5 * d[0] * h[2];
                w[5] -= 0.5 * d[0] * h[1];
            }
        }
    }
    e = wsum(e);
    ph = wsum(ph);
    for (int x = 0; x < 3; ++x) acc[x] = wsum(acc[x]);
    for (int k = 0; k < 6; ++k) w[k] = wsum(w[k]);
    if (i >= 0 && lane == 0) {
        // self term (pme.f90:467-472; phi: pme.f90:256)
        if (fabs(qi) > 1e-10) ph -= 2.0 * qi * sqrt_alpha_pi;
        e -= qi * qi * sqrt_alpha_pi;
        if (f) {
            f[3 * (size_t)i] += acc[0];
            f[3 * (size_t)i + 1] += acc[1];
            f[3 * (size_t)i + 2] += acc[2];
        }
        if (phi) phi[i] += ph;
    }
    if (lane == 0) {
        s_red[wave][0] = e;
        for (int k = 0; k < 6; ++k) s_red[wave][1 + k] = w[k];
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        int qn = threadIdx.x;
        double v = ((s_red[0][qn] + s_red[1][qn]) + s_red[2][qn]) + s_red[3][qn];
        double *rec = partial + (size_t)blockIdx.x * 10;
        if (qn == 0) rec[0] = v;
        else if (qn == 1) rec[1] = v;
        else if (qn == 2) rec[5] = v;
        else if (qn == 3) rec[9] = v;
        else if (qn == 4) { rec[6] = v; rec[8] = v; }
        else if (qn == 5) { rec[3] = v; rec[7] = v; }
        else { rec[2] = v; rec[4] = v; }
    }
}

// ---------------------------------------------------------------------------
// reciprocal space
// ---------------------------------------------------------------------------
__global__ void k_frac(int nat, const double *__restrict__ r, CellArg B, double *__restrict__ s4)
{
    // s = Bbox r; B.A holds the column-major image of Bbox
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nat) return;
    double x = r[3 * j], y = r[3 * j + 1], z = r[3 * j + 2];
    for (int d = 0; d < 3; ++d) s4[4 * (size_t)j + d] = B.A[d] * x + B.A[3 + d] * y + B.A[6 + d] * z;
    s4[4 * (size_t)j + 3] = 0.0;
}

__device__ __forceinline__ double frac_turn(double t) { return t - rint(t); }

__global__ void __launch_bounds__(256)
k_ewald_sk(int nat, const double *__restrict__ s4, const double *__restrict__ q,
           const EwItem *__restrict__ items, int nitems, double *__restrict__ S)
{
    __shared__ double sh_s[EW_CHUNK][3];
    __shared__ double sh_q[EW_CHUNK];
    __shared__ double sh_e[EW_CHUNK][2];
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    EwItem I = {0, 0, 0, 0, 0, 0, 0, 0};
    if (it < nitems) I = items[it];
    double Sc[EW_L], Ss[EW_L];
#pragma unroll
    for (int l = 0; l < EW_L; ++l) Sc[l] = Ss[l] = 0.0;
    for (int base = 0; base < nat; base += EW_CHUNK) {
        int jj = base + threadIdx.x;
        __syncthreads();
        if (threadIdx.x < EW_CHUNK) {
            if (jj < nat) {
                double s1 = s4[4 * (size_t)jj], s2 = s4[4 * (size_t)jj + 1], s3 = s4[4 * (size_t)jj + 2];
                sh_s[threadIdx.x][0] = s1;
                sh_s[threadIdx.x][1] = s2;
                sh_s[threadIdx.x][2] = s3;
                sh_q[threadIdx.x] = q[jj];
                double sn, cs;
                sincospi(2.0 * frac_turn(s3), &sn, &cs);
                sh_e[threadIdx.x][0] = cs;
                sh_e[threadIdx.x][1] = sn;
            } else {
                sh_s[threadIdx.x][0] = sh_s[threadIdx.x][1] = sh_s[threadIdx.x][2] = 0.0;
                sh_q[threadIdx.x] = 0.0;
                sh_e[threadIdx.x][0] = 1.0;
                sh_e[threadIdx.x][1] = 0.0;
            }
        }
        __syncthreads();
        if (it < nitems) {
            const int m = min(EW_CHUNK, nat - base);
            for (int t = 0; t < m; ++t) {
                double ph = frac_turn((double)I.n1 * sh_s[t][0] + (double)I.n2 * sh_s[t][1]) +
                            frac_turn((double)I.n3s * sh_s[t][2]);
                double sn, cs;
                sincospi(2.0 * frac_turn(ph), &sn, &cs);
                const double qj = sh_q[t], ec = sh_e[t][0], es = sh_e[t][1];
                double c = qj * cs, s = qj * sn;
#pragma unroll
                for (int l = 0; l < EW_L; ++l) {
                    Sc[l] += c;
                    Ss[l] += s;
                    double c2 = c * ec - s * es;
                    s = c * es + s * ec;
                    c = c2;
                }
            }
        }
    }
    if (it < nitems)
        for (int l = 0; l < EW_L; ++l)
            if (l < I.cnt) {
                S[2 * (size_t)(I.kofs + l)] = Sc[l];
                S[2 * (size_t)(I.kofs + l) + 1] = Ss[l];
            }
}

__global__ void __launch_bounds__(256)
k_ewald_kenergy(int K, const int *__restrict__ kpack, CellArg Bt, double inv4b2, double pref,
                const double *__restrict__ S, double *__restrict__ T, double *__restrict__ partial)
{
    __shared__ double s_red[4][10];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double rec[10];
    for (int q = 0; q < 10; ++q) rec[q] = 0.0;
    if (k < K) {
        int n1, n2, n3;
        dc_unpack(kpack[k], n1, n2, n3);
        double kv[3];
        for (int x = 0; x < 3; ++x)
            kv[x] = 2.0 * M_PI * ((double)n1 * Bt.A[3 * x + 0] + (double)n2 * Bt.A[3 * x + 1] +
                                  (double)n3 * Bt.A[3 * x + 2]);
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double c = exp(-k2 * inv4b2) / k2;
        double sc = S[2 * (size_t)k], ss = S[2 * (size_t)k + 1];
        double s2 = sc * sc + ss * ss;
        T[2 * (size_t)k] = c * sc;
        T[2 * (size_t)k + 1] = c * ss;
        double ek = pref * c * s2;
        rec[0] = ek;
        double fac = 2.0 * (inv4b2 + 1.0 / k2);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) rec[1 + a + 3 * b] = ek * (fac * kv[a] * kv[b] - (a == b ? 1.0 : 0.0));
    }
    for (int q = 0; q < 10; ++q) rec[q] = wsum(rec[q]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 10; ++q) s_red[wave][q] = rec[q];
    __syncthreads();
    if (threadIdx.x < 10)
        partial[(size_t)blockIdx.x * 10 + threadIdx.x] =
            ((s_red[0][threadIdx.x] + s_red[1][threadIdx.x]) + s_red[2][threadIdx.x]) + s_red[3][threadIdx.x];
}

constexpr int EW_ICHUNK = 64;   // runs per LDS chunk in k_ewald_kforce

// Thread per ion: sum over the same runs of k as k_ewald_sk -- one sincospi
// per run, complex rotation by exp(2 pi i s3) along it -- of T(k) exp(-i k.r).
__global__ void __launch_bounds__(256)
k_ewald_kforce(int nat, int nitems, const double *__restrict__ s4, const double *__restrict__ q,
               const EwItem *__restrict__ items, const double *__restrict__ T, CellArg Bt, double pref2,
               double *__restrict__ f, double *__restrict__ phi)
{
    __shared__ int sh_i[EW_ICHUNK][5];
    __shared__ __align__(16) double sh_t[EW_ICHUNK * EW_L][2];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double s1 = 0, s2 = 0, s3 = 0, ec = 1, es = 0;
    if (j < nat) {
        s1 = s4[4 * (size_t)j];
        s2 = s4[4 * (size_t)j + 1];
        s3 = s4[4 * (size_t)j + 2];
        sincospi(2.0 * frac_turn(s3), &es, &ec);
    }
    double a1 = 0, a2 = 0, a3 = 0, ap = 0;
    for (int ib = 0; ib < nitems; ib += EW_ICHUNK) {
        const int mi = min(EW_ICHUNK, nitems - ib);
        const int k0 = items[ib].kofs;
        const int nk = items[ib + mi - 1].kofs + items[ib + mi - 1].cnt - k0;
        __syncthreads();
        if ((int)threadIdx.x < mi) {
            const EwItem I = items[ib + threadIdx.x];
            sh_i[threadIdx.x][0] = I.n1;
            sh_i[threadIdx.x][1] = I.n2;
            sh_i[threadIdx.x][2] = I.n3s;
            sh_i[threadIdx.x][3] = I.cnt;
            sh_i[threadIdx.x][4] = I.kofs - k0;
        }
        for (int t = threadIdx.x; t < nk; t += blockDim.x) {
            sh_t[t][0] = T[2 * (size_t)(k0 + t)];
            sh_t[t][1] = T[2 * (size_t)(k0 + t) + 1];
        }
        __syncthreads();
        for (int it = 0; it < mi; ++it) {
            const int n1 = sh_i[it][0], n2 = sh_i[it][1], n3s = sh_i[it][2], cnt = sh_i[it][3];
            const int ko = sh_i[it][4];
            const double ph = frac_turn((double)n1 * s1 + (double)n2 * s2) + frac_turn((double)n3s * s3);
            double c, s;
            sincospi(2.0 * frac_turn(ph), &s, &c);
            // the run's terms, unrolled over the run length (EW_L, shorter
            // runs predicated): sim = sum im, lim = sum l im, so that the
            // n3 moment sum_l (n3s + l) im is n3s sim + lim
            double sim = 0.0, lim = 0.0;
#pragma unroll
            for (int l = 0; l < EW_L; ++l) {
                if (l < cnt) {
                    const double2 tv = *reinterpret_cast<const double2 *>(sh_t[ko + l]);
                    const double im = s * tv.x - c * tv.y;
                    ap += c * tv.x + s * tv.y;
                    sim += im;
                    lim += (double)l * im;
                    const double c2 = c * ec - s * es;
                    s = c * es + s * ec;
                    c = c2;
                }
            }
            const double s3im = (double)n3s * sim + lim;
            a1 += (double)n1 * sim;
            a2 += (double)n2 * sim;
            a3 += s3im;
        }
    }
    if (j < nat) {
        const double qj = q[j];
        // k = 2 pi sum_d n_d b_d, b_d = row d of Bbox
        if (f)
            for (int x = 0; x < 3; ++x) {
                double kx = 2.0 * M_PI * (a1 * Bt.A[3 * x + 0] + a2 * Bt.A[3 * x + 1] + a3 * Bt.A[3 * x + 2]);
                f[3 * (size_t)j + x] += pref2 * qj * kx;
            }
        if (phi) phi[j] += pref2 * ap;
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static double solve_beta(double rc, double tol)
{
    double lo = 0.0, hi = 10.0 / rc;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (std::erfc(mid * rc) > tol) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

static int build_kvectors(EwaldState &st, const agpu_particles *p)
{
    // Bt.A[3*x + d] = Bbox(d, x): row d of Bbox is the reciprocal vector b_d / (2 pi)
    const double *B = p->cell.B;   // B[3*k + x] = Bbox(x, k)
    double b[3][3];
    for (int d = 0; d < 3; ++d)
        for (int x = 0; x < 3; ++x) b[d][x] = 2.0 * M_PI * B[3 * x + d];
    const double kmax2 = st.kmax * st.kmax;
    int nmax[3];
    for (int d = 0; d < 3; ++d) {
        const double *a = &p->cell.A[3 * d];
        double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        nmax[d] = (int)std::floor(st.kmax * len / (2.0 * M_PI)) + 1;
        if (nmax[d] > 500) return set_error(AGPU_ERR_ARG, "Ewald: more than 500 k-vectors per axis");
    }
    st.items.clear();
    st.kpack.clear();
    for (int n1 = 0; n1 <= nmax[0]; ++n1)
        for (int n2 = (n1 == 0 ? 0 : -nmax[1]); n2 <= nmax[1]; ++n2) {
            int lo = (n1 == 0 && n2 == 0) ? 1 : -nmax[2];
            int run_start = 0, run_len = 0;
            auto flush = [&]() {
                for (int o = 0; o < run_len; o += EW_L) {
                    EwItem it{};
                    it.n1 = n1;
                    it.n2 = n2;
                    it.n3s = run_start + o;
                    it.cnt = std::min(EW_L, run_len - o);
                    it.kofs = (int)st.kpack.size() - run_len + o;
                    st.items.push_back(it);
                }
                run_len = 0;
            };
            for (int n3 = lo; n3 <= nmax[2]; ++n3) {
                double kv[3];
                for (int x = 0; x < 3; ++x) kv[x] = n1 * b[0][x] + n2 * b[1][x] + n3 * b[2][x];
                double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                if (k2 <= kmax2) {
                    if (run_len == 0) run_start = n3;
                    st.kpack.push_back(((n1 + 512) << 20) | ((n2 + 512) << 10) | (n3 + 512));
                    ++run_len;
                } else if (run_len) {
                    flush();
                }
            }
            if (run_len) flush();
        }
    st.K = (int)st.kpack.size();
    st.nitems = (int)st.items.size();
    AGPU_CHECK(st.d_items.reserve(sizeof(EwItem) * (st.nitems + 1)));
    AGPU_CHECK(st.d_kpack.reserve(sizeof(int) * (st.K + 1)));
    AGPU_CHECK(st.d_S.reserve(sizeof(double) * 2 * (st.K + 1)));
    AGPU_CHECK(st.d_T.reserve(sizeof(double) * 2 * (st.K + 1)));
    hipStream_t s = stream();
    if (st.nitems)
        AGPU_HIP(hipMemcpyAsync(st.d_items.ptr, st.items.data(), sizeof(EwItem) * st.nitems, hipMemcpyHostToDevice, s));
    if (st.K) AGPU_HIP(hipMemcpyAsync(st.d_kpack.ptr, st.kpack.data(), sizeof(int) * st.K, hipMemcpyHostToDevice, s));
    AGPU_HIP(hipStreamSynchronize(s));
    std::memcpy(st.cell_key, p->cell.A, sizeof st.cell_key);
    st.have_k = true;
    return AGPU_OK;
}

int ewald_bind(agpu_potential *pot)
{
    EwaldState *st = static_cast<EwaldState *>(pot->impl);
    agpu_particles *p = pot->p;
    if (!(p->cell.pbc[0] && p->cell.pbc[1] && p->cell.pbc[2]))
        return set_error(AGPU_ERR_PBC, "Ewald summation only works for 3d periodic systems.");
    // beta <= 0: the reference's rule alpha = ln(10)*12/cutoff^2, beta = sqrt(alpha)
    st->beta = st->beta_req > 0.0 ? st->beta_req : std::sqrt(std::log(10.0) * 12.0 / (st->cutoff * st->cutoff));
    st->kmax = st->kmax_req > 0.0 ? st->kmax_req : 2.0 * st->beta * std::sqrt(-std::log(1e-13));
    AGPU_CHECK(agpu_neighbors_request_interaction_range(pot->nl, st->cutoff));
    st->have_k = false;
    return AGPU_OK;
}

static int ewald_run(agpu_potential *pot, const double *q, double *epot, double *f, double *wpot,
                     double *phi)
{
    EwaldState *st = static_cast<EwaldState *>(pot->impl);
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    AGPU_CHECK(nl_update(nl, p, 0, false));
    if (!st->pme && (!st->have_k || std::memcmp(st->cell_key, p->cell.A, sizeof st->cell_key) != 0))
        AGPU_CHECK(build_kvectors(*st, p));
    const int nat = p->nat;
    hipStream_t s = stream();
    if (epot) AGPU_HIP(hipMemsetAsync(epot, 0, sizeof(double), s));
    if (nat == 0) return AGPU_OK;
    CellArg C, Bt;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    std::memcpy(Bt.A, p->cell.B, sizeof Bt.A);
    const double sap = std::sqrt(st->beta * st->beta / M_PI);
    const int nblk_r = (nat + 3) / 4;
    const int nblk_k = (st->K + 255) / 256;
    AGPU_CHECK(st->partial.reserve(sizeof(double) * 10 * (size_t)(nblk_r + nblk_k + 1)));
    double *part = st->partial.as<double>();
    const int *order = nl->idx_sorted.as<int>();
    const NlRows L = nl_rows(nl);
    AGPU_CHECK(st->q4.reserve(sizeof(double4) * (size_t)(nat + 1)));
    double4 *q4 = st->q4.as<double4>();
    AGPU_LAUNCH("ewald_pack", k_ewald_pack, dim3((nat + 255) / 256), dim3(256), 0, s, nat, L.atoms, order, q, q4);
    if (pot->precision == AGPU_PREC_MIXED)
        AGPU_LAUNCH("ewald_real_mixed", k_ewald_real<float>, dim3(nblk_r), dim3(256), 0, s, nat, order, q4, C, L,
                    st->cutoff * st->cutoff, st->beta, sap, f, phi, part);
    else
        AGPU_LAUNCH("ewald_real_fp64", k_ewald_real<double>, dim3(nblk_r), dim3(256), 0, s, nat, order, q4, C, L,
                    st->cutoff * st->cutoff, st->beta, sap, f, phi, part);
    if (st->pme) {
        // reciprocal record after the real-space block records
        AGPU_CHECK(pme_recip(st->pme, p, q, st->beta, part + 10 * (size_t)nblk_r, f, phi));
        AGPU_CHECK(reduce_ev(part, nblk_r + 1, 10, epot, wpot, nullptr));
        AGPU_HIP(hipStreamSynchronize(s));
        return AGPU_OK;
    }
    AGPU_CHECK(st->d_frac.reserve(sizeof(double) * 4 * (size_t)(nat + 1)));
    AGPU_LAUNCH("ewald_frac", k_frac, dim3((nat + 255) / 256), dim3(256), 0, s, nat, p->d_r, Bt,
                st->d_frac.as<double>());
    double V = std::fabs(C.A[0] * (C.A[4] * C.A[8] - C.A[5] * C.A[7]) - C.A[1] * (C.A[3] * C.A[8] - C.A[5] * C.A[6]) +
                         C.A[2] * (C.A[3] * C.A[7] - C.A[4] * C.A[6]));
    if (st->K > 0) {
        AGPU_LAUNCH("ewald_sk", k_ewald_sk, dim3((st->nitems + 255) / 256), dim3(256), 0, s, nat,
                    st->d_frac.as<double>(), q, st->d_items.as<EwItem>(), st->nitems, st->d_S.as<double>());
        AGPU_LAUNCH("ewald_kenergy", k_ewald_kenergy, dim3(nblk_k), dim3(256), 0, s, st->K, st->d_kpack.as<int>(),
                    Bt, 1.0 / (4.0 * st->beta * st->beta), 4.0 * M_PI / V, st->d_S.as<double>(),
                    st->d_T.as<double>(), part + 10 * (size_t)nblk_r);
        AGPU_LAUNCH("ewald_kforce", k_ewald_kforce, dim3((nat + 255) / 256), dim3(256), 0, s, nat, st->nitems,
                    st->d_frac.as<double>(), q, st->d_items.as<EwItem>(), st->d_T.as<double>(), Bt, 8.0 * M_PI / V,
                    f, phi);
    }
    AGPU_CHECK(reduce_ev(part, nblk_r + (st->K > 0 ? nblk_k : 0), 10, epot, wpot, nullptr));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

int ewald_compute(agpu_potential *pot, const agpu_eval_args *a)
{
    if (a->mask || a->epot_per_at || a->wpot_per_at)
        return set_error(AGPU_ERR_UNSUPPORTED, "Ewald: mask and per-atom properties are not supported");
    return ewald_run(pot, a->q, a->epot, a->f, a->wpot, nullptr);
}

int ewald_potential(agpu_potential *pot, const double *q, double *phi, int)
{
    EwaldState *st = static_cast<EwaldState *>(pot->impl);
    AGPU_CHECK(st->phi_tmp.reserve(sizeof(double) * 16));
    double *tmp = st->phi_tmp.as<double>();   // [0] epot, [1..9] wpot scratch
    AGPU_HIP(hipMemsetAsync(tmp, 0, sizeof(double) * 16, stream()));
    return ewald_run(pot, q, tmp, nullptr, tmp + 1, phi);
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_ewald_create(agpu_potential **pot, double cutoff, double beta, double kmax)
{
    if (!pot || !(cutoff > 0.0)) return set_error(AGPU_ERR_ARG, "ewald_create: bad arguments");
    EwaldState *st = new EwaldState();
    st->cutoff = cutoff;
    st->beta_req = beta;
    st->kmax_req = kmax;
    agpu_potential *p = new agpu_potential();
    p->kind = AGPU_KIND_EWALD;
    p->cutoff = cutoff;
    p->impl = st;
    p->free_impl = ewald_free;
    *pot = p;
    return AGPU_OK;
}

extern "C" int agpu_pme_create(agpu_potential **pot, double cutoff, int order, const int *grid)
{
    if (!pot) return set_error(AGPU_ERR_ARG, "pme_create: pot is NULL");
    // pme_t defaults (pme.f90:36-46): cutoff 10, grid 64^3, order 8
    if (!(cutoff > 0.0)) cutoff = 10.0;
    if (order <= 0) order = 8;
    int K[3] = {64, 64, 64};
    if (grid)
        for (int d = 0; d < 3; ++d) K[d] = grid[d];
    if (order < 3 || order > 12) return set_error(AGPU_ERR_ARG, "PME: order must be 3 .. 12");
    for (int d = 0; d < 3; ++d)
        if (K[d] < order || K[d] > 1000)
            return set_error(AGPU_ERR_ARG, "PME: every grid dimension must be >= order and <= 1000");
    if ((long long)K[0] * K[1] * K[2] > (1ll << 30)) return set_error(AGPU_ERR_ARG, "PME: grid too large");
    EwaldState *st = new EwaldState();
    st->cutoff = cutoff;
    st->pme = pme_state_new(order, K);
    agpu_potential *p = new agpu_potential();
    p->kind = AGPU_KIND_PME;
    p->cutoff = cutoff;
    p->impl = st;
    p->free_impl = ewald_free;
    *pot = p;
    return AGPU_OK;
}
