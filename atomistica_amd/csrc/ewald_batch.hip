// Batched Ewald Coulomb sum (e^2/Angstrom) over many small periodic cells of
// uneven size, one launch per pass for the whole batch.
//
// Per structure, the same sums as the single-structure Ewald path (ewald.hip;
// real space and self term of pme.f90:215-321, 386-437, 467-472; reciprocal
// space as the direct structure-factor sum over the half space of
// k = 2 pi Bbox^T n, |k| <= kmax, S(k) accumulated in FP64):
//   k_ewb_real<R>  one wavefront per tile of <= 64 ions of one structure:
//                  every ion of the structure is staged through LDS and every
//                  periodic image within the cut-off is summed (the image
//                  range per axis comes from the reciprocal vectors), so no
//                  neighbour list is stored.  erfc/exp in R (double, or float
//                  with FP64 distances and FP64 accumulation).
//   k_ewb_frac     fractional coordinates s = Bbox r per ion.
//   k_ewb_sk       one wavefront per tile of <= 64 runs of one structure
//                  (a run = up to 16 consecutive n3 in one (n1, n2) column);
//                  ions staged through LDS; one sincospi per run and ion, the
//                  rest of the run by complex rotation.
//   k_ewb_kenergy  thread per k: c(k) = exp(-k^2/4beta^2)/k^2, T(k) = c S(k),
//                  energy and virial record of k.
//   k_ewb_kforce   one wavefront per ion tile: sum over the structure's runs
//                  of T(k) exp(-i k.r) (forces and potential per ion).
//   k_ewb_sum      one wavefront per structure: energy and virial from the
//                  per-ion and per-k records in a fixed order.
// There is no counterpart of batching in the reference, which evaluates one
// structure per call.  No atomics anywhere: every output has one writer.
#include <cmath>
#include <cstring>
#include <vector>

#include "agpu_internal.hpp"
#include "batch_internal.hpp"
#include "profile.hpp"

namespace agpu {

constexpr int EWB_L = 16;      // k per run
constexpr int EWB_T = 64;      // ions (real space, forces) or runs (S(k)) per wavefront tile

struct EwbRun {
    int s, n1, n2, n3s, cnt, kofs, pad0, pad1;
};

struct EwbTile {
    int s, b, e, pad;      // structure, [b, e): ions or runs
};

struct EwbState {
    double beta = 0.0, kmax = 0.0;
    uint64_t struct_rev = ~0ull;
    int K = 0;
    std::vector<EwbRun> runs;
    std::vector<EwbTile> itiles, rtiles;
    std::vector<int> koff, kstruct;
    DevBuf d_runs, d_itiles, d_rtiles, d_koff, d_kstruct, d_kpack, d_S, d_T, d_s4, d_ionrec, d_krec, d_q;
};

static void ewb_free(void *p) { delete static_cast<EwbState *>(p); }

// record layout (7 doubles): e, w_xx, w_yy, w_zz, w_yz, w_xz, w_xy
constexpr int EWB_REC = 7;

template <typename R> __device__ __forceinline__ R ewb_erfc(R x);
template <> __device__ __forceinline__ double ewb_erfc<double>(double x) { return erfc(x); }
template <> __device__ __forceinline__ float ewb_erfc<float>(float x) { return erfcf(x); }
template <typename R> __device__ __forceinline__ R ewb_exp(R x);
template <> __device__ __forceinline__ double ewb_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float ewb_exp<float>(float x) { return expf(x); }

struct EwbGeom {
    const int *off;        // atom offsets per structure
    const double *cell;    // 9 per structure, cell[3k+x] = a_k[x]
    const double *rec;     // 9 per structure, rec[3k+x] = b_k[x], a_k . b_l = delta_kl
};

template <typename R>
__global__ void __launch_bounds__(EWB_T)
k_ewb_real(const EwbTile *__restrict__ tiles, EwbGeom G, const double *__restrict__ r,
           const double *__restrict__ q, double cutoff, double beta, double *__restrict__ f,
           double *__restrict__ phi, double *__restrict__ ionrec)
{
    __shared__ double sh_r[EWB_T][3];
    __shared__ double sh_q[EWB_T];
    const EwbTile T = tiles[blockIdx.x];
    const int lane = threadIdx.x;
    const int i = T.b + lane;
    const bool act = i < T.e;
    double A[9], b[9], t[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        A[k] = G.cell[9 * (size_t)T.s + k];
        b[k] = G.rec[9 * (size_t)T.s + k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        t[k] = cutoff * sqrt(b[3 * k] * b[3 * k] + b[3 * k + 1] * b[3 * k + 1] + b[3 * k + 2] * b[3 * k + 2]);
    const double rc2 = cutoff * cutoff;
    const double sap = beta / sqrt(M_PI);
    const R bR = (R)beta, sapR = (R)sap, alR = (R)(beta * beta);
    double ri[3] = {0.0, 0.0, 0.0}, qi = 0.0;
    if (act) {
        ri[0] = r[3 * (size_t)i];
        ri[1] = r[3 * (size_t)i + 1];
        ri[2] = r[3 * (size_t)i + 2];
        qi = q[i];
    }
    const bool live = act && fabs(qi) > 1e-10;
    double e = 0.0, ph = 0.0, acc[3] = {0.0, 0.0, 0.0}, w[6] = {0, 0, 0, 0, 0, 0};
    const int a0 = G.off[T.s], a1 = G.off[T.s + 1];
    for (int base = a0; base < a1; base += EWB_T) {
        const int jj = base + lane;
        __syncthreads();
        if (jj < a1) {
            sh_r[lane][0] = r[3 * (size_t)jj];
            sh_r[lane][1] = r[3 * (size_t)jj + 1];
            sh_r[lane][2] = r[3 * (size_t)jj + 2];
            sh_q[lane] = q[jj];
        }
        __syncthreads();
        if (!live) continue;
        const int m = min(EWB_T, a1 - base);
        for (int tt = 0; tt < m; ++tt) {
            const double qj = sh_q[tt];
            if (!(fabs(qj) > 1e-10)) continue;
            const int j = base + tt;
            const double d0[3] = {ri[0] - sh_r[tt][0], ri[1] - sh_r[tt][1], ri[2] - sh_r[tt][2]};
            int lo[3], hi[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double ds = b[3 * k] * d0[0] + b[3 * k + 1] * d0[1] + b[3 * k + 2] * d0[2];
                lo[k] = (int)ceil(-ds - t[k] - 1e-9);
                hi[k] = (int)floor(-ds + t[k] + 1e-9);
            }
            for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
                for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
                    for (int n2 = lo[2]; n2 <= hi[2]; ++n2) {
                        if (j == i && n0 == 0 && n1 == 0 && n2 == 0) continue;
                        double d[3];
#pragma unroll
                        for (int x = 0; x < 3; ++x)
                            d[x] = d0[x] + (A[x] * n0 + A[3 + x] * n1 + A[6 + x] * n2);
                        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (!(r2 < rc2)) continue;
                        const double ar = sqrt(r2);
                        const R arr = (R)ar;
                        const R ec = ewb_erfc<R>(bR * arr);
                        const R rinv = (R)1 / arr;
                        const R hlp = ec * rinv;
                        const R fr = rinv * rinv * rinv * (ec + (R)2 * sapR * arr * ewb_exp<R>(-alR * (R)r2));
                        const double qq = qi * qj;
                        e += 0.5 * qq * (double)hlp;
                        ph += qj * (double)hlp;
                        const double fs = qq * (double)fr;
                        const double h[3] = {fs * d[0], fs * d[1], fs * d[2]};
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
    if (!act) return;
    // self term (pme.f90:467-472; phi: pme.f90:256)
    if (live) ph -= 2.0 * qi * sap;
    e -= qi * qi * sap;
    if (f) {
        f[3 * (size_t)i] += acc[0];
        f[3 * (size_t)i + 1] += acc[1];
        f[3 * (size_t)i + 2] += acc[2];
    }
    if (phi) phi[i] += ph;
    double *rec = ionrec + EWB_REC * (size_t)i;
    rec[0] = e;
#pragma unroll
    for (int k = 0; k < 6; ++k) rec[1 + k] = w[k];
}

__device__ __forceinline__ double ewb_turn(double t) { return t - rint(t); }

// s = Bbox r per ion: s_k = b_k . r
__global__ void k_ewb_frac(int n, const int *__restrict__ sof, const double *__restrict__ rec,
                           const double *__restrict__ r, double4 *__restrict__ s4)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double *b = rec + 9 * (size_t)sof[j];
    const double x = r[3 * (size_t)j], y = r[3 * (size_t)j + 1], z = r[3 * (size_t)j + 2];
    s4[j] = make_double4(b[0] * x + b[1] * y + b[2] * z, b[3] * x + b[4] * y + b[5] * z,
                         b[6] * x + b[7] * y + b[8] * z, 0.0);
}

__global__ void __launch_bounds__(EWB_T)
k_ewb_sk(const EwbTile *__restrict__ tiles, const EwbRun *__restrict__ runs, const int *__restrict__ off,
         const double4 *__restrict__ s4, const double *__restrict__ q, double *__restrict__ S)
{
    __shared__ double sh_s[EWB_T][3];
    __shared__ double sh_q[EWB_T];
    __shared__ double sh_e[EWB_T][2];
    const EwbTile T = tiles[blockIdx.x];
    const int lane = threadIdx.x;
    const int it = T.b + lane;
    const bool act = it < T.e;
    EwbRun I{0, 0, 0, 0, 0, 0, 0, 0};
    if (act) I = runs[it];
    double Sc[EWB_L], Ss[EWB_L];
#pragma unroll
    for (int l = 0; l < EWB_L; ++l) Sc[l] = Ss[l] = 0.0;
    const int a0 = off[T.s], a1 = off[T.s + 1];
    for (int base = a0; base < a1; base += EWB_T) {
        const int jj = base + lane;
        __syncthreads();
        if (jj < a1) {
            const double4 sj = s4[jj];
            sh_s[lane][0] = sj.x;
            sh_s[lane][1] = sj.y;
            sh_s[lane][2] = sj.z;
            sh_q[lane] = q[jj];
            double sn, cs;
            sincospi(2.0 * ewb_turn(sj.z), &sn, &cs);
            sh_e[lane][0] = cs;
            sh_e[lane][1] = sn;
        }
        __syncthreads();
        if (!act) continue;
        const int m = min(EWB_T, a1 - base);
        for (int tt = 0; tt < m; ++tt) {
            const double ph = ewb_turn((double)I.n1 * sh_s[tt][0] + (double)I.n2 * sh_s[tt][1]) +
                              ewb_turn((double)I.n3s * sh_s[tt][2]);
            double sn, cs;
            sincospi(2.0 * ewb_turn(ph), &sn, &cs);
            const double qj = sh_q[tt], ec = sh_e[tt][0], es = sh_e[tt][1];
            double c = qj * cs, s = qj * sn;
#pragma unroll
            for (int l = 0; l < EWB_L; ++l) {
                Sc[l] += c;
                Ss[l] += s;
                const double c2 = c * ec - s * es;
                s = c * es + s * ec;
                c = c2;
            }
        }
    }
    if (act)
        for (int l = 0; l < EWB_L; ++l)
            if (l < I.cnt) {
                S[2 * (size_t)(I.kofs + l)] = Sc[l];
                S[2 * (size_t)(I.kofs + l) + 1] = Ss[l];
            }
}

__global__ void __launch_bounds__(256)
k_ewb_kenergy(int K, const int *__restrict__ kpack, const int *__restrict__ kstruct, const double *__restrict__ rec,
              const double *__restrict__ cell, double inv4b2, const double *__restrict__ S, double *__restrict__ T,
              double *__restrict__ krec)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int s = kstruct[k];
    const double *b = rec + 9 * (size_t)s, *A = cell + 9 * (size_t)s;
    int n1, n2, n3;
    dc_unpack(kpack[k], n1, n2, n3);
    double kv[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) kv[x] = 2.0 * M_PI * ((double)n1 * b[x] + (double)n2 * b[3 + x] + (double)n3 * b[6 + x]);
    const double V = fabs(A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                          A[2] * (A[3] * A[7] - A[4] * A[6]));
    const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    const double c = exp(-k2 * inv4b2) / k2;
    const double sc = S[2 * (size_t)k], ss = S[2 * (size_t)k + 1];
    T[2 * (size_t)k] = c * sc;
    T[2 * (size_t)k + 1] = c * ss;
    const double ek = (4.0 * M_PI / V) * c * (sc * sc + ss * ss);
    const double fac = 2.0 * (inv4b2 + 1.0 / k2);
    double *o = krec + EWB_REC * (size_t)k;
    o[0] = ek;
    o[1] = ek * (fac * kv[0] * kv[0] - 1.0);
    o[2] = ek * (fac * kv[1] * kv[1] - 1.0);
    o[3] = ek * (fac * kv[2] * kv[2] - 1.0);
    o[4] = ek * fac * kv[1] * kv[2];
    o[5] = ek * fac * kv[0] * kv[2];
    o[6] = ek * fac * kv[0] * kv[1];
}

constexpr int EWB_RCHUNK = 16;   // runs staged per LDS chunk in k_ewb_kforce

__global__ void __launch_bounds__(EWB_T)
k_ewb_kforce(const EwbTile *__restrict__ itiles, const int *__restrict__ rfirst, const EwbRun *__restrict__ runs,
             const double4 *__restrict__ s4, const double *__restrict__ q, const double *__restrict__ T,
             const double *__restrict__ rec, const double *__restrict__ cell, double *__restrict__ f,
             double *__restrict__ phi)
{
    __shared__ int sh_i[EWB_RCHUNK][5];
    __shared__ __align__(16) double sh_t[EWB_RCHUNK * EWB_L][2];
    const EwbTile Ti = itiles[blockIdx.x];
    const int lane = threadIdx.x;
    const int j = Ti.b + lane;
    const bool act = j < Ti.e;
    double s1 = 0, s2 = 0, s3 = 0, ec = 1, es = 0;
    if (act) {
        const double4 sj = s4[j];
        s1 = sj.x;
        s2 = sj.y;
        s3 = sj.z;
        sincospi(2.0 * ewb_turn(s3), &es, &ec);
    }
    const int r0 = rfirst[Ti.s], r1 = rfirst[Ti.s + 1];
    double a1 = 0, a2 = 0, a3 = 0, ap = 0;
    for (int ib = r0; ib < r1; ib += EWB_RCHUNK) {
        const int mi = min(EWB_RCHUNK, r1 - ib);
        const int k0 = runs[ib].kofs;
        const int nk = runs[ib + mi - 1].kofs + runs[ib + mi - 1].cnt - k0;
        __syncthreads();
        if (lane < mi) {
            const EwbRun I = runs[ib + lane];
            sh_i[lane][0] = I.n1;
            sh_i[lane][1] = I.n2;
            sh_i[lane][2] = I.n3s;
            sh_i[lane][3] = I.cnt;
            sh_i[lane][4] = I.kofs - k0;
        }
        for (int t = lane; t < nk; t += EWB_T) {
            sh_t[t][0] = T[2 * (size_t)(k0 + t)];
            sh_t[t][1] = T[2 * (size_t)(k0 + t) + 1];
        }
        __syncthreads();
        if (!act) continue;
        for (int it = 0; it < mi; ++it) {
            const int n1 = sh_i[it][0], n2 = sh_i[it][1], n3s = sh_i[it][2], cnt = sh_i[it][3];
            const int ko = sh_i[it][4];
            const double ph = ewb_turn((double)n1 * s1 + (double)n2 * s2) + ewb_turn((double)n3s * s3);
            double c, s;
            sincospi(2.0 * ewb_turn(ph), &s, &c);
            // unrolled over the run length (shorter runs predicated):
            // sum_l (n3s + l) im = n3s sim + lim
            double sim = 0.0, lim = 0.0;
#pragma unroll
            for (int l = 0; l < EWB_L; ++l) {
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
    if (!act) return;
    const double *A = cell + 9 * (size_t)Ti.s, *b = rec + 9 * (size_t)Ti.s;
    const double V = fabs(A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                          A[2] * (A[3] * A[7] - A[4] * A[6]));
    const double pref2 = 8.0 * M_PI / V;
    if (f) {
        const double qj = q[j];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            // k = 2 pi sum_d n_d b_d
            const double kx = 2.0 * M_PI * (a1 * b[x] + a2 * b[3 + x] + a3 * b[6 + x]);
            f[3 * (size_t)j + x] += pref2 * qj * kx;
        }
    }
    if (phi) phi[j] += pref2 * ap;
}

// energy and virial per structure: one wavefront per structure, lane-strided
// partial sums of the ion and k records, then a fixed butterfly
__global__ void __launch_bounds__(256)
k_ewb_sum(int nstruct, const int *__restrict__ off, const int *__restrict__ koff,
          const double *__restrict__ ionrec, const double *__restrict__ krec, double *__restrict__ es,
          double *__restrict__ ws)
{
    const int s = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x & 63;
    if (s >= nstruct) return;
    double v[EWB_REC];
#pragma unroll
    for (int c = 0; c < EWB_REC; ++c) v[c] = 0.0;
    for (int a = off[s] + lane; a < off[s + 1]; a += 64)
#pragma unroll
        for (int c = 0; c < EWB_REC; ++c) v[c] += ionrec[EWB_REC * (size_t)a + c];
    for (int k = koff[s] + lane; k < koff[s + 1]; k += 64)
#pragma unroll
        for (int c = 0; c < EWB_REC; ++c) v[c] += krec[EWB_REC * (size_t)k + c];
#pragma unroll
    for (int c = 0; c < EWB_REC; ++c) v[c] = group_sum<64>(v[c]);
    if (lane != 0) return;
    es[s] = v[0];
    if (ws) {
        double *w = ws + 9 * (size_t)s;
        w[0] = v[1];
        w[4] = v[2];
        w[8] = v[3];
        w[5] = w[7] = v[4];
        w[2] = w[6] = v[5];
        w[1] = w[3] = v[6];
    }
}

// k-vector runs of every structure (half space, as the single-structure path)
static int ewb_kvectors(agpu_batch *b, EwbState &st)
{
    const int ns = b->nstruct;
    st.runs.clear();
    st.rtiles.clear();
    st.koff.assign(ns + 1, 0);
    st.kstruct.clear();
    std::vector<int> kpack, rfirst(ns + 1, 0);
    const double kmax2 = st.kmax * st.kmax;
    for (int s = 0; s < ns; ++s) {
        const double *A = &b->h_cell[9 * (size_t)s];
        double B[9], bv[3][3];
        if (compute_inverse(A, B) != AGPU_OK) return set_error(AGPU_ERR_ARG, "Ewald batch: singular cell");
        // B[3*x + d] = (A^-1)(d, x)... rows of A^-T: b_d[x] = B[3*x + d]
        for (int d = 0; d < 3; ++d)
            for (int x = 0; x < 3; ++x) bv[d][x] = 2.0 * M_PI * B[3 * x + d];
        int nmax[3];
        for (int d = 0; d < 3; ++d) {
            const double *a = A + 3 * d;
            const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            nmax[d] = (int)std::floor(st.kmax * len / (2.0 * M_PI)) + 1;
            if (nmax[d] > 500) return set_error(AGPU_ERR_ARG, "Ewald batch: more than 500 k-vectors per axis");
        }
        rfirst[s] = (int)st.runs.size();
        st.koff[s] = (int)kpack.size();
        for (int n1 = 0; n1 <= nmax[0]; ++n1)
            for (int n2 = (n1 == 0 ? 0 : -nmax[1]); n2 <= nmax[1]; ++n2) {
                const int lo = (n1 == 0 && n2 == 0) ? 1 : -nmax[2];
                int run_start = 0, run_len = 0;
                auto flush = [&]() {
                    for (int o = 0; o < run_len; o += EWB_L) {
                        EwbRun it{};
                        it.s = s;
                        it.n1 = n1;
                        it.n2 = n2;
                        it.n3s = run_start + o;
                        it.cnt = std::min(EWB_L, run_len - o);
                        it.kofs = (int)kpack.size() - run_len + o;
                        st.runs.push_back(it);
                    }
                    run_len = 0;
                };
                for (int n3 = lo; n3 <= nmax[2]; ++n3) {
                    double kv[3];
                    for (int x = 0; x < 3; ++x) kv[x] = n1 * bv[0][x] + n2 * bv[1][x] + n3 * bv[2][x];
                    if (kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2] <= kmax2) {
                        if (run_len == 0) run_start = n3;
                        kpack.push_back(((n1 + 512) << 20) | ((n2 + 512) << 10) | (n3 + 512));
                        st.kstruct.push_back(s);
                        ++run_len;
                    } else if (run_len) {
                        flush();
                    }
                }
                if (run_len) flush();
            }
        for (int t = rfirst[s]; t < (int)st.runs.size(); t += EWB_T)
            st.rtiles.push_back(EwbTile{s, t, std::min(t + EWB_T, (int)st.runs.size()), 0});
    }
    rfirst[ns] = (int)st.runs.size();
    st.koff[ns] = (int)kpack.size();
    st.K = (int)kpack.size();
    hipStream_t hs = stream();
    AGPU_CHECK(st.d_runs.reserve(sizeof(EwbRun) * (st.runs.size() + 1)));
    AGPU_CHECK(st.d_rtiles.reserve(sizeof(EwbTile) * (st.rtiles.size() + 1)));
    AGPU_CHECK(st.d_koff.reserve(sizeof(int) * 2 * (ns + 1)));
    AGPU_CHECK(st.d_kpack.reserve(sizeof(int) * (st.K + 1)));
    AGPU_CHECK(st.d_kstruct.reserve(sizeof(int) * (st.K + 1)));
    AGPU_CHECK(st.d_S.reserve(sizeof(double) * 2 * (st.K + 1)));
    AGPU_CHECK(st.d_T.reserve(sizeof(double) * 2 * (st.K + 1)));
    AGPU_CHECK(st.d_krec.reserve(sizeof(double) * EWB_REC * (st.K + 1)));
    if (!st.runs.empty())
        AGPU_HIP(hipMemcpyAsync(st.d_runs.ptr, st.runs.data(), sizeof(EwbRun) * st.runs.size(), hipMemcpyHostToDevice, hs));
    if (!st.rtiles.empty())
        AGPU_HIP(hipMemcpyAsync(st.d_rtiles.ptr, st.rtiles.data(), sizeof(EwbTile) * st.rtiles.size(),
                                hipMemcpyHostToDevice, hs));
    // d_koff: [koff (ns + 1) | rfirst (ns + 1)]
    AGPU_HIP(hipMemcpyAsync(st.d_koff.ptr, st.koff.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice, hs));
    AGPU_HIP(hipMemcpyAsync(st.d_koff.as<int>() + ns + 1, rfirst.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice,
                            hs));
    if (st.K) {
        AGPU_HIP(hipMemcpyAsync(st.d_kpack.ptr, kpack.data(), sizeof(int) * st.K, hipMemcpyHostToDevice, hs));
        AGPU_HIP(hipMemcpyAsync(st.d_kstruct.ptr, st.kstruct.data(), sizeof(int) * st.K, hipMemcpyHostToDevice, hs));
    }
    AGPU_HIP(hipStreamSynchronize(hs));
    return AGPU_OK;
}

static int ewb_ion_tiles(agpu_batch *b, EwbState &st)
{
    st.itiles.clear();
    for (int s = 0; s < b->nstruct; ++s)
        for (int a = b->h_off[s]; a < b->h_off[s + 1]; a += EWB_T)
            st.itiles.push_back(EwbTile{s, a, std::min(a + EWB_T, b->h_off[s + 1]), 0});
    AGPU_CHECK(st.d_itiles.reserve(sizeof(EwbTile) * (st.itiles.size() + 1)));
    hipStream_t hs = stream();
    if (!st.itiles.empty())
        AGPU_HIP(hipMemcpyAsync(st.d_itiles.ptr, st.itiles.data(), sizeof(EwbTile) * st.itiles.size(),
                                hipMemcpyHostToDevice, hs));
    AGPU_HIP(hipStreamSynchronize(hs));
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_ewald_batch_compute(agpu_batch *b, const double *q, double cutoff, double beta, double kmax,
                                        int precision, double *e_struct, double *f, double *wpot_struct,
                                        double *phi, int flags)
{
    if (!b || !q || !e_struct) return set_error(AGPU_ERR_ARG, "ewald_batch_compute: bad arguments");
    if (!(cutoff > 0.0) || !(beta > 0.0) || !(kmax > 0.0))
        return set_error(AGPU_ERR_ARG, "ewald_batch_compute: need cutoff, beta, kmax > 0");
    if (precision != AGPU_PREC_FP64 && precision != AGPU_PREC_MIXED)
        return set_error(AGPU_ERR_ARG, "ewald_batch_compute: unknown precision");
    const int n = b->ntot, ns = b->nstruct;
    if (ns == 0) return AGPU_OK;
    if (!b->ewb) {
        b->ewb = new EwbState();
        b->ewb_free = ewb_free;
    }
    EwbState &st = *static_cast<EwbState *>(b->ewb);
    // k-vector runs and tiles follow the cells and atom counts (set_structures)
    if (st.struct_rev != b->struct_rev || st.beta != beta || st.kmax != kmax) {
        st.beta = beta;
        st.kmax = kmax;
        AGPU_CHECK(ewb_kvectors(b, st));
        AGPU_CHECK(ewb_ion_tiles(b, st));
        st.struct_rev = b->struct_rev;
    }
    hipStream_t s = stream();
    const bool host = flags & AGPU_HOST;
    // device staging for host arrays: [q (n) | f (3n) | phi (n) | e (ns) | w (9 ns)]
    AGPU_CHECK(st.d_q.reserve(sizeof(double) * ((size_t)5 * n + 10 * (size_t)ns + 8)));
    double *dq = st.d_q.as<double>(), *df = dq + n, *dphi = df + 3 * (size_t)n, *de = dphi + n, *dw = de + ns;
    const double *qv = q;
    double *fv = f, *pv = phi, *ev = e_struct, *wv = wpot_struct;
    if (host) {
        if (n) AGPU_HIP(hipMemcpyAsync(dq, q, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (f && n) AGPU_HIP(hipMemcpyAsync(df, f, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        if (phi && n) AGPU_HIP(hipMemcpyAsync(dphi, phi, sizeof(double) * n, hipMemcpyHostToDevice, s));
        qv = dq;
        fv = f ? df : nullptr;
        pv = phi ? dphi : nullptr;
        ev = de;
        wv = wpot_struct ? dw : nullptr;
    }
    AGPU_CHECK(st.d_ionrec.reserve(sizeof(double) * EWB_REC * ((size_t)n + 1)));
    AGPU_CHECK(st.d_s4.reserve(sizeof(double4) * ((size_t)n + 1)));
    EwbGeom G{b->off.as<int>(), b->cell.as<double>(), b->rec.as<double>()};
    const int nit = (int)st.itiles.size();
    if (nit > 0) {
        if (precision == AGPU_PREC_MIXED)
            AGPU_LAUNCH("ewald_batch_real_mixed", k_ewb_real<float>, dim3(nit), dim3(EWB_T), 0, s,
                        st.d_itiles.as<EwbTile>(), G, b->r.as<double>(), qv, cutoff, beta, fv, pv,
                        st.d_ionrec.as<double>());
        else
            AGPU_LAUNCH("ewald_batch_real_fp64", k_ewb_real<double>, dim3(nit), dim3(EWB_T), 0, s,
                        st.d_itiles.as<EwbTile>(), G, b->r.as<double>(), qv, cutoff, beta, fv, pv,
                        st.d_ionrec.as<double>());
        AGPU_LAUNCH("ewald_batch_frac", k_ewb_frac, dim3((n + 255) / 256), dim3(256), 0, s, n, b->sof.as<int>(),
                    b->rec.as<double>(), b->r.as<double>(), st.d_s4.as<double4>());
    }
    const int *koff = st.d_koff.as<int>(), *rfirst = koff + ns + 1;
    if (st.K > 0) {
        AGPU_LAUNCH("ewald_batch_sk", k_ewb_sk, dim3((unsigned)st.rtiles.size()), dim3(EWB_T), 0, s,
                    st.d_rtiles.as<EwbTile>(), st.d_runs.as<EwbRun>(), b->off.as<int>(), st.d_s4.as<double4>(), qv,
                    st.d_S.as<double>());
        AGPU_LAUNCH("ewald_batch_kenergy", k_ewb_kenergy, dim3((st.K + 255) / 256), dim3(256), 0, s, st.K,
                    st.d_kpack.as<int>(), st.d_kstruct.as<int>(), b->rec.as<double>(), b->cell.as<double>(),
                    1.0 / (4.0 * beta * beta), st.d_S.as<double>(), st.d_T.as<double>(), st.d_krec.as<double>());
        if (nit > 0 && (fv || pv))
            AGPU_LAUNCH("ewald_batch_kforce", k_ewb_kforce, dim3(nit), dim3(EWB_T), 0, s, st.d_itiles.as<EwbTile>(),
                        rfirst, st.d_runs.as<EwbRun>(), st.d_s4.as<double4>(), qv, st.d_T.as<double>(),
                        b->rec.as<double>(), b->cell.as<double>(), fv, pv);
    }
    AGPU_LAUNCH("ewald_batch_sum", k_ewb_sum, dim3((ns + 3) / 4), dim3(256), 0, s, ns, b->off.as<int>(), koff,
                st.d_ionrec.as<double>(), st.d_krec.as<double>(), ev, wv);
    if (host) {
        if (f && n) AGPU_HIP(hipMemcpyAsync(f, df, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        if (phi && n) AGPU_HIP(hipMemcpyAsync(phi, dphi, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        AGPU_HIP(hipMemcpyAsync(e_struct, de, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
        if (wpot_struct) AGPU_HIP(hipMemcpyAsync(wpot_struct, dw, sizeof(double) * 9 * ns, hipMemcpyDeviceToHost, s));
    }
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}
