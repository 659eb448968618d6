// Cell/bin neighbour-list builder for gfx950.
//
// Same list, in the same row order, as the reference's fill_neighbor_list
// (src/python/f90/python_neighbors.f90:570-754) over the bins of
// binning_init/binning_update (:765-959):
//   1. k_bin:      bin of every atom (floor(rec_cell_size.r) + 1, wrapped on
//                  periodic axes, clamped otherwise) and its wrap count.
//   2. counting sort by bin (counts and in-bin places from k_bin, exclusive
//                  scan, k_cell_fill), then k_cell_order puts each bin back
//                  into ascending atom order: the reference's linked lists,
//                  as arrays, in row-major bin order.  A bin too full for
//                  k_cell_order flags a rebuild with a radix sort of
//                  (bin, atom) pairs + k_cell_start.
//   3. the bin-sorted copy of (r, atom, wrap) (and the FP32 wrapped position
//                  for the thread-per-atom search), written by k_cell_order
//                  (k_gather after the radix sort), so that candidate scans
//                  read contiguous records.
//   4. nl_pairs:   the accepted (j, image) pairs of every atom over its
//                  stencil (x-major, then y, then z; atoms of a bin ascending)
//                  into a padded staging row of stride S per bin-order slot,
//                  neighbours named by slot, and the row counts by slot.
//                  k_bin_pairs (a wavefront per block of home bins,
//                  candidates shared by the wavefront) when bins are dense,
//                  k_atom_pairs (a thread per home atom) when they hold < 4
//                  atoms.  The force kernels read these rows and the slot
//                  records directly (bin-order storage: a neighbour's
//                  position is a slot record near the home atom's own).
//   5. exclusive scan of the counts -> slot row offsets; the caller-order
//                  offsets and k_nl_compact's caller-order CSR list (slots
//                  turned back into caller indices) only when asked.
// The distance test is the reference's FP64 expression without contraction,
// so the accepted set is bit-identical.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

struct BinGeom {
    double rec[9];   // rec_cell_size, column-major image: rec[3*k + x]
    int nc[3];
    int nd[3];
    int pbc[3];
};

__device__ __forceinline__ int floordiv(int a, int n)
{
    int q = a / n;
    return (a % n != 0 && ((a < 0) != (n < 0))) ? q - 1 : q;
}

// floor(matmul(rec, r)) + 1, then wrap (periodic) or clamp; returns the
// 1-based bin and the number of +n wraps applied per axis.
__device__ __forceinline__ void bin_of(const BinGeom &g, const double *r, int *c, int *w)
{
#pragma clang fp contract(off)
    for (int x = 0; x < 3; ++x) {
        double s = 0.0;
        s = s + g.rec[0 + x] * r[0];
        s = s + g.rec[3 + x] * r[1];
        s = s + g.rec[6 + x] * r[2];
        double fl = floor(s);
        // guard absurd coordinates (NaN / beyond int range)
        if (!(fl > -1.0e9 && fl < 1.0e9)) fl = 0.0;
        int ci = (int)fl + 1;
        if (g.pbc[x]) {
            int t = floordiv(ci - 1, g.nc[x]);
            c[x] = ci - t * g.nc[x];
            w[x] = -t;
        } else {
            c[x] = ci < 1 ? 1 : (ci > g.nc[x] ? g.nc[x] : ci);
            w[x] = 0;
        }
    }
}

// cnt (counting sort): the bin populations are counted here as well
__global__ void k_bin(int nat, const double *__restrict__ r, BinGeom g, int *__restrict__ key,
                      int *__restrict__ idx, int *__restrict__ cnt, int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nat) return;
    double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    int c[3], w[3];
    bin_of(g, ri, c, w);
    if (!dc_pack_ok(w[0]) || !dc_pack_ok(w[1]) || !dc_pack_ok(w[2])) atomicOr(flags, DERR_IMAGE_RANGE);
    const int b = ((c[0] - 1) * g.nc[1] + (c[1] - 1)) * g.nc[2] + (c[2] - 1);
    key[i] = b;
    // counting sort: the atom's provisional place in its bin; the radix
    // sort takes (key, index) pairs
    idx[i] = cnt ? atomicAdd(&cnt[b], 1) : i;
}

__global__ void k_cell_start(int nat, int ncell, const int *__restrict__ key_sorted,
                             int *__restrict__ cell_start)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nat) return;
    int kprev = (p == 0) ? -1 : key_sorted[p - 1];
    int kcur = (p == nat) ? ncell : key_sorted[p];
    for (int c = kprev + 1; c <= kcur; ++c) cell_start[c] = p;
}

// Counting sort of the atoms by bin: k_bin counts and takes each atom's
// provisional place in its bin with the same atomic, an exclusive scan gives
// the bin starts, k_cell_fill scatters, then each bin's atoms are put back
// into ascending index order, which is the order a stable sort gives.  The
// reordering takes one group of lanes per bin: every element's rank is the
// number of smaller indices in the bin (indices are distinct), counted
// through lane broadcasts; the same pass writes the sorted records.  A bin
// too full for its group flags a fallback to the radix sort.
__global__ void k_zero_i(int *__restrict__ a, long long n)
{
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        a[q] = 0;
}

__global__ void k_cell_fill(int nat, int ncell, const int *__restrict__ key, const int *__restrict__ sub,
                            int *__restrict__ cell_start, int *__restrict__ idx_sorted)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) cell_start[ncell] = nat;
    if (i >= nat) return;
    idx_sorted[cell_start[key[i]] + sub[i]] = i;
}

// the bin-sorted record of atom j at slot p (and the optional FP32 wrapped
// position q = r + A.w with max(|q_x|, |q_y|, |q_z|) in .w, its sign bit
// the wrap flag (w != 0))
__device__ __forceinline__ void gather_atom(int p, int j, const double *__restrict__ r, const BinGeom &g,
                                            const CellArg &C, SortedAtom *__restrict__ out,
                                            float4 *__restrict__ pos32)
{
    double rj[3] = {r[3 * (size_t)j], r[3 * (size_t)j + 1], r[3 * (size_t)j + 2]};
    int c[3], w[3];
    bin_of(g, rj, c, w);
    SortedAtom a;
    a.x = rj[0];
    a.y = rj[1];
    a.z = rj[2];
    a.j = j;
    a.w = dc_pack(w[0], w[1], w[2]);
    out[p] = a;
    if (pos32) {
        float q[3];
        for (int x = 0; x < 3; ++x) q[x] = (float)(rj[x] + image_x(C, x, w[0], w[1], w[2]));
        const int flag = (w[0] | w[1] | w[2]) != 0 ? (int)0x80000000u : 0;
        const float qm = fmaxf(fmaxf(fabsf(q[0]), fabsf(q[1])), fabsf(q[2]));
        pos32[p] = make_float4(q[0], q[1], q[2], __int_as_float(__float_as_int(qm) | flag));
    }
}

// G lanes per bin, up to 4 atoms per lane: bins of up to 4G atoms are put
// into ascending index order in registers, the bin counter is cleared for
// the next build, and each atom's sorted record is written at its slot; a
// fuller bin flags the radix-sort fallback.
template <int G>
__global__ void __launch_bounds__(256)
k_cell_order(int ncell, const int *__restrict__ cell_start, int *__restrict__ cnt, int *__restrict__ idx_sorted,
             int *__restrict__ flags, const double *__restrict__ r, BinGeom g, CellArg C,
             SortedAtom *__restrict__ out, float4 *__restrict__ pos32)
{
    const int lane = threadIdx.x % G;
    const int c = blockIdx.x * (256 / G) + threadIdx.x / G;
    if (c >= ncell) return;                      // uniform per group
    const int b = cell_start[c], n = cell_start[c + 1] - b;
    if (lane == 0 && n > 0) cnt[c] = 0;
    if (n == 0) return;
    if (n > 4 * G) {
        // the build is redone with the radix sort; until then every slot
        // still holds one of the bin's atoms (in scatter order), so the
        // kernels after this one read valid records
        if (lane == 0) atomicOr(flags, DERR_SORT_FALLBACK);
        for (int k = lane; k < n; k += G) gather_atom(b + k, idx_sorted[b + k], r, g, C, out, pos32);
        return;
    }
    int v[4], rk[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = lane + G * q;
        v[q] = k < n ? idx_sorted[b + k] : 0x7fffffff;
    }
    for (int k = 0; k < n; ++k) {
        const int src = k / G;
        const int w = src == 0 ? v[0] : (src == 1 ? v[1] : (src == 2 ? v[2] : v[3]));
        const int u = __shfl(w, k % G, G);
#pragma unroll
        for (int q = 0; q < 4; ++q) rk[q] += u < v[q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = lane + G * q;
        if (k < n) {
            idx_sorted[b + rk[q]] = v[q];
            gather_atom(b + rk[q], v[q], r, g, C, out, pos32);
        }
    }
}

// the sorted records after the radix sort (slot p -> atom idx_sorted[p])
__global__ void k_gather(int nat, const double *__restrict__ r, BinGeom g, CellArg C,
                         const int *__restrict__ idx_sorted, SortedAtom *__restrict__ out,
                         float4 *__restrict__ pos32)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    gather_atom(p, idx_sorted[p], r, g, C, out, pos32);
}

constexpr int NL_R = 4;         // candidate slots per lane (registers): 256 per pass
constexpr int NL_WAVES = 4;     // wavefronts per block
constexpr int NL_HMAX = 64;     // home atoms per pass

// smallest b in [0, 64) with pre[b] > e (pre = inclusive prefix, pre[63] > e)
__device__ __forceinline__ int find_bin(const int *pre, int e)
{
    int lo = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
        if (pre[lo + step - 1] <= e) lo += step;
    return lo;
}

constexpr int DC_ZERO = (512 << 20) | (512 << 10) | 512;   // dc_pack(0, 0, 0)

// a wave-uniform value moved to scalar registers
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int off_pack(int a, int b, int c) { return ((a + 128) << 16) | ((b + 128) << 8) | (c + 128); }

struct NlShared {
    int hi[NL_HMAX], ho[NL_HMAX], hp[NL_HMAX];
    int pre[64], start[64], bu[64], boff[64];
};

// pair shift s = w_home + u_bin - w_cand; distance as the reference writes it
template <bool ORTHO>
__device__ __forceinline__ double pair_d2(const CellArg &C, double xi, double yi, double zi, double xj, double yj,
                                          double zj, int s0, int s1, int s2)
{
#pragma clang fp contract(off)
    double d[3];
    if (ORTHO) {
        // off-diagonal cell entries are exactly zero, so A.s reduces to the
        // diagonal products bit for bit
        d[0] = (xi - xj) + C.A[0] * (double)s0;
        d[1] = (yi - yj) + C.A[4] * (double)s1;
        d[2] = (zi - zj) + C.A[8] * (double)s2;
    } else {
        const double ri[3] = {xi, yi, zi}, rj[3] = {xj, yj, zj};
        delta_r(C, ri, rj, s0, s1, s2, d);
    }
    return dot3_nc(d);
}

// The same distance when the home atom has no wrap (s = the candidate's own
// shift ct) and A.s was formed per candidate: (r_i - r_j) + A.s, then the
// squared norm, rounded step by step as pair_d2<true> does
__device__ __forceinline__ double pair_d2_pre(double xi, double yi, double zi, double xj, double yj, double zj,
                                              double p0, double p1, double p2)
{
#pragma clang fp contract(off)
    double d[3];
    d[0] = (xi - xj) + p0;
    d[1] = (yi - yj) + p1;
    d[2] = (zi - zj) + p2;
    return dot3_nc(d);
}

// Row statistics of a build, one record per wavefront (reduced by
// k_offsets_tail; atomics on one address from every wavefront serialise):
// part[2w] = the longest row, part[2w + 1] = 1 + the largest caller index
// with a non-empty row (the last row of the list, on which the capacity rule
// is decided)
__device__ __forceinline__ void row_stats_wave(int longest, int last, int2 *__restrict__ part, int w)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        longest = max(longest, __shfl_xor(longest, o, 64));
        last = max(last, __shfl_xor(last, o, 64));
    }
    if ((threadIdx.x & 63) == 0) part[w] = make_int2(longest, last);
}

// One wavefront per block of B x B x B home bins (B = 1 for dense bins).
// The block's candidates -- the atoms of its (B + 2 nd)^3 stencil bins in
// x-major bin order, ascending atom order inside each bin -- are held in
// registers, 256 per pass (lane L holds candidates L, L+64, L+128, L+192).
// The wavefront walks its home atoms one at a time (wave-uniform data) and
// tests 64 candidates per instruction; a candidate counts for a home atom
// only if its bin lies in the atom's own (2 nd + 1)^3 stencil, so every atom
// sees exactly the reference's candidates in the reference's order
// (python_neighbors.f90:664-735).  The ballot of accepted lanes ranks each
// accepted pair inside the row; rows go to a padded staging area (stride S
// per sorted slot, neighbours by slot).
// MULTI: blocks of B > 1 bins, whose candidates are checked against each home
// atom's own stencil
template <bool ORTHO, bool MULTI>
__global__ void __launch_bounds__(256)
k_bin_pairs(int B, BinGeom g, CellArg C, double cutoff_sq, const SortedAtom *__restrict__ atoms,
            const int *__restrict__ cell_start, int *__restrict__ scnt,
            int2 *__restrict__ stage, int S, int *__restrict__ flags, int2 *__restrict__ part)
{
    __shared__ NlShared sh[NL_WAVES];
    NlShared &L = sh[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int nbx = (g.nc[0] + B - 1) / B, nby = (g.nc[1] + B - 1) / B, nbz = (g.nc[2] + B - 1) / B;
    const int blk = xcd_block() * NL_WAVES + (threadIdx.x >> 6);
    // every wavefront of the grid leaves its row-statistics record
    if (blk >= nbx * nby * nbz) {
        if (lane == 0) part[blk] = make_int2(0, 0);
        return;
    }
    const int bz = blk % nbz, by = (blk / nbz) % nby, bx = blk / (nbz * nby);
    const int h0[3] = {bx * B + 1, by * B + 1, bz * B + 1};
    const int nh = B * B * B;

    // ---- home bins: lane L owns home bin L (x-major inside the block) ----
    int hk = 0, hs = 0, hb_off = 0;
    if (lane < nh) {
        const int oz = lane % B, oy = (lane / B) % B, ox = lane / (B * B);
        const int cx = h0[0] + ox, cy = h0[1] + oy, cz = h0[2] + oz;
        if (cx <= g.nc[0] && cy <= g.nc[1] && cz <= g.nc[2]) {
            const int b = ((cx - 1) * g.nc[1] + (cy - 1)) * g.nc[2] + (cz - 1);
            hs = cell_start[b];
            hk = cell_start[b + 1] - hs;
        }
        hb_off = off_pack(ox, oy, oz);
    }
    int hpre = hk;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int v = __shfl_up(hpre, o, 64);
        if (lane >= o) hpre += v;
    }
    const int MH = __shfl(hpre, 63, 64);
    if (MH == 0) {
        if (lane == 0) part[blk] = make_int2(0, 0);
        return;
    }
    const int sx = B + 2 * g.nd[0], sy = B + 2 * g.nd[1], sz = B + 2 * g.nd[2];
    const int nst = sx * sy * sz;
    int wmax = 0, wlast = 0;
    bool stage_over = false, img_bad = false;

    for (int a0 = 0; a0 < MH; a0 += NL_HMAX) {
        const int ma = min(NL_HMAX, MH - a0);
        __builtin_amdgcn_wave_barrier();
        L.pre[lane] = hpre;
        L.start[lane] = hs;
        L.boff[lane] = hb_off;
        __builtin_amdgcn_wave_barrier();
        if (lane < ma) {
            const int e = a0 + lane;
            const int hb = find_bin(L.pre, e);
            const int p = L.start[hb] + e - (hb ? L.pre[hb - 1] : 0);
            L.hi[lane] = atoms[p].j;
            L.ho[lane] = L.boff[hb];
            L.hp[lane] = p;
        }
        __builtin_amdgcn_wave_barrier();
        // home slots, stencil offsets and row counts held lane-wise (lane a:
        // home atom a), read with lane moves and written by lane a, so the home loop
        // below has no LDS access and its record loads can run ahead
        const int hp_lane = L.hp[lane];
        const int ho_lane = MULTI ? L.ho[lane] : 0;
        int hcnt_lane = 0;

        for (int st0 = 0; st0 < nst; st0 += 64) {
            // lane L owns stencil bin st0 + L of the block (x-major order)
            const int st = st0 + lane;
            int kL = 0, startL = 0, uL = 0, oL = 0;
            if (st < nst) {
                const int oz = st % sz - g.nd[2];
                const int oy = (st / sz) % sy - g.nd[1];
                const int ox = st / (sz * sy) - g.nd[0];
                int cc[3] = {h0[0] + ox, h0[1] + oy, h0[2] + oz};
                int uu[3] = {0, 0, 0};
                bool ok = true;
                for (int k = 0; k < 3; ++k) {
                    if (g.pbc[k]) {
                        int t = floordiv(cc[k] - 1, g.nc[k]);
                        cc[k] -= t * g.nc[k];
                        uu[k] = -t;
                    } else if (cc[k] < 1 || cc[k] > g.nc[k]) {
                        ok = false;
                    }
                }
                if (ok) {
                    const int b = ((cc[0] - 1) * g.nc[1] + (cc[1] - 1)) * g.nc[2] + (cc[2] - 1);
                    startL = cell_start[b];
                    kL = cell_start[b + 1] - startL;
                    uL = dc_pack(uu[0], uu[1], uu[2]);
                    oL = off_pack(ox, oy, oz);
                }
            }
            int incl = kL;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int v = __shfl_up(incl, o, 64);
                if (lane >= o) incl += v;
            }
            __builtin_amdgcn_wave_barrier();
            L.pre[lane] = incl;
            L.start[lane] = startL;
            L.bu[lane] = uL;
            L.boff[lane] = oL;
            __builtin_amdgcn_wave_barrier();
            const int T = L.pre[63];
            for (int c0 = 0; c0 < T; c0 += 64 * NL_R) {
                // ---- load this pass's candidates into registers ----
                double cx[NL_R], cy[NL_R], cz[NL_R], cp0[NL_R], cp1[NL_R], cp2[NL_R];
                int cj[NL_R], co[NL_R], cpk[NL_R];
                bool cv[NL_R];
#pragma unroll
                for (int r = 0; r < NL_R; ++r) {
                    const int e = c0 + 64 * r + lane;
                    cv[r] = e < T;
                    cx[r] = cy[r] = cz[r] = 0.0;
                    cp0[r] = cp1[r] = cp2[r] = 0.0;
                    cj[r] = -1;
                    co[r] = 0;
                    cpk[r] = -1;
                    if (cv[r]) {
                        const int b = find_bin(L.pre, e);
                        const int src = L.start[b] + e - (b ? L.pre[b - 1] : 0);
                        const SortedAtom A = atoms[src];
                        int u0, u1, u2, w0, w1, w2;
                        dc_unpack(L.bu[b], u0, u1, u2);
                        dc_unpack(A.w, w0, w1, w2);
                        cx[r] = A.x;
                        cy[r] = A.y;
                        cz[r] = A.z;
                        cj[r] = src;
                        const int t0 = u0 - w0, t1 = u1 - w1, t2 = u2 - w2;
                        if (MULTI) co[r] = L.boff[b];
                        if (ORTHO) {
                            cp0[r] = C.A[0] * (double)t0;
                            cp1[r] = C.A[4] * (double)t1;
                            cp2[r] = C.A[8] * (double)t2;
                        }
                        // -1: beyond the packed range (read again below)
                        if (dc_pack_ok(t0) && dc_pack_ok(t1) && dc_pack_ok(t2)) cpk[r] = dc_pack(t0, t1, t2);
                    }
                }
                // ---- home atoms one at a time (wave-uniform) ----
                // each home record is loaded (scalar) one atom ahead: the
                // index is clamped, so the load is unconditional
                int hpn = __builtin_amdgcn_readlane(hp_lane, 0);
                SortedAtom Hn = atoms[hpn];
                for (int a = 0; a < ma; ++a) {
                    const int hp = hpn;
                    const SortedAtom H = Hn;
                    hpn = __builtin_amdgcn_readlane(hp_lane, min(a + 1, ma - 1));
                    Hn = atoms[hpn];
                    int w0, w1, w2;
                    dc_unpack(H.w, w0, w1, w2);
                    const int ho = MULTI ? __builtin_amdgcn_readlane(ho_lane, a) : 0;
                    int2 *row = stage + (size_t)hp * S;
                    int cnt = __builtin_amdgcn_readlane(hcnt_lane, a);
                    // accepted lanes of one slot row, ranked after the row's
                    // entries so far
                    auto emit = [&](bool acc, int j, int pk) {
                        const unsigned long long m = __ballot(acc);
                        if (acc) {
                            const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                                (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)cnt));
                            img_bad = img_bad || pk < 0;
                            if (pos < (unsigned)S) row[pos] = make_int2(j, pk);
                        }
                        cnt += __popcll(m);
                    };
                    // a candidate of a multi-bin block counts only inside the
                    // home atom's own stencil
                    auto member = [&](int r) {
                        if (!MULTI) return cv[r];
                        const int d0 = ((co[r] >> 16) & 255) - ((ho >> 16) & 255);
                        const int d1 = ((co[r] >> 8) & 255) - ((ho >> 8) & 255);
                        const int d2 = (co[r] & 255) - (ho & 255);
                        return cv[r] && d0 >= -g.nd[0] && d0 <= g.nd[0] && d1 >= -g.nd[1] && d1 <= g.nd[1] &&
                               d2 >= -g.nd[2] && d2 <= g.nd[2];
                    };
                    if (ORTHO && (w0 | w1 | w2) == 0) {
                        // home atom inside the cell: s = ct, whose A.s and
                        // packed form were made at the candidate load; the
                        // tests combine without short-circuit branches and
                        // one exec region guards the store
#pragma unroll
                        for (int r = 0; r < NL_R; ++r) {
                            if (c0 + 64 * r >= T) break;     // slot row past the last candidate
                            const double d2 = pair_d2_pre(H.x, H.y, H.z, cx[r], cy[r], cz[r], cp0[r], cp1[r], cp2[r]);
                            const bool acc = (bool)((int)member(r) & (int)(d2 < cutoff_sq) &
                                                    (int)((cj[r] != hp) | (cpk[r] != DC_ZERO)));
                            const unsigned long long m = __ballot(acc);
                            const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                                (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)cnt));
                            img_bad = (bool)((int)img_bad | ((int)acc & (int)(cpk[r] < 0)));
                            if ((int)acc & (int)(pos < (unsigned)S)) row[pos] = make_int2(cj[r], cpk[r]);
                            cnt += __popcll(m);
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < NL_R; ++r) {
                            if (c0 + 64 * r >= T) break;
                            int t0 = 0, t1 = 0, t2 = 0;
                            if (cpk[r] >= 0) {
                                dc_unpack(cpk[r], t0, t1, t2);
                            } else if (cv[r]) {
                                // shift beyond the packed range: read it again
                                const int e = c0 + 64 * r + lane;
                                const int b = find_bin(L.pre, e);
                                const int src = L.start[b] + e - (b ? L.pre[b - 1] : 0);
                                int u0, u1, u2, v0, v1, v2;
                                dc_unpack(L.bu[b], u0, u1, u2);
                                dc_unpack(atoms[src].w, v0, v1, v2);
                                t0 = u0 - v0;
                                t1 = u1 - v1;
                                t2 = u2 - v2;
                            }
                            const int s0 = w0 + t0, s1 = w1 + t1, s2 = w2 + t2;
                            const double d2 = pair_d2<ORTHO>(C, H.x, H.y, H.z, cx[r], cy[r], cz[r], s0, s1, s2);
                            const int pk = dc_pack_ok(s0) && dc_pack_ok(s1) && dc_pack_ok(s2) ? dc_pack(s0, s1, s2) : -1;
                            emit(member(r) && !(cj[r] == hp && s0 == 0 && s1 == 0 && s2 == 0) && d2 < cutoff_sq,
                                 cj[r], pk);
                        }
                    }
                    if (lane == a) hcnt_lane = cnt;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < ma) {
            const int c = hcnt_lane;
            scnt[hp_lane] = c;
            wmax = max(wmax, c);
            if (c > 0) wlast = max(wlast, L.hi[lane] + 1);
            if (c > S) stage_over = true;
        }
    }
    row_stats_wave(wmax, wlast, part, blk);
    const bool any_over = __ballot(stage_over) != 0ull;
    const bool any_bad = __ballot(img_bad) != 0ull;
    if (lane == 0 && any_over) atomicOr(flags, DERR_STAGE_OVERFLOW);
    if (lane == 0 && any_bad) atomicOr(flags, DERR_IMAGE_RANGE);
}

// Sparse bins (a few atoms each): one thread per home atom in bin order,
// walking its own (2 nd + 1)^3 stencil in x-major bin order and each bin's
// atoms in ascending order -- the reference's candidate order
// (python_neighbors.f90:664-735) -- with the same non-contracted distance
// test, so rows equal those of the wavefront kernel above.
//
// A z column of the stencil is at most three runs of consecutive bins (bins
// are z-major in memory, so each run is one range of the sorted atoms): the
// bins wrapped from the top of the cell (shift u_z = +1), the unwrapped
// ones, the bins wrapped from the bottom (u_z = -1), in that order -- the
// order of the stencil offsets.  Every thread runs the same three scans
// (most of them empty), so wavefronts do not split at the z faces.  Cells
// with fewer than 2 nd + 1 bins along z take the bin-by-bin walk.
//
// Candidates are screened in FP32 on the wrapped positions q = r + A.w
// (r_i - r_j + A.s = q_i - q_j + A.u for the bin shift u).  Per pair, the
// error of each FP32 difference component is below
// E = 2^-24 (8 max(|q_i|, |q_j|) + 4 |A.u|_1 + 8 r_c) (the last term
// covers r_c rounded to FP32 as well): a squared distance
// below (r_c - 2E)^2 (1 - 1e-6) is inside the cut-off in exact arithmetic
// and in the FP64 test, one at or above (r_c + 2E)^2 (1 + 1e-6) outside
// (E / 4 of slack either side after the rounding of the FP32 sum of
// squares); the band between takes the FP64 test on the caller's
// positions.  A band wider than 1e-3 r_c (or a NaN / infinite coordinate)
// sends the pair to the FP64 test as well.
template <bool ORTHO>
__global__ void __launch_bounds__(256)
k_atom_pairs(int nat, BinGeom g, CellArg C, double cutoff_sq, const SortedAtom *__restrict__ atoms,
             const float4 *__restrict__ pos32, const int *__restrict__ cell_start, int *__restrict__ scnt, int2 *__restrict__ stage, int S, int *__restrict__ flags,
             int2 *__restrict__ part)
{
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    bool over = false, bad = false;
    int longest = 0, last = 0;
    if (p < nat) {
        const float4 H = pos32[p];
        int w0, w1, w2, hc[3], ih;
        {
            const SortedAtom A = atoms[p];
            const double ri[3] = {A.x, A.y, A.z};
            int hw[3];
            bin_of(g, ri, hc, hw);
            dc_unpack(A.w, w0, w1, w2);
            ih = A.j;
        }
        const float rc = (float)sqrt(cutoff_sq);
        const float hm = fabsf(H.w);
        const size_t srow = (size_t)p * S;
        int cnt = 0;
        // the candidates q0 .. q1 - 1 (atoms of consecutive bins of one
        // shift u), in order
        auto scan = [&](int q0, int q1, int ux, int uy, int uz) {
            double au[3];
            if (ORTHO) {
                // off-diagonal cell entries are zero: A.u is the diagonal
                // products, bit for bit
                au[0] = C.A[0] * (double)ux;
                au[1] = C.A[4] * (double)uy;
                au[2] = C.A[8] * (double)uz;
            } else {
                for (int x = 0; x < 3; ++x) au[x] = image_x(C, x, ux, uy, uz);
            }
            const float hx = H.x + (float)au[0], hy = H.y + (float)au[1], hz = H.z + (float)au[2];
            const float e0 = 5.9604645e-8f * (4.0f * (float)(fabs(au[0]) + fabs(au[1]) + fabs(au[2])) + 8.0f * rc);
            const int s0u = w0 + ux, s1u = w1 + uy, s2u = w2 + uz;
            auto test = [&](const float4 B, int q) {
                const float dx = hx - B.x, dy = hy - B.y, dz = hz - B.z;
                const float d2f = dx * dx + dy * dy + dz * dz;
                const float bm = fabsf(B.w);
                const float E2 = 2.0f * (e0 + 4.76837158e-7f * fmaxf(hm, bm));
                const float th = rc + E2, tl = rc - E2;
                const bool exact = !(E2 < 2e-3f * rc);
                if (!exact && !(d2f < th * th * 1.000001f)) return;
                const int jb = __float_as_int(B.w);
                int s0 = s0u, s1 = s1u, s2 = s2u;
                if (jb < 0) {
                    int v0, v1, v2;
                    dc_unpack(atoms[q].w, v0, v1, v2);
                    s0 -= v0;
                    s1 -= v1;
                    s2 -= v2;
                }
                if (q == p && s0 == 0 && s1 == 0 && s2 == 0) return;
                if (exact || !(d2f < tl * tl * 0.999999f)) {
                    const SortedAtom Ap = atoms[p], Bq = atoms[q];
                    if (!(pair_d2<ORTHO>(C, Ap.x, Ap.y, Ap.z, Bq.x, Bq.y, Bq.z, s0, s1, s2) < cutoff_sq))
                        return;
                }
                if (!dc_pack_ok(s0) || !dc_pack_ok(s1) || !dc_pack_ok(s2)) bad = true;
                if (cnt < S) stage[srow + cnt] = make_int2(q, dc_pack(s0, s1, s2));
                ++cnt;
            };
            // two candidates per trip, both records loaded before either test
            // (the second index clamped into the run, so both loads are in
            // bounds and unconditional)
            for (int q = q0; q < q1; q += 2) {
                const float4 B0 = pos32[q];
                const float4 B1 = pos32[min(q + 1, q1 - 1)];
                test(B0, q);
                if (q + 1 < q1) test(B1, q + 1);
            }
        };
        const int nz = g.nc[2], ndz = g.nd[2];
        const bool runs = !g.pbc[2] || 2 * ndz + 1 <= nz;
        const int zlo = hc[2] - ndz, zhi = hc[2] + ndz;
        if (runs && g.nd[0] == 1 && g.nd[1] == 1 && ndz == 1) {
            // a 3 x 3 x 3 stencil (bins of at least the cut-off): the nine
            // column ranges are read first, all loads in flight together,
            // then scanned in stencil order.  Column k = 3 (ox + 1) + (oy + 1)
            // lies one bin at most outside the cell, so its wrap is a compare
            const int m0 = max(zlo, 1), m1 = min(zhi, nz);
            int bcs[9], uxs[9], uys[9], q0[9], q1[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                int cx = hc[0] + k / 3 - 1, cy = hc[1] + k % 3 - 1, ux = 0, uy = 0;
                bool ok = true;
                if (g.pbc[0]) {
                    if (cx < 1) { cx += g.nc[0]; ux = 1; }
                    else if (cx > g.nc[0]) { cx -= g.nc[0]; ux = -1; }
                } else {
                    ok = cx >= 1 && cx <= g.nc[0];
                }
                if (g.pbc[1]) {
                    if (cy < 1) { cy += g.nc[1]; uy = 1; }
                    else if (cy > g.nc[1]) { cy -= g.nc[1]; uy = -1; }
                } else {
                    ok = ok && cy >= 1 && cy <= g.nc[1];
                }
                const int bc = ok ? ((cx - 1) * g.nc[1] + (cy - 1)) * nz : -1;
                bcs[k] = bc;
                uxs[k] = ux;
                uys[k] = uy;
                q0[k] = bc >= 0 ? cell_start[bc + m0 - 1] : 0;
                q1[k] = bc >= 0 ? cell_start[bc + m1] : 0;
            }
            const bool tw = g.pbc[2] && zlo < 1, uw = g.pbc[2] && zhi > nz;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int bc = bcs[k];
                if (bc < 0) continue;
                // wrapped from the top (u_z = +1), unwrapped, wrapped from the bottom (u_z = -1)
                if (tw) scan(cell_start[bc + zlo + nz - 1], cell_start[bc + nz], uxs[k], uys[k], 1);
                scan(q0[k], q1[k], uxs[k], uys[k], 0);
                if (uw) scan(cell_start[bc], cell_start[bc + zhi - nz], uxs[k], uys[k], -1);
            }
        } else
        for (int ox = -g.nd[0]; ox <= g.nd[0]; ++ox) {
            int cx = hc[0] + ox, ux = 0;
            if (g.pbc[0]) {
                const int t = floordiv(cx - 1, g.nc[0]);
                cx -= t * g.nc[0];
                ux = -t;
            } else if (cx < 1 || cx > g.nc[0]) {
                continue;
            }
            for (int oy = -g.nd[1]; oy <= g.nd[1]; ++oy) {
                int cy = hc[1] + oy, uy = 0;
                if (g.pbc[1]) {
                    const int t = floordiv(cy - 1, g.nc[1]);
                    cy -= t * g.nc[1];
                    uy = -t;
                } else if (cy < 1 || cy > g.nc[1]) {
                    continue;
                }
                const int bcol = ((cx - 1) * g.nc[1] + (cy - 1)) * nz;
                if (runs) {
                    // bins (1-based) of the three runs; an empty run has b1 < b0
                    const int m0 = max(zlo, 1), m1 = min(zhi, nz);
                    int t0 = 1, t1 = 0, u0 = 1, u1 = 0;
                    if (g.pbc[2]) {
                        t0 = zlo + nz;            // wrapped from the top: zlo .. 0
                        t1 = nz;
                        u0 = 1;                   // wrapped from the bottom: nz+1 .. zhi
                        u1 = zhi - nz;
                    }
                    if (t0 <= t1) scan(cell_start[bcol + t0 - 1], cell_start[bcol + t1], ux, uy, 1);
                    scan(cell_start[bcol + m0 - 1], cell_start[bcol + m1], ux, uy, 0);
                    if (u0 <= u1) scan(cell_start[bcol + u0 - 1], cell_start[bcol + u1], ux, uy, -1);
                    continue;
                }
                for (int oz = -ndz; oz <= ndz; ++oz) {
                    int cz = hc[2] + oz, uz = 0;
                    const int t = floordiv(cz - 1, nz);
                    cz -= t * nz;
                    uz = -t;
                    const int b = bcol + (cz - 1);
                    scan(cell_start[b], cell_start[b + 1], ux, uy, uz);
                }
            }
        }
        scnt[p] = cnt;
        over = cnt > S;
        longest = cnt;
        last = cnt > 0 ? ih + 1 : 0;
    }
    // one row-statistics record per block
    {
        __shared__ int2 sm[4];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            longest = max(longest, __shfl_xor(longest, o, 64));
            last = max(last, __shfl_xor(last, o, 64));
        }
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = make_int2(longest, last);
        __syncthreads();
        if (threadIdx.x == 0)
            part[blockIdx.x] = make_int2(max(max(sm[0].x, sm[1].x), max(sm[2].x, sm[3].x)),
                                         max(max(sm[0].y, sm[1].y), max(sm[2].y, sm[3].y)));
    }
    const unsigned long long mo = __ballot(over), mb = __ballot(bad);
    if ((threadIdx.x & 63) == 0) {
        if (mo) atomicOr(flags, DERR_STAGE_OVERFLOW);
        if (mb) atomicOr(flags, DERR_IMAGE_RANGE);
    }
}

// Copy staged rows (sorted slot p, stride S, neighbours by slot) into the
// CSR rows of the caller atoms, neighbours by caller index (nothing when the
// build overflowed).
__global__ void __launch_bounds__(256)
k_nl_compact(int nat, const SortedAtom *__restrict__ atoms, const int *__restrict__ scnt,
             const int *__restrict__ offsets, const int2 *__restrict__ stage, int S, int2 *__restrict__ nbr,
             const int *__restrict__ flags)
{
    constexpr int G = 16;
    const int lane = threadIdx.x % G;
    const int p = blockIdx.x * (256 / G) + threadIdx.x / G;
    if (p >= nat) return;
    if (*flags & (DERR_NL_OVERFLOW | DERR_STAGE_OVERFLOW)) return;
    const int i = atoms[p].j;
    const int cnt = scnt[p];
    if (cnt > S) return;
    const int2 *src = stage + (size_t)p * S;
    int2 *dst = nbr + offsets[i];
    for (int o = lane; o < cnt; o += G) {
        const int2 e = src[o];
        dst[o] = make_int2(atoms[e.x].j, e.y);
    }
}

// the CSR rows of the current build, copied from the staged rows if needed
static int nl_compact(agpu_neighbors *nl)
{
    if (nl->csr) return AGPU_OK;
    AGPU_CHECK(nl_caller_offsets(nl));
    if (nl->nat > 0)
        AGPU_LAUNCH("nl_compact", k_nl_compact, dim3((nl->nat + 15) / 16), dim3(256), 0, stream(), nl->nat,
                    nl->atoms_sorted.as<SortedAtom>(), nl->scnt.as<int>(), nl->offsets.as<int>(),
                    nl->stage.as<int2>(), nl->stage_s, nl->nbr.as<int2>(), nl->rowmax.as<int>() + 3);
    nl->csr = true;
    return AGPU_OK;
}

// the positions of the slots from the caller's positions (caller atom
// order[p] at slot p); caller index and wrap counts stay
__global__ void __launch_bounds__(256)
k_slot_positions(int nat, const int *__restrict__ order, const double *__restrict__ r,
                 SortedAtom *__restrict__ atoms)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    const int i = order[p];
    SortedAtom a = atoms[p];
    a.x = r[3 * (size_t)i];
    a.y = r[3 * (size_t)i + 1];
    a.z = r[3 * (size_t)i + 2];
    atoms[p] = a;
}

// The slot positions from the particles where they may lag them: a list
// held over moving atoms (positions change without a revision), or an
// imported one whose particles moved since.  Built lists write them.
static int nl_slots_refresh(agpu_neighbors *nl, const agpu_particles *p)
{
    const int nat = nl->nat;
    if (nat == 0 || (!nl->held && nl->slot_rev == p->pos_rev)) return AGPU_OK;
    AGPU_LAUNCH("nl_slot_positions", k_slot_positions, dim3((nat + 255) / 256), dim3(256), 0, stream(), nat,
                nl->idx_sorted.as<int>(), p->d_r, nl->atoms_sorted.as<SortedAtom>());
    nl->slot_rev = p->pos_rev;
    return AGPU_OK;
}

int nl_slots_current(agpu_neighbors *nl, const agpu_particles *p) { return nl_slots_refresh(nl, p); }

NlRows nl_rows(const agpu_neighbors *nl)
{
    if (nl->external)
        return NlRows{nl->nbr.as<int2>(), nullptr, nl->offsets.as<int>(), 0, nl->atoms_sorted.as<SortedAtom>()};
    return NlRows{nl->stage.as<int2>(), nl->scnt.as<int>(), nl->soffs.as<int>(), nl->stage_s,
                  nl->atoms_sorted.as<SortedAtom>()};
}

// row lengths by caller atom (caller atom order[p] at slot p)
__global__ void k_counts_caller(int nat, const int *__restrict__ order, const int *__restrict__ scnt,
                                int *__restrict__ counts)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nat) counts[order[p]] = scnt[p];
}

// offsets[nat] = offsets[nat - 1] + counts[nat - 1] after the exclusive
// scan of nat entries
__global__ void k_offsets_end(int nat, const int *__restrict__ counts, int *__restrict__ offsets)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) offsets[nat] = nat > 0 ? offsets[nat - 1] + counts[nat - 1] : 0;
}

int nl_caller_offsets(agpu_neighbors *nl)
{
    if (nl->caller_off || nl->external) return AGPU_OK;
    const int nat = nl->nat;
    hipStream_t s = stream();
    if (nat > 0) {
        AGPU_LAUNCH("nl_counts", k_counts_caller, dim3((nat + 255) / 256), dim3(256), 0, s, nat,
                    nl->idx_sorted.as<int>(), nl->scnt.as<int>(), nl->counts.as<int>());
        size_t tmp = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, nl->counts.as<int>(), nl->offsets.as<int>(), 0, (size_t)nat,
                                         rocprim::plus<int>(), s));
        // reserved by the build for both scans
        if (tmp + 16 > nl->sort_tmp.bytes) return set_error(AGPU_ERR_STATE, "neighbour list: scan scratch too small");
        AGPU_HIP(rocprim::exclusive_scan(nl->sort_tmp.ptr, tmp, nl->counts.as<int>(), nl->offsets.as<int>(), 0,
                                         (size_t)nat, rocprim::plus<int>(), s));
        AGPU_LAUNCH("nl_offsets_end", k_offsets_end, dim3(1), dim3(64), 0, s, nat, nl->counts.as<int>(),
                    nl->offsets.as<int>());
    }
    nl->caller_off = true;
    return AGPU_OK;
}

// The row statistics of the pair search (npart records {longest row,
// 1 + last non-empty caller row}) reduced into stats[0] and the last row;
// the total (stats[1..2], from the scan of the row lengths in slot order,
// offsets[nat] = total); and the reference's capacity rule
// (python_neighbors.f90:713-718): caller row i ends at 1-based slot
// (entries of rows before i) + cnt + i (one separator per earlier atom),
// increasing with i, so the list overflows exactly when the last non-empty
// row il ends at or beyond nat*avgn -- and that row ends at total + il,
// every later row being empty.  One block of 1024 threads.
__global__ void __launch_bounds__(1024)
k_offsets_tail(int nat, const int *__restrict__ counts, int *__restrict__ offsets, int *__restrict__ stats,
               long long cap, const int2 *__restrict__ part, int npart)
{
    __shared__ int2 sm[16];
    int longest = 0, last = 0;
    for (int k = threadIdx.x; k < npart; k += 1024) {
        const int2 v = part[k];
        longest = max(longest, v.x);
        last = max(last, v.y);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        longest = max(longest, __shfl_xor(longest, o, 64));
        last = max(last, __shfl_xor(last, o, 64));
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = make_int2(longest, last);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; ++w) {
            longest = max(longest, sm[w].x);
            last = max(last, sm[w].y);
        }
        stats[0] = longest;
        stats[5] = last;
    }
    // offsets[nat] = total.  The 32-bit scan cannot wrap while
    // nat * rowmax < 2^31; larger lists are refused as an overflow (the
    // capacity nat*avgn is below 2^31 anyway).
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if ((long long)nat * (long long)stats[0] >= 0x7fffffffll) {
            stats[3] |= DERR_NL_OVERFLOW;
            offsets[nat] = 0;
            stats[1] = 0x7fffffff;
            stats[2] = 0;
            return;
        }
        long long t = (nat > 0) ? (long long)offsets[nat - 1] + counts[nat - 1] : 0;
        offsets[nat] = (int)t;
        stats[1] = (int)(t & 0x7fffffff);
        stats[2] = (int)(t >> 31);
        const int il = stats[5] - 1;
        if (il >= 0 && t + il >= cap) stats[3] |= DERR_NL_OVERFLOW;
    }
}

static void geometry(const agpu_particles *p, agpu_neighbors *nl, BinGeom &g)
{
    double bin = nl->bin_size > 0.0 ? nl->bin_size : nl->cutoff;
    const double *A = p->cell.A;
    double cs[3][3];
    for (int x = 0; x < 3; ++x) {
        double s = 0.0;
        s = s + A[3 * x] * A[3 * x];
        s = s + A[3 * x + 1] * A[3 * x + 1];
        s = s + A[3 * x + 2] * A[3 * x + 2];
        int n = (int)(std::sqrt(s) / bin);
        nl->nc[x] = n < 3 ? 3 : n;
    }
    for (int x = 0; x < 3; ++x)
        for (int c = 0; c < 3; ++c) cs[x][c] = A[3 * x + c] / nl->nc[x];
    auto cross = [](const double *u, const double *v, double *w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    auto dot = [](const double *u, const double *v) {
        double s = 0.0;
        s = s + u[0] * v[0];
        s = s + u[1] * v[1];
        s = s + u[2] * v[2];
        return s;
    };
    double nn[3][3];
    cross(cs[1], cs[2], nn[0]);
    cross(cs[2], cs[0], nn[1]);
    cross(cs[0], cs[1], nn[2]);
    double cv = dot(cs[0], nn[0]);
    for (int a = 0; a < 3; ++a) {
        double d2 = dot(nn[a], nn[a]);
        for (int c = 0; c < 3; ++c) nn[a][c] = cv * nn[a][c] / d2;
        nl->nd[a] = (int)(nl->cutoff / std::sqrt(dot(nn[a], nn[a]))) + 1;
    }
    for (int x = 0; x < 3; ++x)
        for (int k = 0; k < 3; ++k) nl->rec[3 * k + x] = p->cell.B[3 * k + x] * nl->nc[x];
    std::memcpy(g.rec, nl->rec, sizeof g.rec);
    for (int x = 0; x < 3; ++x) {
        g.nc[x] = nl->nc[x];
        g.nd[x] = nl->nd[x];
        g.pbc[x] = p->cell.pbc[x] != 0;
    }
}

bool nl_current(const agpu_neighbors *nl, const agpu_particles *p)
{
    if (nl->pending) return false;      // a build whose flags nobody checked
    if (nl->external || nl->held) return nl->p == p && nl->nat == p->nat;
    return nl->p == p && nl->pos_rev == p->pos_rev && nl->cell_rev == p->cell_rev && nl->nat == p->nat;
}

const int *nl_flags(const agpu_neighbors *nl) { return nl->rowmax.as<int>() + 3; }

int nl_settle(agpu_neighbors *nl, bool *redo)
{
    *redo = false;
    if (!nl->pending) return AGPU_OK;
    nl->pending = false;
    const int fl = nl->stats.h[3];
    nl->maxrow = nl->stats.h[0];
    if (fl & DERR_SORT_FALLBACK) {
        nl->radix = true;
        nl->p = nullptr;
        *redo = true;
        return AGPU_OK;
    }
    if (fl & DERR_STAGE_OVERFLOW) {
        nl->stage_w = ((int)(1.25 * nl->maxrow) + 16 + 7) & ~7;
        nl->p = nullptr;
        *redo = true;
        return AGPU_OK;
    }
    nl->total = (int64_t)nl->stats.h[1] | ((int64_t)nl->stats.h[2] << 31);
    {
        const int want = (nl->maxrow + std::max(2, nl->maxrow / 4) + 3) & ~3;
        if (want >= 4 && 4 * want <= 3 * nl->stage_w) nl->stage_w = want;
    }
    if (fl & DERR_IMAGE_RANGE) {
        nl->p = nullptr;
        return set_error(AGPU_ERR_DOMAIN, "Particle outside simulation domain: image shift beyond +-511 cells.");
    }
    if (fl & DERR_NL_OVERFLOW) {
        nl->p = nullptr;
        return set_error(AGPU_ERR_NL_OVERFLOW,
                         "Neighbor list overflow: " + std::to_string(nl->total) + " neighbours + " +
                             std::to_string(nl->nat) + " separators exceed nat*avgn = " +
                             std::to_string(nl->cap) + ". Increase avgn.");
    }
    return AGPU_OK;
}

int nl_update(agpu_neighbors *nl, agpu_particles *p, int force, bool csr, bool defer)
{
    if (nl->pending) {
        // an earlier deferred build that nobody checked
        AGPU_HIP(hipStreamSynchronize(stream()));
        bool redo = false;
        AGPU_CHECK(nl_settle(nl, &redo));
    }
    if (nl->external) {
        // an imported list stays as the caller gave it
        if (nl->p != p || nl->nat != p->nat)
            return set_error(AGPU_ERR_STATE, "neighbour list: imported for other particles");
        return nl_slots_refresh(nl, p);
    }
    if (nl->held && force == 0 && nl->p == p && nl->nat == p->nat) {
        AGPU_CHECK(nl_slots_refresh(nl, p));
        return csr ? nl_compact(nl) : AGPU_OK;
    }
    if (!p->cell_set) return set_error(AGPU_ERR_STATE, "particles: cell not set");
    if (p->nat > 0 && !p->d_r) return set_error(AGPU_ERR_STATE, "particles: positions not set");
    if (!force && nl->p == p && nl->pos_rev == p->pos_rev && nl->cell_rev == p->cell_rev &&
        nl->nat == p->nat)
        return csr ? nl_compact(nl) : AGPU_OK;
    // invalid until this build completes
    nl->p = nullptr;
    nl->csr = true;
    // verlet shell 0 (aseinterface.py:255) unless a driver set one
    nl->cutoff = nl->interaction_range + nl->skin;
    if (nl->cutoff <= 0.0)
        return set_error(AGPU_ERR_ARG, "Cutoff needs to be larger than zero.");
    const int nat = p->nat;
    hipStream_t s = stream();
    BinGeom g;
    geometry(p, nl, g);
    long long ncell = (long long)g.nc[0] * g.nc[1] * g.nc[2];
    if (ncell > (1ll << 30)) return set_error(AGPU_ERR_ARG, "neighbour list: too many bins");
    nl->cap = (int64_t)nat * nl->avgn;
    if (nl->cap > 0x7fffffffll)
        return set_error(AGPU_ERR_ARG, "neighbour list: nat*avgn exceeds 2^31 entries");
    AGPU_CHECK(nl->stats.init());
    AGPU_CHECK(nl->key.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->key_sorted.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->idx.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->idx_sorted.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->cell_start.reserve(sizeof(int) * (ncell + 1)));
    AGPU_CHECK(nl->atoms_sorted.reserve(sizeof(SortedAtom) * (nat + 1)));
    AGPU_CHECK(nl->counts.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->scnt.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->soffs.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->offsets.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->rowmax.reserve(sizeof(int) * 8));
    AGPU_CHECK(nl->nbr.reserve(sizeof(int2) * (nl->cap > 0 ? nl->cap : 1)));
    // [0] rowmax, [1..2] total, [3] flags, [5] 1 + the last caller index
    // with a non-empty row
    int *flags = nl->rowmax.as<int>();
    AGPU_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 8, s));
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const int TB = 256;
    if (nat > 0) {
        int nb = (nat + TB - 1) / TB;
        int *cnt = nullptr;
        long long cnt_zero = 0;
        if (!nl->radix) {
            // the bin counters are zero again after k_cell_fill; they are
            // cleared only when new, grown or left by an unfinished build
            void *old = nl->cell_cnt.ptr;
            AGPU_CHECK(nl->cell_cnt.reserve(sizeof(int) * (ncell + 1)));
            cnt = nl->cell_cnt.as<int>();
            cnt_zero = nl->cell_cnt.ptr == old ? nl->cnt_zero : 0;
            nl->cnt_zero = 0;
            if (cnt_zero < ncell) {
                const int nzb = (int)std::min<long long>(1024, (ncell + TB - 1) / TB);
                AGPU_LAUNCH("nl_sort_zero", k_zero_i, dim3(nzb), dim3(TB), 0, s, cnt, ncell);
            }
        }
        const double per_bin = (double)nat / (double)ncell;
        const bool per_atom = nl->pair_mode == 1 || (nl->pair_mode == 0 && per_bin < 4.0);
        if (per_atom) AGPU_CHECK(nl->pos32.reserve(sizeof(float4) * (nat + 1)));
        float4 *pos32 = per_atom ? nl->pos32.as<float4>() : nullptr;
        AGPU_LAUNCH("nl_bin", k_bin, dim3(nb), dim3(TB), 0, s, nat, p->d_r, g,
                    nl->key.as<int>(), nl->idx.as<int>(), cnt, flags + 3);
        if (!nl->radix) {
            size_t tmp = 0;
            AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp, cnt, nl->cell_start.as<int>(), 0, (size_t)ncell,
                                             rocprim::plus<int>(), s));
            AGPU_CHECK(nl->sort_tmp.reserve(tmp + 16));
            AGPU_HIP(rocprim::exclusive_scan(nl->sort_tmp.ptr, tmp, cnt, nl->cell_start.as<int>(), 0,
                                             (size_t)ncell, rocprim::plus<int>(), s));
            AGPU_LAUNCH("nl_sort_fill", k_cell_fill, dim3(nb), dim3(TB), 0, s, nat, (int)ncell, nl->key.as<int>(),
                        nl->idx.as<int>(), nl->cell_start.as<int>(), nl->idx_sorted.as<int>());
            // lanes per bin from the mean occupancy (capacity 4 atoms per
            // lane); the same pass writes the sorted records
            const double occ = (double)nat / (double)ncell;
            auto ord = occ > 8.0 ? k_cell_order<64> : (occ > 2.0 ? k_cell_order<16> : k_cell_order<4>);
            const int bpb = occ > 8.0 ? 4 : (occ > 2.0 ? 16 : 64);
            AGPU_LAUNCH("nl_sort_order", ord, dim3((unsigned)((ncell + bpb - 1) / bpb)), dim3(TB), 0, s, (int)ncell,
                        nl->cell_start.as<int>(), cnt, nl->idx_sorted.as<int>(), flags + 3, p->d_r, g, C,
                        nl->atoms_sorted.as<SortedAtom>(), pos32);
            nl->cnt_zero = std::max(cnt_zero, ncell);
        } else {
            size_t tmp = 0;
            int end_bit = 1;
            while ((1ll << end_bit) <= ncell) ++end_bit;
            AGPU_HIP(rocprim::radix_sort_pairs(nullptr, tmp, nl->key.as<int>(), nl->key_sorted.as<int>(),
                                               nl->idx.as<int>(), nl->idx_sorted.as<int>(),
                                               (size_t)nat, 0, end_bit, s));
            AGPU_CHECK(nl->sort_tmp.reserve(tmp + 16));
            AGPU_HIP(rocprim::radix_sort_pairs(nl->sort_tmp.ptr, tmp, nl->key.as<int>(),
                                               nl->key_sorted.as<int>(), nl->idx.as<int>(),
                                               nl->idx_sorted.as<int>(), (size_t)nat, 0, end_bit, s));
            AGPU_LAUNCH("nl_cell_start", k_cell_start, dim3((nat + 1 + TB - 1) / TB), dim3(TB), 0, s,
                        nat, (int)ncell, nl->key_sorted.as<int>(), nl->cell_start.as<int>());
        }
        if (nl->radix)
            AGPU_LAUNCH("nl_gather", k_gather, dim3(nb), dim3(TB), 0, s, nat, p->d_r, g, C,
                        nl->idx_sorted.as<int>(), nl->atoms_sorted.as<SortedAtom>(), pos32);
        double cutoff_sq = nl->cutoff * nl->cutoff;
        // bins per wavefront edge: one bin when bins are dense, a 2^3 or 3^3
        // block of bins when they hold few atoms (stencil of +-1 bin only)
        int B = 1;
        if (g.nd[0] == 1 && g.nd[1] == 1 && g.nd[2] == 1 && per_bin < 4.0) B = 2;
        const bool ortho = C.A[1] == 0.0 && C.A[2] == 0.0 && C.A[3] == 0.0 && C.A[5] == 0.0 &&
                           C.A[6] == 0.0 && C.A[7] == 0.0;
        const long long nblk_b = (long long)((g.nc[0] + B - 1) / B) * ((g.nc[1] + B - 1) / B) * ((g.nc[2] + B - 1) / B);
        const int nbin_blk = (int)((nblk_b + NL_WAVES - 1) / NL_WAVES);
        // staging stride: expected neighbours from the density, grown on demand
        if (nl->stage_w <= 0) {
            const double *A = p->cell.A;
            double V = std::fabs(A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                                 A[2] * (A[3] * A[7] - A[4] * A[6]));
            double est = (V > 0 ? nat / V : 0.0) * 4.18879 * nl->cutoff * nl->cutoff * nl->cutoff;
            nl->stage_w = ((int)(1.4 * est) + 16 + 7) & ~7;
        }
        const int S = nl->stage_w;
        AGPU_CHECK(nl->stage.reserve(sizeof(int2) * (size_t)nat * S));
        // row-statistics records: one per block (thread per atom) or per
        // wavefront (wavefront per bin block)
        const int na = (nat + TB - 1) / TB;
        const int npart = per_atom ? na : nbin_blk * NL_WAVES;
        AGPU_CHECK(nl->part.reserve(sizeof(int2) * (size_t)npart));
        int2 *part = nl->part.as<int2>();
        if (per_atom) {
            if (ortho)
                AGPU_LAUNCH("nl_pairs", k_atom_pairs<true>, dim3(na), dim3(TB), 0, s, nat, g, C, cutoff_sq,
                            nl->atoms_sorted.as<SortedAtom>(), nl->pos32.as<float4>(), nl->cell_start.as<int>(),
                            nl->scnt.as<int>(), nl->stage.as<int2>(), S, flags + 3, part);
            else
                AGPU_LAUNCH("nl_pairs", k_atom_pairs<false>, dim3(na), dim3(TB), 0, s, nat, g, C, cutoff_sq,
                            nl->atoms_sorted.as<SortedAtom>(), nl->pos32.as<float4>(), nl->cell_start.as<int>(),
                            nl->scnt.as<int>(), nl->stage.as<int2>(), S, flags + 3, part);
        } else {
            auto kern = ortho ? (B > 1 ? k_bin_pairs<true, true> : k_bin_pairs<true, false>)
                              : (B > 1 ? k_bin_pairs<false, true> : k_bin_pairs<false, false>);
            AGPU_LAUNCH("nl_pairs", kern, dim3(nbin_blk), dim3(TB), 0, s, B, g, C, cutoff_sq,
                        nl->atoms_sorted.as<SortedAtom>(), nl->cell_start.as<int>(), nl->scnt.as<int>(),
                        nl->stage.as<int2>(), S, flags + 3, part);
        }
        // row offsets in slot order (the per-entry arrays of the force
        // kernels); the caller-order offsets follow when a caller asks
        size_t tmp2 = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp2, nl->scnt.as<int>(), nl->soffs.as<int>(), 0,
                                         (size_t)nat, rocprim::plus<int>(), s));
        AGPU_CHECK(nl->sort_tmp.reserve(tmp2 + 16));
        AGPU_HIP(rocprim::exclusive_scan(nl->sort_tmp.ptr, tmp2, nl->scnt.as<int>(),
                                         nl->soffs.as<int>(), 0, (size_t)nat, rocprim::plus<int>(), s));
        AGPU_LAUNCH("nl_offsets_tail", k_offsets_tail, dim3(1), dim3(1024), 0, s, nat,
                    nl->scnt.as<int>(), nl->soffs.as<int>(), flags, (long long)nl->cap, part, npart);
    } else {
        AGPU_HIP(hipMemsetAsync(nl->soffs.ptr, 0, sizeof(int), s));
        AGPU_HIP(hipMemsetAsync(nl->offsets.ptr, 0, sizeof(int), s));
    }
    AGPU_HIP(hipMemcpyAsync(nl->stats.h, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    if (defer && !csr && nat > 0 && !capturing()) {
        // the rows are taken as built; nl_settle checks the flags after the
        // caller's synchronisation
        nl->pending = true;
        nl->stage_s = nl->stage_w;
        nl->csr = false;
        nl->caller_off = false;
        nl->p = p;
        nl->nat = nat;
        nl->pos_rev = p->pos_rev;
        nl->cell_rev = p->cell_rev;
        nl->slot_rev = p->pos_rev;
        nl->n_updates++;
        return AGPU_OK;
    }
    AGPU_HIP(hipStreamSynchronize(s));
    int fl = nl->stats.h[3];
    nl->maxrow = nl->stats.h[0];
    if (fl & DERR_SORT_FALLBACK) {
        // a bin holds more atoms than the per-bin ordering handles: redo the
        // binning with the radix sort from now on
        nl->radix = true;
        nl->p = nullptr;
        return nl_update(nl, p, force > 0 ? force : 1, csr);
    }
    if (fl & DERR_STAGE_OVERFLOW) {
        // a row outgrew the staging stride: widen it and rebuild once
        nl->stage_w = ((int)(1.25 * nl->maxrow) + 16 + 7) & ~7;
        nl->p = nullptr;
        if (force > 1) return set_error(AGPU_ERR_NL_OVERFLOW, "neighbour list: staging stride overflow");
        return nl_update(nl, p, 2, csr);
    }
    nl->total = (int64_t)nl->stats.h[1] | ((int64_t)nl->stats.h[2] << 31);
    nl->stage_s = nl->stage_w;
    // the staging stride follows the rows: the next build narrows it when
    // the longest row leaves a quarter of it unused (a quarter of the row,
    // at least 2 entries, kept as headroom; a row that outgrows it later
    // widens it again through the overflow rebuild)
    {
        const int want = (nl->maxrow + std::max(2, nl->maxrow / 4) + 3) & ~3;
        if (want >= 4 && 4 * want <= 3 * nl->stage_w) nl->stage_w = want;
    }
    nl->csr = nat == 0;
    nl->caller_off = nat == 0;
    nl->p = p;
    nl->nat = nat;
    nl->pos_rev = p->pos_rev;
    nl->cell_rev = p->cell_rev;
    nl->slot_rev = p->pos_rev;
    nl->n_updates++;
    if (fl & DERR_IMAGE_RANGE) {
        nl->p = nullptr;
        return set_error(AGPU_ERR_DOMAIN,
                         "Particle outside simulation domain: image shift beyond +-511 cells.");
    }
    if (fl & DERR_NL_OVERFLOW) {
        nl->p = nullptr;
        return set_error(AGPU_ERR_NL_OVERFLOW,
                         "Neighbor list overflow: " + std::to_string(nl->total) + " neighbours + " +
                             std::to_string(nat) + " separators exceed nat*avgn = " +
                             std::to_string(nl->cap) + ". Increase avgn.");
    }
    return csr ? nl_compact(nl) : AGPU_OK;
}

__global__ void k_iota(int n, int *__restrict__ a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// slots of an imported list: the caller's atoms in their own order
__global__ void k_slots_identity(int nat, const double *__restrict__ r, SortedAtom *__restrict__ atoms)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    SortedAtom a;
    a.x = r[3 * (size_t)p];
    a.y = r[3 * (size_t)p + 1];
    a.z = r[3 * (size_t)p + 2];
    a.j = p;
    a.w = DC_ZERO;
    atoms[p] = a;
}

// The caller's own full list (LAMMPS: firstneigh/numneigh of local and ghost
// atoms, as PairAtomistica::Atomistica_neigh hands it to the reference,
// pair_atomistica.cpp:385-470): row i is neigh[seed[i] .. seed[i] +
// numneigh[i]), indices into the particles, no periodic images (the caller
// supplies ghost atoms instead).  LAMMPS keeps special-bond flags in the top
// two bits of an entry; they are masked off.  The rows are kept in the
// caller's order and are used until the next import or rebuild request.
int nl_import(agpu_neighbors *nl, agpu_particles *p, const int64_t *seed, const int *numneigh, const int *neigh)
{
    const int nat = p->nat;
    std::vector<int> off(nat + 1, 0);
    int64_t total = 0;
    int maxrow = 0;
    for (int i = 0; i < nat; ++i) {
        if (numneigh[i] < 0) return set_error(AGPU_ERR_ARG, "neighbour import: negative row length");
        off[i] = (int)total;
        total += numneigh[i];
        maxrow = std::max(maxrow, numneigh[i]);
        if (total > 0x7fffffffll) return set_error(AGPU_ERR_ARG, "neighbour import: more than 2^31 entries");
    }
    off[nat] = (int)total;
    std::vector<int2> rows((size_t)total + 1);
    const int zero = (512 << 20) | (512 << 10) | 512;     // image (0, 0, 0)
    for (int i = 0; i < nat; ++i) {
        const int *src = neigh + seed[i];
        for (int k = 0; k < numneigh[i]; ++k) {
            const int j = src[k] & 0x3fffffff;
            if (j >= nat) return set_error(AGPU_ERR_ARG, "neighbour import: index beyond the atoms");
            rows[off[i] + k] = make_int2(j, zero);
        }
    }
    AGPU_CHECK(nl->offsets.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->nbr.reserve(sizeof(int2) * ((size_t)total + 1)));
    AGPU_CHECK(nl->idx_sorted.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->atoms_sorted.reserve(sizeof(SortedAtom) * (nat + 1)));
    AGPU_CHECK(nl->rowmax.reserve(sizeof(int) * 8));
    if (nat > 0 && !p->d_r) return set_error(AGPU_ERR_STATE, "particles: positions not set");
    hipStream_t s = stream();
    // no build flags: the imported rows are usable as they are
    AGPU_HIP(hipMemsetAsync(nl->rowmax.ptr, 0, sizeof(int) * 8, s));
    AGPU_HIP(hipMemcpyAsync(nl->offsets.ptr, off.data(), sizeof(int) * (nat + 1), hipMemcpyHostToDevice, s));
    if (total) AGPU_HIP(hipMemcpyAsync(nl->nbr.ptr, rows.data(), sizeof(int2) * total, hipMemcpyHostToDevice, s));
    // atoms in the caller's order: slot = atom
    if (nat) {
        AGPU_LAUNCH("nl_iota", k_iota, dim3((nat + 255) / 256), dim3(256), 0, s, nat, nl->idx_sorted.as<int>());
        AGPU_LAUNCH("nl_slots", k_slots_identity, dim3((nat + 255) / 256), dim3(256), 0, s, nat, p->d_r,
                    nl->atoms_sorted.as<SortedAtom>());
    }
    AGPU_HIP(hipStreamSynchronize(s));
    nl->external = true;
    nl->csr = true;
    nl->p = p;
    nl->nat = nat;
    nl->pos_rev = p->pos_rev;
    nl->cell_rev = p->cell_rev;
    nl->slot_rev = p->pos_rev;
    nl->total = total;
    nl->maxrow = maxrow;
    nl->cap = total + nat;
    nl->n_updates++;
    return AGPU_OK;
}

__global__ void k_export(int nat, const int *__restrict__ offsets, const int2 *__restrict__ nbr,
                         long long *__restrict__ off_out, int *__restrict__ j_out, int *__restrict__ dc_out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nat) return;
    off_out[i] = offsets[i];
    if (i == nat) return;
    for (int s = offsets[i]; s < offsets[i + 1]; ++s) {
        int2 e = nbr[s];
        j_out[s] = e.x;
        int a, b, c;
        dc_unpack(e.y, a, b, c);
        dc_out[3 * s] = a;
        dc_out[3 * s + 1] = b;
        dc_out[3 * s + 2] = c;
    }
}

int nl_export(agpu_neighbors *nl, int64_t *offsets, int *nbr, int *dc, int on_device)
{
    AGPU_CHECK(nl_compact(nl));
    hipStream_t s = stream();
    const int nat = nl->nat;
    const int64_t n = nl->total;
    long long *d_off;
    int *d_j, *d_dc;
    DevBuf tmp;
    if (on_device) {
        d_off = (long long *)offsets;
        d_j = nbr;
        d_dc = dc;
    } else {
        size_t bytes = sizeof(long long) * (nat + 1) + sizeof(int) * 4 * (n + 1);
        AGPU_CHECK(tmp.reserve(bytes));
        d_off = tmp.as<long long>();
        d_j = (int *)(d_off + nat + 1);
        d_dc = d_j + (n + 1);
    }
    AGPU_LAUNCH("nl_export", k_export, dim3((nat + 1 + 255) / 256), dim3(256), 0, s, nat,
                nl->offsets.as<int>(), nl->nbr.as<int2>(), d_off, d_j, d_dc);
    if (!on_device) {
        AGPU_HIP(hipMemcpyAsync(offsets, d_off, sizeof(long long) * (nat + 1), hipMemcpyDeviceToHost, s));
        if (n > 0) {
            AGPU_HIP(hipMemcpyAsync(nbr, d_j, sizeof(int) * n, hipMemcpyDeviceToHost, s));
            AGPU_HIP(hipMemcpyAsync(dc, d_dc, sizeof(int) * 3 * n, hipMemcpyDeviceToHost, s));
        }
    }
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

}  // namespace agpu
