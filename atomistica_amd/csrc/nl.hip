// Cell/bin neighbour-list builder for gfx950.
//
// Same list, in the same row order, as the reference's fill_neighbor_list
// (src/python/f90/python_neighbors.f90:570-754) over the bins of
// binning_init/binning_update (:765-959):
//   1. k_bin:      bin of every atom (floor(rec_cell_size.r) + 1, wrapped on
//                  periodic axes, clamped otherwise) and its wrap count.
//   2. radix sort of (bin, atom) pairs: atoms of a bin become contiguous and
//                  stay in ascending atom order -- the reference's linked
//                  lists, as arrays (Morton-free row-major bin order).
//   3. k_cell_start: first sorted slot of each bin.
//   4. k_gather:   bin-sorted copy of (r, atom, wrap) so that candidate scans
//                  read contiguous 32-byte records.
//   5. k_scan_bins<false>: per atom, count accepted (j, image) pairs over
//                  the stencil (x-major, then y, then z) -> exclusive scan.
//   6. k_scan_bins<true>:  same walk, writes (j, packed image) in order.
// Threads walk atoms in bin order, so a wavefront scans the same candidate
// bins and its loads broadcast from L1/L2.  The distance test is the
// reference's FP64 expression without contraction, so the accepted set is
// bit-identical.
#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

struct __align__(16) SortedAtom {
    double x, y, z;
    int j;       // caller index
    int w;       // packed wrap counts of its bin (dc_pack layout)
};

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

__global__ void k_bin(int nat, const double *__restrict__ r, BinGeom g, int *__restrict__ key,
                      int *__restrict__ idx, int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nat) return;
    double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    int c[3], w[3];
    bin_of(g, ri, c, w);
    if (!dc_pack_ok(w[0]) || !dc_pack_ok(w[1]) || !dc_pack_ok(w[2])) atomicOr(flags, DERR_IMAGE_RANGE);
    key[i] = ((c[0] - 1) * g.nc[1] + (c[1] - 1)) * g.nc[2] + (c[2] - 1);
    idx[i] = i;
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

__global__ void k_gather(int nat, const double *__restrict__ r, BinGeom g,
                         const int *__restrict__ idx_sorted, SortedAtom *__restrict__ out,
                         int *__restrict__ rank)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    int j = idx_sorted[p];
    double rj[3] = {r[3 * j], r[3 * j + 1], r[3 * j + 2]};
    int c[3], w[3];
    bin_of(g, rj, c, w);
    SortedAtom a;
    a.x = rj[0];
    a.y = rj[1];
    a.z = rj[2];
    a.j = j;
    a.w = dc_pack(w[0], w[1], w[2]);
    out[p] = a;
    rank[j] = p;
}

// FILL = false: count rows; FILL = true: write them.
template <bool FILL>
__global__ void __launch_bounds__(256)
k_scan_bins(int nat, BinGeom g, CellArg C, double cutoff_sq, const SortedAtom *__restrict__ atoms,
            const int *__restrict__ cell_start, int *__restrict__ counts,
            const int *__restrict__ offsets, int2 *__restrict__ nbr, long long cap,
            int *__restrict__ flags, int *__restrict__ rowmax)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    const SortedAtom ai = atoms[p];
    const int i = ai.j;
    const double ri[3] = {ai.x, ai.y, ai.z};
    int wi0, wi1, wi2;
    dc_unpack(ai.w, wi0, wi1, wi2);
    // bin of i, recomputed exactly as the reference does per atom
    int ci[3], wtmp[3];
    bin_of(g, ri, ci, wtmp);

    int out = 0, cnt = 0;
    long long base = 0;
    bool ok = true;
    if (FILL) {
        if (*flags & DERR_NL_OVERFLOW) return;
        base = offsets[i];
        cnt = counts[i];
        // serial overflow rule of python_neighbors.f90:713-718: the last
        // 1-based slot written, counting one separator per earlier atom
        if (cnt > 0 && base + cnt + (long long)i >= cap) {
            atomicOr(flags, DERR_NL_OVERFLOW);
            ok = false;
        }
        if (!ok) return;
    }
    for (int dx = -g.nd[0]; dx <= g.nd[0]; ++dx) {
        int cx = ci[0] + dx, s0 = wi0;
        if (g.pbc[0]) {
            int t = floordiv(cx - 1, g.nc[0]);
            cx -= t * g.nc[0];
            s0 -= t;
        } else if (cx < 1 || cx > g.nc[0]) continue;
        for (int dy = -g.nd[1]; dy <= g.nd[1]; ++dy) {
            int cy = ci[1] + dy, s1 = wi1;
            if (g.pbc[1]) {
                int t = floordiv(cy - 1, g.nc[1]);
                cy -= t * g.nc[1];
                s1 -= t;
            } else if (cy < 1 || cy > g.nc[1]) continue;
            for (int dz = -g.nd[2]; dz <= g.nd[2]; ++dz) {
                int cz = ci[2] + dz, s2 = wi2;
                if (g.pbc[2]) {
                    int t = floordiv(cz - 1, g.nc[2]);
                    cz -= t * g.nc[2];
                    s2 -= t;
                } else if (cz < 1 || cz > g.nc[2]) continue;
                int cell = ((cx - 1) * g.nc[1] + (cy - 1)) * g.nc[2] + (cz - 1);
                int pb = cell_start[cell], pe = cell_start[cell + 1];
                for (int q = pb; q < pe; ++q) {
                    const SortedAtom aj = atoms[q];
                    int wj0, wj1, wj2;
                    dc_unpack(aj.w, wj0, wj1, wj2);
                    int t0 = s0 - wj0, t1 = s1 - wj1, t2 = s2 - wj2;
                    if (aj.j == i && t0 == 0 && t1 == 0 && t2 == 0) continue;
                    const double rj[3] = {aj.x, aj.y, aj.z};
                    double d[3];
                    delta_r(C, ri, rj, t0, t1, t2, d);
                    double d2 = dot3_nc(d);
                    if (d2 < cutoff_sq) {
                        if (FILL) {
                            if (out >= cnt) continue;   // never taken: same walk as the count
                            if (!dc_pack_ok(t0) || !dc_pack_ok(t1) || !dc_pack_ok(t2))
                                atomicOr(flags, DERR_IMAGE_RANGE);
                            nbr[base + out] = make_int2(aj.j, dc_pack(t0, t1, t2));
                        }
                        ++out;
                    }
                }
            }
        }
    }
    if (!FILL) {
        counts[i] = out;
        atomicMax(rowmax, out);
    }
}

__global__ void k_offsets_tail(int nat, const int *__restrict__ counts, int *__restrict__ offsets,
                               int *__restrict__ stats)
{
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

int nl_update(agpu_neighbors *nl, agpu_particles *p, int force)
{
    if (!p->cell_set) return set_error(AGPU_ERR_STATE, "particles: cell not set");
    if (p->nat > 0 && !p->d_r) return set_error(AGPU_ERR_STATE, "particles: positions not set");
    if (!force && nl->p == p && nl->pos_rev == p->pos_rev && nl->cell_rev == p->cell_rev &&
        nl->nat == p->nat)
        return AGPU_OK;
    nl->cutoff = nl->interaction_range;   // verlet shell 0 (aseinterface.py:255)
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
    AGPU_CHECK(nl->rank.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->cell_start.reserve(sizeof(int) * (ncell + 1)));
    AGPU_CHECK(nl->atoms_sorted.reserve(sizeof(SortedAtom) * (nat + 1)));
    AGPU_CHECK(nl->counts.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->offsets.reserve(sizeof(int) * (nat + 1)));
    AGPU_CHECK(nl->rowmax.reserve(sizeof(int) * 4));
    AGPU_CHECK(nl->nbr.reserve(sizeof(int2) * (nl->cap > 0 ? nl->cap : 1)));
    int *flags = nl->rowmax.as<int>();   // [0] rowmax, [1..2] total, [3] flags
    AGPU_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 4, s));
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const int TB = 256;
    if (nat > 0) {
        int nb = (nat + TB - 1) / TB;
        AGPU_LAUNCH("nl_bin", k_bin, dim3(nb), dim3(TB), 0, s, nat, p->d_r, g,
                    nl->key.as<int>(), nl->idx.as<int>(), flags + 3);
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
        AGPU_LAUNCH("nl_gather", k_gather, dim3(nb), dim3(TB), 0, s, nat, p->d_r, g,
                    nl->idx_sorted.as<int>(), nl->atoms_sorted.as<SortedAtom>(), nl->rank.as<int>());
        double cutoff_sq = nl->cutoff * nl->cutoff;
        AGPU_LAUNCH("nl_count", k_scan_bins<false>, dim3(nb), dim3(TB), 0, s, nat, g, C, cutoff_sq,
                    nl->atoms_sorted.as<SortedAtom>(), nl->cell_start.as<int>(), nl->counts.as<int>(),
                    (const int *)nullptr, (int2 *)nullptr, (long long)nl->cap, flags + 3, flags);
        size_t tmp2 = 0;
        AGPU_HIP(rocprim::exclusive_scan(nullptr, tmp2, nl->counts.as<int>(), nl->offsets.as<int>(), 0,
                                         (size_t)nat, rocprim::plus<int>(), s));
        AGPU_CHECK(nl->sort_tmp.reserve(tmp2 + 16));
        AGPU_HIP(rocprim::exclusive_scan(nl->sort_tmp.ptr, tmp2, nl->counts.as<int>(),
                                         nl->offsets.as<int>(), 0, (size_t)nat, rocprim::plus<int>(), s));
        AGPU_LAUNCH("nl_offsets_tail", k_offsets_tail, dim3(1), dim3(64), 0, s, nat,
                    nl->counts.as<int>(), nl->offsets.as<int>(), flags);
        AGPU_LAUNCH("nl_fill", k_scan_bins<true>, dim3(nb), dim3(TB), 0, s, nat, g, C, cutoff_sq,
                    nl->atoms_sorted.as<SortedAtom>(), nl->cell_start.as<int>(), nl->counts.as<int>(),
                    nl->offsets.as<int>(), nl->nbr.as<int2>(), (long long)nl->cap, flags + 3,
                    (int *)nullptr);
    } else {
        AGPU_HIP(hipMemsetAsync(nl->offsets.ptr, 0, sizeof(int), s));
    }
    AGPU_HIP(hipMemcpyAsync(nl->stats.h, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    int fl = nl->stats.h[3];
    nl->maxrow = nl->stats.h[0];
    nl->total = (int64_t)nl->stats.h[1] | ((int64_t)nl->stats.h[2] << 31);
    nl->p = p;
    nl->nat = nat;
    nl->pos_rev = p->pos_rev;
    nl->cell_rev = p->cell_rev;
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
