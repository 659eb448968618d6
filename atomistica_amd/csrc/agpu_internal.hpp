// Internal declarations shared by the HIP translation units of
// libatomistica_amd: handle layouts, device helpers and error plumbing.
// Nothing here is part of the C-ABI (see include/atomistica_amd.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/atomistica_amd.h"

namespace agpu {

// ---------------------------------------------------------------------------
// errors: the last message per host thread, returned by agpu_get_error()
// ---------------------------------------------------------------------------
int set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define AGPU_HIP(call)                                                       \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) return agpu::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define AGPU_CHECK(expr)                                                     \
    do {                                                                     \
        int rc_ = (expr);                                                    \
        if (rc_ != AGPU_OK) return rc_;                                      \
    } while (0)

// device-side error bits, OR-ed into a per-call flag word
enum DevErr : int {
    DERR_NL_OVERFLOW = 1,
    DERR_IMAGE_RANGE = 2,
    DERR_SPLINE_RANGE = 4,
    DERR_BOND_OVERFLOW = 8,
    DERR_STAGE_OVERFLOW = 16,
    DERR_FIXED_RANGE = 32,
    DERR_SORT_FALLBACK = 64,
    DERR_SCR_ROW = 128,           // screened bond order: row beyond the staged length
    // a list build whose rows must not be used (nl_settle redoes or refuses it)
    DERR_NL_UNUSABLE = DERR_NL_OVERFLOW | DERR_IMAGE_RANGE | DERR_STAGE_OVERFLOW | DERR_SORT_FALLBACK,
    // a pass run ahead of its checks found it must be redone (an unusable
    // list build, a row longer than its lane group): later passes write
    // nothing to the caller's outputs
    DERR_REDO = 256,
};

hipStream_t stream();

// Grow-only device buffer owned by a handle.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t nbytes);
    void release();
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
    ~DevBuf() { release(); }
};

// Pinned host staging word for flags and small statistics.
struct HostStats {
    int *h = nullptr;
    int init();
    ~HostStats();
};

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct Cell {
    double A[9];     // A[3*k + x] = component x of lattice vector k
    double B[9];     // B[3*k + x] = Bbox(x, k): column-major image of A^-1
    int pbc[3];
};

}  // namespace agpu

struct agpu_particles {
    int nat = 0;
    int nalloc = 0;
    agpu::Cell cell{};
    int cell_set = 0;
    double *d_r = nullptr;        // (nat, 3) device positions
    int *d_Z = nullptr;           // (nat) device atomic numbers
    bool own_r = true, own_Z = true;
    agpu::DevBuf r_buf, Z_buf;
    uint64_t pos_rev = 0;         // bumped on every position change
    uint64_t cell_rev = 0;
    uint64_t Z_rev = 0;
    std::vector<int> h_Z;         // host copy of atomic numbers
    // LAMMPS step: global atom IDs of the local atoms and ghosts and the
    // number of local atoms (the REBO2 half rule tag(j) >= tag(i) over
    // local i, bop_kernel_rebo2.f90:1328-1333); natloc < 0 when unset
    agpu::DevBuf d_tag;
    int natloc = -1;
};

struct agpu_neighbors {
    int avgn = 100;
    double interaction_range = 0.0;
    double cutoff = 0.0;
    double bin_size = 0.0;        // <= 0: equal to the cutoff
    // geometry of the last build
    int nc[3] = {0, 0, 0};
    int nd[3] = {0, 0, 0};
    double rec[9];                // rec_cell_size, column-major image
    // provenance of the last build
    const agpu_particles *p = nullptr;
    uint64_t pos_rev = ~0ull, cell_rev = ~0ull;
    int nat = 0;
    int64_t cap = 0;              // nat * avgn entries
    int64_t total = 0;            // neighbours in the list
    int maxrow = 0;
    // device storage
    agpu::DevBuf key, key_sorted, idx, idx_sorted, cell_start, atoms_sorted,
        counts, offsets, nbr, sort_tmp, rowmax, stage, pos32, part, scnt, soffs;
    int stage_w = 0;              // staging stride of the next build (entries per atom)
    int stage_s = 0;              // staging stride of the current rows
    // particles' pos_rev when the slot positions (atoms_sorted) were written
    uint64_t slot_rev = ~0ull;
    // offsets holds the caller-order row offsets of the last build
    bool caller_off = false;
    // a build whose flags are on their way to the host and not yet checked
    // (nl_update with defer): nl_settle checks them after the caller's sync
    bool pending = false;
    bool radix = false;           // bin sort by radix sort (after a very full bin)
    // the list was imported from the caller (LAMMPS) and is not rebuilt
    bool external = false;
    // Verlet shell added to the interaction range, and a hold flag: while
    // held, the list is kept although positions move (the MD driver tracks
    // the displacements against the shell and rebuilds)
    double skin = 0.0;
    bool held = false;
    // pair search kernel: 0 chosen from the bin occupancy, 1 thread per
    // atom, 2 wavefront per bin block (both give the same rows)
    int pair_mode = 0;
    // the CSR rows (offsets, nbr: caller indices) hold the last build; false
    // while its rows are only staged (slot p's row at stage + p*stage_s,
    // scnt[p] entries naming slots) because every caller since the build
    // read them there
    bool csr = true;
    agpu::DevBuf cell_cnt;
    long long cnt_zero = 0;       // leading cell_cnt entries known to be zero
    agpu::HostStats stats;
    int64_t n_updates = 0;
};

namespace agpu {

// Rebuilds the list when positions, cell or atoms changed (force > 0:
// always).  csr = false: the caller reads the rows through NlRows only, so
// the copy of staged rows into the CSR rows is left until a caller asks.
int nl_update(agpu_neighbors *nl, agpu_particles *p, int force, bool csr = true, bool defer = false);
// defer = true (staged rows only, not under capture): a build returns
// without waiting for its flags; the caller's kernels test the device flag
// word (nl_flags) and write nothing when it marks the build unusable, and
// after its own stream synchronisation the caller calls nl_settle, which
// checks the flags, records the list size and sets *redo when the build
// has to be redone (the caller then rebuilds and runs again)
int nl_settle(agpu_neighbors *nl, bool *redo);
const int *nl_flags(const agpu_neighbors *nl);
// true when nl_update would keep the list as it is (no rebuild, nothing
// allocated): imported, held by a driver, or built from these positions
bool nl_current(const agpu_neighbors *nl, const agpu_particles *p);
// graph-capture mode (no host synchronisation, no memset nodes)
bool capturing();
void set_capturing(bool on);
int zero_dev(void *ptr, size_t nbytes);
int nl_import(agpu_neighbors *nl, agpu_particles *p, const int64_t *seed, const int *numneigh, const int *neigh);
int nl_export(agpu_neighbors *nl, int64_t *offsets, int *nbr, int *dc, int on_device);
int compute_inverse(const double *A, double *B);

// ---- device helpers ----
struct CellArg {
    double A[9];
};

// An atom of the list's bin order ("slot"): its position as the caller gave
// it, its caller index and the packed wrap counts of its bin.
struct __align__(16) SortedAtom {
    double x, y, z;
    int j;       // caller index
    int w;       // packed wrap counts of its bin (dc_pack layout)
};

// The rows of the current list in bin-order storage.  Atoms are named by
// slot (caller atom order[p] at slot p), in the rows as well: an entry is
// {slot of the neighbour, packed image}.  Row x has its entries at
// rows + x*S (cnt[x] of them) when S > 0 (the builder's staging rows), else
// at rows + off[x] (off[x + 1] - off[x] of them: CSR by slot, an imported
// list, whose slots are the caller's atoms).  off is the exclusive scan of
// the row lengths in slot order in both cases: entry k of row x is entry
// off[x] + k of the per-entry arrays, which are dense and ordered by slot.
// atoms[x] holds slot x's position, read where a kernel gathers neighbours.
struct NlRows {
    const int2 *rows;
    const int *cnt;
    const int *off;
    int S;
    const SortedAtom *atoms;
};
// the rows of the current list (nl_update brings the slot positions up to
// date before it returns)
NlRows nl_rows(const agpu_neighbors *nl);
// the slot positions brought up to date without a rebuild (for callers that
// do not go through nl_update)
int nl_slots_current(agpu_neighbors *nl, const agpu_particles *p);
// offsets of the caller-order rows (the reference's list layout: per-bond
// arrays are indexed by them); computed on the first request after a build
int nl_caller_offsets(agpu_neighbors *nl);

__device__ __forceinline__ const int2 *nl_row(const NlRows &R, int x, int &n)
{
    const int b = R.off[x];
    if (R.S > 0) {
        n = R.cnt[x];
        return R.rows + (size_t)x * R.S;
    }
    n = R.off[x + 1] - b;
    return R.rows + b;
}

// index of entry k of row x in the per-entry arrays
__device__ __forceinline__ size_t nl_at(const NlRows &R, int x, int k) { return (size_t)R.off[x] + k; }

__device__ __forceinline__ int dc_pack_ok(int a) { return a >= -512 && a <= 511; }
__device__ __forceinline__ int dc_pack(int a, int b, int c)
{
    return ((a + 512) << 20) | ((b + 512) << 10) | (c + 512);
}
__device__ __forceinline__ void dc_unpack(int code, int &a, int &b, int &c)
{
    a = ((code >> 20) & 1023) - 512;
    b = ((code >> 10) & 1023) - 512;
    c = (code & 1023) - 512;
}

// A.s summed k = 0..2 in the order the reference's matmul writes it
__device__ __forceinline__ double image_x(const CellArg &C, int x, int s0, int s1, int s2)
{
#pragma clang fp contract(off)
    double as = 0.0;
    as = as + C.A[0 + x] * (double)s0;
    as = as + C.A[3 + x] * (double)s1;
    as = as + C.A[6 + x] * (double)s2;
    return as;
}

// dr = r_i - r_j + A.dc (delta_r of python_neighbors.f90:708, GET_DRJ of
// macros.inc:76), non-contracted
__device__ __forceinline__ void delta_r(const CellArg &C, const double *ri, const double *rj,
                                        int s0, int s1, int s2, double *d)
{
#pragma clang fp contract(off)
    for (int x = 0; x < 3; ++x) d[x] = (ri[x] - rj[x]) + image_x(C, x, s0, s1, s2);
}

__device__ __forceinline__ double dot3_nc(const double *d)
{
#pragma clang fp contract(off)
    double s = 0.0;
    s = s + d[0] * d[0];
    s = s + d[1] * d[1];
    s = s + d[2] * d[2];
    return s;
}

// The work block of this workgroup when consecutive work blocks share data
// (neighbouring atoms in bin order): workgroups go to the 8 XCDs in turn and
// each XCD has its own L2, so the first 8 floor(n / 8) workgroups are dealt
// out so that XCD x takes one contiguous eighth of them; the rest keep their
// index.  A bijection on [0, n).
__device__ __forceinline__ int xcd_block()
{
    const int n = gridDim.x, b = blockIdx.x;
    const int per = n / 8;
    if (b >= 8 * per) return b;
    return (b % 8) * per + b / 8;
}

// warp-level (wave64) helpers
template <typename T> __device__ __forceinline__ T shfl_xor(T v, int m)
{
    return __shfl_xor(v, m, 64);
}

template <int G, typename T> __device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

}  // namespace agpu

// potential handle (one struct for every class; kind selects the kernels)
struct agpu_potential {
    int kind = 0;                 // AGPU_KIND_*
    int precision = AGPU_PREC_FP64;
    int bound = 0;
    double cutoff = 0.0;          // interaction range requested at bind_to
    agpu_particles *p = nullptr;
    agpu_neighbors *nl = nullptr;
    void *impl = nullptr;         // class-specific state
    void (*free_impl)(void *) = nullptr;
    agpu::DevBuf d_el, flags, partial, partial_out;
    agpu::HostStats stats;
};

namespace agpu {
// class hooks
int tersoff_bind(agpu_potential *pot);
int tersoff_compute(agpu_potential *pot, const agpu_eval_args *a);
// true for a bond-order potential with screening (its records are sized on
// the host every call, so it cannot run under stream capture)
bool tersoff_screened(const agpu_potential *pot);
int eam_bind(agpu_potential *pot);
int eam_compute(agpu_potential *pot, const agpu_eval_args *a);
// selects launch variant k (lanes per atom g, entries per lane u) when k is
// in range and returns the number of variants (0: this binding has none)
int eam_variant(agpu_potential *pot, int k, int *g, int *u);
// unscreened bond-order classes: k = -1 returns the number of variants (2;
// 0 for other classes), k = 0 selects the row path (a lane per row entry,
// reverse entries found in the neighbour's row), k = 1 the bond-compacting
// path (bond slots, reverse-slot pass, slot gather)
int bop_variant(agpu_potential *pot, int k);
int ewald_bind(agpu_potential *pot);
int ewald_compute(agpu_potential *pot, const agpu_eval_args *a);
int ewald_potential(agpu_potential *pot, const double *q, double *phi, int flags);
// smooth PME reciprocal sum (pme.hip): rec10 := [E, W column-major], f +=
// q E, phi += phi_rec (f, phi may be NULL)
struct PmeState;
PmeState *pme_state_new(int order, const int *grid);
void pme_state_free(PmeState *pm);
int pme_recip(PmeState *pm, const agpu_particles *p, const double *q, double beta, double *rec10, double *f,
              double *phi);
int rebo2_bind(agpu_potential *pot);
int rebo2_compute(agpu_potential *pot, const agpu_eval_args *a);
// REBO2 launch variants: k = 0, 1, 2 select 4, 8 or 16 lanes per half pair
// in the torsion pass (the per-pair sums change their association order);
// returns the number of variants (3), *g (optional) the current width
int rebo2_variant(agpu_potential *pot, int k, int *g);

// Deterministic fixed-order reduction of nblk per-block records of width w
// (double): column 0 -> *epot (+=, or = with set_epot), columns 1..9 ->
// wpot[0..8] (+=), columns 10.. -> extra[0..] (+=).  Runs on the library
// stream.
// tmp: room for the second-level records, ceil(nblk / 256) * w doubles
// (NULL: the library's own buffer, grown on demand).
int reduce_ev(const double *partials, int nblk, int w, double *epot, double *wpot, double *extra,
              bool set_epot = false, double *tmp = nullptr, const int *skip = nullptr);

// *out = max(*out, max(v[0..n))) on the library stream (device pointers)
int int_max(const int *v, int n, int *out);

// map atomic numbers to database element indices (-1 = not handled)
int map_elements(agpu_potential *pot, const int *z2db, int nz);
}  // namespace agpu
