// C-ABI entry points of libatomistica_amd and the small runtime behind them:
// per-thread error messages, the launch stream, grow-only device buffers,
// per-kernel event timing, particles / neighbours handles and the
// potential dispatch of init / bind_to / energy_and_forces.
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

static thread_local std::string g_err;
static hipStream_t g_stream = nullptr;
static bool g_own_stream = false;

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
                  hipGetErrorString(e), what, file, line);
    return set_error(AGPU_ERR_HIP, buf);
}

hipStream_t stream()
{
    if (!g_stream) {
        if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) == hipSuccess)
            g_own_stream = true;
        else
            g_stream = nullptr;
    }
    return g_stream;
}

int DevBuf::reserve(size_t nbytes)
{
    if (nbytes <= bytes && ptr) return AGPU_OK;
    release();
    size_t n = nbytes < 256 ? 256 : nbytes;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    bytes = n;
    return AGPU_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int HostStats::init()
{
    if (h) return AGPU_OK;
    void *p = nullptr;
    AGPU_HIP(hipHostMalloc(&p, 64 * sizeof(int), hipHostMallocDefault));
    h = static_cast<int *>(p);
    std::memset(h, 0, 64 * sizeof(int));
    return AGPU_OK;
}

HostStats::~HostStats()
{
    if (h) (void)hipHostFree(h);
}

// ---------------------------------------------------------------------------
// per-kernel event timing
// ---------------------------------------------------------------------------
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
static bool g_prof = false;
static std::vector<ProfRec> g_prof_recs;
static std::map<std::string, std::pair<double, int>> g_prof_done;

// while a step is being captured into a graph: no host synchronisation, no
// memset nodes, no profiling events (set by the MD driver)
static bool g_capture = false;
bool capturing() { return g_capture; }
void set_capturing(bool on) { g_capture = on; }

bool profiling() { return g_prof && !g_capture; }

__global__ void k_zero_words(unsigned *__restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// zero a device range with a kernel (a fill kernel records as a kernel node
// when captured; nbytes a multiple of 4)
int zero_dev(void *ptr, size_t nbytes)
{
    const size_t n = nbytes / 4;
    if (n == 0) return AGPU_OK;
    const unsigned nb = (unsigned)std::min<size_t>(1024, (n + 255) / 256);
    AGPU_LAUNCH("zero", k_zero_words, dim3(nb), dim3(256), 0, stream(), (unsigned *)ptr, n);
    return AGPU_OK;
}

int prof_begin(const char *name, hipStream_t s)
{
    ProfRec r;
    r.name = name;
    if (hipEventCreate(&r.a) != hipSuccess) return -1;
    if (hipEventCreate(&r.b) != hipSuccess) {
        (void)hipEventDestroy(r.a);
        return -1;
    }
    (void)hipEventRecord(r.a, s);
    g_prof_recs.push_back(r);
    return (int)g_prof_recs.size() - 1;
}

void prof_end(int token, hipStream_t s)
{
    if (token >= 0 && token < (int)g_prof_recs.size()) (void)hipEventRecord(g_prof_recs[token].b, s);
}

static void prof_collect()
{
    for (auto &r : g_prof_recs) {
        (void)hipEventSynchronize(r.b);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &d = g_prof_done[r.name];
            d.first += ms;
            d.second += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
}

// ---------------------------------------------------------------------------
// reductions and element maps
// ---------------------------------------------------------------------------
// Stage 1: block b sums records [b*chunk, min(nblk, (b+1)*chunk)) column by
// column; every thread walks a fixed record stride, then a fixed LDS tree.
// Stage 2 (one block) sums the stage-1 records the same way and adds them to
// the outputs.  The grouping depends only on nblk, so results are bitwise
// reproducible.
constexpr int RED_W_MAX = 16;

__global__ void __launch_bounds__(256)
k_reduce_rec(const double *__restrict__ part, int nblk, int w, int chunk,
             double *__restrict__ out, double *__restrict__ epot, double *__restrict__ wpot,
             double *__restrict__ extra, int set_epot, const int *__restrict__ skip)
{
    // a call that will be redone leaves the outputs as they are
    if (skip && (*skip & DERR_REDO)) return;
    __shared__ double sh[256];
    const int lanes = 256 / w;         // records read per sweep
    const int c = threadIdx.x % w;
    const int rr = threadIdx.x / w;
    const int b0 = blockIdx.x * chunk;
    const int b1 = min(nblk, b0 + chunk);
    double s = 0.0;
    if (rr < lanes)
        for (int b = b0 + rr; b < b1; b += lanes) s += part[(size_t)b * w + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < w) {
        double t = 0.0;
        for (int k = 0; k < lanes; ++k) t += sh[k * w + threadIdx.x];
        if (out) {
            out[(size_t)blockIdx.x * w + threadIdx.x] = t;
        } else {
            int col = threadIdx.x;
            if (col == 0) {
                if (epot) *epot = set_epot ? t : *epot + t;
            } else if (col < 10) {
                if (wpot) wpot[col - 1] += t;
            } else if (extra) {
                extra[col - 10] += t;
            }
        }
    }
}

// process-lifetime scratch (never freed: no device frees from static destructors)
static DevBuf *g_red_tmp = new DevBuf();

int reduce_ev(const double *partials, int nblk, int w, double *epot, double *wpot, double *extra, bool set_epot,
              double *tmp, const int *skip)
{
    if (nblk <= 0 || w <= 0) return AGPU_OK;
    if (w > RED_W_MAX) return set_error(AGPU_ERR_ARG, "reduce_ev: record too wide");
    hipStream_t s = stream();
    const int chunk = 256;
    int n1 = (nblk + chunk - 1) / chunk;
    if (n1 <= 1) {
        AGPU_LAUNCH("reduce_ev", k_reduce_rec, dim3(1), dim3(256), 0, s, partials, nblk, w, nblk,
                    (double *)nullptr, epot, wpot, extra, (int)set_epot, skip);
        return AGPU_OK;
    }
    if (!tmp) {
        AGPU_CHECK(g_red_tmp->reserve(sizeof(double) * (size_t)n1 * w));
        tmp = g_red_tmp->as<double>();
    }
    AGPU_LAUNCH("reduce_ev", k_reduce_rec, dim3(n1), dim3(256), 0, s, partials, nblk, w, chunk,
                tmp, (double *)nullptr, (double *)nullptr, (double *)nullptr, 0, (const int *)nullptr);
    AGPU_LAUNCH("reduce_ev", k_reduce_rec, dim3(1), dim3(256), 0, s, (const double *)tmp, n1, w,
                n1, (double *)nullptr, epot, wpot, extra, (int)set_epot, skip);
    return AGPU_OK;
}

// max over n ints into *out (atomicMax, one atomic per block of 4096 values)
__global__ void __launch_bounds__(256) k_int_max(const int *__restrict__ v, int n, int *__restrict__ out)
{
    __shared__ int sm[4];
    int m = 0;
    const int base = blockIdx.x * 4096;
    for (int k = threadIdx.x; k < 4096; k += 256)
        if (base + k < n) m = max(m, v[base + k]);
    for (int o = 32; o >= 1; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(sm[0], sm[1]), max(sm[2], sm[3])));
}

int int_max(const int *v, int n, int *out)
{
    if (n <= 0) return AGPU_OK;
    AGPU_LAUNCH("int_max", k_int_max, dim3((n + 4095) / 4096), dim3(256), 0, stream(), v, n, out);
    return AGPU_OK;
}

__global__ void k_map_el(int nat, const int *__restrict__ Z, const int *__restrict__ z2db, int nz,
                         int *__restrict__ el)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nat) return;
    int z = Z[i];
    el[i] = (z >= 0 && z < nz) ? z2db[z] : -1;
}

int map_elements(agpu_potential *pot, const int *z2db, int nz)
{
    agpu_particles *p = pot->p;
    AGPU_CHECK(pot->d_el.reserve(sizeof(int) * (p->nat + 1) + sizeof(int) * nz));
    int *d_el = pot->d_el.as<int>();
    int *d_map = d_el + p->nat + 1;
    hipStream_t s = stream();
    AGPU_HIP(hipMemcpyAsync(d_map, z2db, sizeof(int) * nz, hipMemcpyHostToDevice, s));
    if (p->nat > 0)
        AGPU_LAUNCH("map_elements", k_map_el, dim3((p->nat + 255) / 256), dim3(256), 0, s, p->nat,
                    p->d_Z, d_map, nz, d_el);
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

int compute_inverse(const double *cell, double *bbox)
{
    // Partial-pivot LU solve of Abox X = I (the gaussn -> dgesv call of
    // python_particles.f90:331); Abox(x,k) = cell[3k+x], same image for X.
    double a[9], b[9];
    int ipiv[3];
    std::memcpy(a, cell, sizeof a);
    for (int k = 0; k < 9; ++k) b[k] = 0.0;
    b[0] = b[4] = b[8] = 1.0;
#define AX(m, x, k) (m)[3 * (k) + (x)]
    for (int j = 0; j < 3; ++j) {
        int p = j;
        double amax = std::fabs(AX(a, j, j));
        for (int i = j + 1; i < 3; ++i)
            if (std::fabs(AX(a, i, j)) > amax) {
                amax = std::fabs(AX(a, i, j));
                p = i;
            }
        ipiv[j] = p;
        if (AX(a, p, j) == 0.0) return set_error(AGPU_ERR_ARG, "singular cell");
        if (p != j)
            for (int k = 0; k < 3; ++k) std::swap(AX(a, j, k), AX(a, p, k));
        double rp = 1.0 / AX(a, j, j);
        for (int i = j + 1; i < 3; ++i) AX(a, i, j) = AX(a, i, j) * rp;
        for (int k = j + 1; k < 3; ++k)
            for (int i = j + 1; i < 3; ++i) AX(a, i, k) = AX(a, i, k) - AX(a, i, j) * AX(a, j, k);
    }
    for (int j = 0; j < 3; ++j)
        if (ipiv[j] != j)
            for (int c = 0; c < 3; ++c) std::swap(AX(b, j, c), AX(b, ipiv[j], c));
    for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < 3; ++k)
            if (AX(b, k, c) != 0.0)
                for (int i = k + 1; i < 3; ++i) AX(b, i, c) = AX(b, i, c) - AX(b, k, c) * AX(a, i, k);
        for (int k = 2; k >= 0; --k)
            if (AX(b, k, c) != 0.0) {
                AX(b, k, c) = AX(b, k, c) / AX(a, k, k);
                for (int i = 0; i < k; ++i) AX(b, i, c) = AX(b, i, c) - AX(b, k, c) * AX(a, i, k);
            }
    }
#undef AX
    std::memcpy(bbox, b, sizeof b);
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

// ===========================================================================
// library / device
// ===========================================================================
extern "C" {

const char *agpu_version(void) { return "atomistica_amd 0.1 (gfx950)"; }

int agpu_get_error(char *buf, int len)
{
    if (!buf || len <= 0) return AGPU_ERR_ARG;
    std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return AGPU_OK;
}

int agpu_device_count(int *n)
{
    if (!n) return set_error(AGPU_ERR_ARG, "n is NULL");
    AGPU_HIP(hipGetDeviceCount(n));
    return AGPU_OK;
}

int agpu_set_device(int device)
{
    AGPU_HIP(hipSetDevice(device));
    return AGPU_OK;
}

int agpu_set_stream(void *s)
{
    if (g_own_stream && g_stream) {
        (void)hipStreamSynchronize(g_stream);
        (void)hipStreamDestroy(g_stream);
    }
    g_own_stream = false;
    g_stream = static_cast<hipStream_t>(s);
    if (!g_stream) stream();
    return AGPU_OK;
}

int agpu_get_stream(void **s)
{
    if (!s) return set_error(AGPU_ERR_ARG, "s is NULL");
    *s = (void *)stream();
    return AGPU_OK;
}

int agpu_synchronize(void)
{
    AGPU_HIP(hipStreamSynchronize(stream()));
    return AGPU_OK;
}

int agpu_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return set_error(AGPU_ERR_ARG, "ptr is NULL");
    AGPU_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return AGPU_OK;
}

int agpu_free(void *ptr)
{
    if (ptr) AGPU_HIP(hipFree(ptr));
    return AGPU_OK;
}

int agpu_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                      : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    AGPU_HIP(hipMemcpyAsync(dst, src, bytes, k, stream()));
    AGPU_HIP(hipStreamSynchronize(stream()));
    return AGPU_OK;
}

int agpu_profile_enable(int on)
{
    if (!on && g_prof) prof_collect();
    g_prof = on != 0;
    return AGPU_OK;
}

int agpu_profile_reset(void)
{
    prof_collect();
    g_prof_done.clear();
    return AGPU_OK;
}

int agpu_profile_list(char *buf, int len)
{
    if (!buf || len <= 0) return set_error(AGPU_ERR_ARG, "profile_list: bad buffer");
    prof_collect();
    std::string s;
    for (auto &kv : g_prof_done) {
        if (!s.empty()) s += ",";
        s += kv.first;
    }
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
    return AGPU_OK;
}

int agpu_profile_query(const char *kernel, double *total_ms, int *count)
{
    prof_collect();
    auto it = g_prof_done.find(kernel ? kernel : "");
    if (total_ms) *total_ms = it == g_prof_done.end() ? 0.0 : it->second.first;
    if (count) *count = it == g_prof_done.end() ? 0 : it->second.second;
    return AGPU_OK;
}

// ===========================================================================
// particles
// ===========================================================================
int agpu_particles_create(agpu_particles **p)
{
    if (!p) return set_error(AGPU_ERR_ARG, "p is NULL");
    *p = new agpu_particles();
    return AGPU_OK;
}

int agpu_particles_destroy(agpu_particles *p)
{
    delete p;
    return AGPU_OK;
}

int agpu_particles_allocate(agpu_particles *p, int nat)
{
    if (!p || nat < 0) return set_error(AGPU_ERR_ARG, "particles_allocate: bad arguments");
    AGPU_CHECK(p->r_buf.reserve(sizeof(double) * 3 * (size_t)(nat + 1)));
    AGPU_CHECK(p->Z_buf.reserve(sizeof(int) * (size_t)(nat + 1)));
    p->d_r = p->r_buf.as<double>();
    p->d_Z = p->Z_buf.as<int>();
    p->own_r = p->own_Z = true;
    p->nat = nat;
    p->nalloc = nat;
    p->h_Z.assign(nat, 0);
    AGPU_HIP(hipMemsetAsync(p->d_Z, 0, sizeof(int) * (nat + 1), stream()));
    AGPU_HIP(hipMemsetAsync(p->d_r, 0, sizeof(double) * 3 * (nat + 1), stream()));
    AGPU_HIP(hipStreamSynchronize(stream()));
    p->pos_rev++;
    p->Z_rev++;
    return AGPU_OK;
}

int agpu_particles_set_cell(agpu_particles *p, const double *cell, const int *pbc)
{
    if (!p || !cell) return set_error(AGPU_ERR_ARG, "particles_set_cell: bad arguments");
    std::memcpy(p->cell.A, cell, sizeof p->cell.A);
    AGPU_CHECK(compute_inverse(cell, p->cell.B));
    for (int k = 0; k < 3; ++k) p->cell.pbc[k] = pbc ? (pbc[k] != 0) : 1;
    p->cell_set = 1;
    p->cell_rev++;
    return AGPU_OK;
}

int agpu_particles_set_atomic_numbers(agpu_particles *p, const int *Z, int flags)
{
    if (!p || (!Z && p->nat > 0)) return set_error(AGPU_ERR_ARG, "set_atomic_numbers: bad arguments");
    if (p->nat == 0) return AGPU_OK;
    if (!p->own_Z || !p->d_Z) {
        AGPU_CHECK(p->Z_buf.reserve(sizeof(int) * (size_t)(p->nat + 1)));
        p->d_Z = p->Z_buf.as<int>();
        p->own_Z = true;
    }
    hipStream_t s = stream();
    if (flags & AGPU_HOST) {
        p->h_Z.assign(Z, Z + p->nat);
        AGPU_HIP(hipMemcpyAsync(p->d_Z, Z, sizeof(int) * p->nat, hipMemcpyHostToDevice, s));
    } else {
        p->h_Z.resize(p->nat);
        AGPU_HIP(hipMemcpyAsync(p->d_Z, Z, sizeof(int) * p->nat, hipMemcpyDeviceToDevice, s));
        AGPU_HIP(hipMemcpyAsync(p->h_Z.data(), Z, sizeof(int) * p->nat, hipMemcpyDeviceToHost, s));
    }
    AGPU_HIP(hipStreamSynchronize(s));
    p->Z_rev++;
    return AGPU_OK;
}

int agpu_particles_set_positions(agpu_particles *p, const double *r, int flags)
{
    if (!p || (!r && p->nat > 0)) return set_error(AGPU_ERR_ARG, "set_positions: bad arguments");
    if (!p->own_r || !p->d_r) {
        AGPU_CHECK(p->r_buf.reserve(sizeof(double) * 3 * (size_t)(p->nat + 1)));
        p->d_r = p->r_buf.as<double>();
        p->own_r = true;
    }
    if (p->nat > 0) {
        hipStream_t s = stream();
        AGPU_HIP(hipMemcpyAsync(p->d_r, r, sizeof(double) * 3 * p->nat,
                                (flags & AGPU_HOST) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
        if (flags & AGPU_HOST) AGPU_HIP(hipStreamSynchronize(s));
    }
    p->pos_rev++;
    return AGPU_OK;
}

int agpu_particles_set_position_pointer(agpu_particles *p, int nat, double *r_device)
{
    if (!p || nat != p->nat || (!r_device && nat > 0))
        return set_error(AGPU_ERR_ARG, "set_position_pointer: nat differs from allocate() or NULL pointer");
    p->d_r = r_device;
    p->own_r = false;
    p->pos_rev++;
    return AGPU_OK;
}

int agpu_particles_changed_positions(agpu_particles *p)
{
    if (!p) return set_error(AGPU_ERR_ARG, "p is NULL");
    p->pos_rev++;
    return AGPU_OK;
}

int agpu_particles_get_nat(agpu_particles *p, int *nat)
{
    if (!p || !nat) return set_error(AGPU_ERR_ARG, "bad arguments");
    *nat = p->nat;
    return AGPU_OK;
}

// ===========================================================================
// neighbours
// ===========================================================================
int agpu_neighbors_create(agpu_neighbors **nl, int avgn)
{
    if (!nl) return set_error(AGPU_ERR_ARG, "nl is NULL");
    *nl = new agpu_neighbors();
    (*nl)->avgn = avgn > 0 ? avgn : 100;
    return AGPU_OK;
}

int agpu_neighbors_destroy(agpu_neighbors *nl)
{
    delete nl;
    return AGPU_OK;
}

int agpu_neighbors_request_interaction_range(agpu_neighbors *nl, double cutoff)
{
    if (!nl) return set_error(AGPU_ERR_ARG, "nl is NULL");
    if (cutoff > nl->interaction_range) {
        nl->interaction_range = cutoff;
        nl->p = nullptr;   // force a rebuild
    }
    return AGPU_OK;
}

int agpu_neighbors_set_bin_size(agpu_neighbors *nl, double bin_size)
{
    if (!nl) return set_error(AGPU_ERR_ARG, "nl is NULL");
    nl->bin_size = bin_size;
    nl->p = nullptr;
    return AGPU_OK;
}

int agpu_neighbors_update(agpu_neighbors *nl, agpu_particles *p)
{
    if (!nl || !p) return set_error(AGPU_ERR_ARG, "neighbors_update: NULL handle");
    return nl_update(nl, p, 0);
}

int agpu_neighbors_get_size(agpu_neighbors *nl, int64_t *total, int *maxrow)
{
    if (!nl) return set_error(AGPU_ERR_ARG, "nl is NULL");
    if (total) *total = nl->total;
    if (maxrow) *maxrow = nl->maxrow;
    return AGPU_OK;
}

int agpu_neighbors_get(agpu_neighbors *nl, int64_t *offsets, int *neighbors, int *dc, int flags)
{
    if (!nl || !offsets || (nl->total > 0 && (!neighbors || !dc)))
        return set_error(AGPU_ERR_ARG, "neighbors_get: bad arguments");
    if (!nl->p) return set_error(AGPU_ERR_STATE, "neighbors_get: list not built");
    return nl_export(nl, offsets, neighbors, dc, !(flags & AGPU_HOST));
}

// ===========================================================================
// potentials
// ===========================================================================
int agpu_potential_destroy(agpu_potential *pot)
{
    if (!pot) return AGPU_OK;
    if (pot->free_impl && pot->impl) pot->free_impl(pot->impl);
    delete pot;
    return AGPU_OK;
}

int agpu_potential_kind(agpu_potential *pot, int *kind)
{
    if (!pot || !kind) return set_error(AGPU_ERR_ARG, "bad arguments");
    *kind = pot->kind;
    return AGPU_OK;
}

int agpu_potential_get_cutoff(agpu_potential *pot, double *cutoff)
{
    if (!pot || !cutoff) return set_error(AGPU_ERR_ARG, "bad arguments");
    *cutoff = pot->cutoff;
    return AGPU_OK;
}

int agpu_potential_set_precision(agpu_potential *pot, int precision)
{
    if (!pot || (precision != AGPU_PREC_FP64 && precision != AGPU_PREC_MIXED))
        return set_error(AGPU_ERR_ARG, "set_precision: unknown mode");
    pot->precision = precision;
    return AGPU_OK;
}

int agpu_potential_bind_to(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl)
{
    if (!pot || !p || !nl) return set_error(AGPU_ERR_ARG, "bind_to: NULL handle");
    pot->p = p;
    pot->nl = nl;
    pot->bound = 0;
    int rc = AGPU_ERR_STATE;
    switch (pot->kind) {
    case AGPU_KIND_TERSOFF:
    case AGPU_KIND_KUMAGAI:
    case AGPU_KIND_BRENNER:
    case AGPU_KIND_TERSOFF_SCR:
    case AGPU_KIND_KUMAGAI_SCR:
    case AGPU_KIND_BRENNER_SCR:
    case AGPU_KIND_JUSLIN:
    case AGPU_KIND_JUSLIN_SCR: rc = tersoff_bind(pot); break;
    case AGPU_KIND_EAM_ALLOY:
    case AGPU_KIND_EAM: rc = eam_bind(pot); break;
    case AGPU_KIND_EWALD:
    case AGPU_KIND_PME: rc = ewald_bind(pot); break;
    case AGPU_KIND_REBO2: rc = rebo2_bind(pot); break;
    default: return set_error(AGPU_ERR_STATE, "bind_to: unknown potential kind");
    }
    if (rc == AGPU_OK) pot->bound = 1;
    return rc;
}

static int stage_and_call(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                          agpu_eval_args a)
{
    if (!pot->bound || pot->p != p || pot->nl != nl)
        return set_error(AGPU_ERR_STATE, "energy_and_forces: potential not bound to these objects");
    const int nat = p->nat;
    int rc = AGPU_OK;
    if (!(a.flags & AGPU_HOST)) {
        switch (pot->kind) {
        case AGPU_KIND_TERSOFF:
        case AGPU_KIND_KUMAGAI:
        case AGPU_KIND_BRENNER:
        case AGPU_KIND_TERSOFF_SCR:
        case AGPU_KIND_KUMAGAI_SCR:
        case AGPU_KIND_BRENNER_SCR:
        case AGPU_KIND_JUSLIN:
        case AGPU_KIND_JUSLIN_SCR: rc = tersoff_compute(pot, &a); break;
        case AGPU_KIND_EAM_ALLOY:
        case AGPU_KIND_EAM: rc = eam_compute(pot, &a); break;
        case AGPU_KIND_EWALD:
        case AGPU_KIND_PME: rc = ewald_compute(pot, &a); break;
        case AGPU_KIND_REBO2: rc = rebo2_compute(pot, &a); break;
        default: rc = set_error(AGPU_ERR_STATE, "unknown potential kind");
        }
        return rc;
    }
    // host arrays: stage through one device block; per-bond arrays hold one
    // entry per list entry of the updated list
    int64_t nbond = 0;
    if (a.epot_per_bond || a.f_per_bond || a.wpot_per_bond) {
        AGPU_CHECK(nl_update(nl, p, 0));
        nbond = nl->total;
    }
    DevBuf st;
    size_t n = 1 + 3 * (size_t)nat + 9 + (a.q ? nat : 0) + (a.epot_per_at ? nat : 0) +
               (a.wpot_per_at ? 9 * (size_t)nat : 0) + (a.epot_per_bond ? (size_t)nbond : 0) +
               (a.f_per_bond ? 3 * (size_t)nbond : 0) + (a.wpot_per_bond ? 9 * (size_t)nbond : 0);
    size_t nmask = a.mask ? (size_t)nat : 0;
    AGPU_CHECK(st.reserve(sizeof(double) * n + sizeof(int) * (nmask + 1)));
    double *b = st.as<double>();
    agpu_eval_args d = a;
    d.flags &= ~AGPU_HOST;
    double *cur = b;
    d.epot = cur;
    cur += 1;
    d.f = cur;
    cur += 3 * (size_t)nat;
    d.wpot = cur;
    cur += 9;
    double *qd = nullptr, *epa = nullptr, *wpa = nullptr;
    if (a.q) { qd = cur; cur += nat; }
    if (a.epot_per_at) { epa = cur; cur += nat; }
    if (a.wpot_per_at) { wpa = cur; cur += 9 * (size_t)nat; }
    double *epb = nullptr, *fpb = nullptr, *wpb = nullptr;
    if (a.epot_per_bond) { epb = cur; cur += nbond; }
    if (a.f_per_bond) { fpb = cur; cur += 3 * (size_t)nbond; }
    if (a.wpot_per_bond) { wpb = cur; cur += 9 * (size_t)nbond; }
    d.epot_per_bond = epb;
    d.f_per_bond = fpb;
    d.wpot_per_bond = wpb;
    int *md = a.mask ? reinterpret_cast<int *>(cur) : nullptr;
    d.q = qd;
    d.epot_per_at = epa;
    d.wpot_per_at = wpa;
    d.mask = md;
    hipStream_t s = stream();
    AGPU_HIP(hipMemsetAsync(d.epot, 0, sizeof(double), s));
    if (a.f) AGPU_HIP(hipMemcpyAsync(d.f, a.f, sizeof(double) * 3 * nat, hipMemcpyHostToDevice, s));
    else AGPU_HIP(hipMemsetAsync(d.f, 0, sizeof(double) * 3 * nat, s));
    if (a.wpot) AGPU_HIP(hipMemcpyAsync(d.wpot, a.wpot, sizeof(double) * 9, hipMemcpyHostToDevice, s));
    else AGPU_HIP(hipMemsetAsync(d.wpot, 0, sizeof(double) * 9, s));
    if (a.q) AGPU_HIP(hipMemcpyAsync(qd, a.q, sizeof(double) * nat, hipMemcpyHostToDevice, s));
    if (epa) AGPU_HIP(hipMemcpyAsync(epa, a.epot_per_at, sizeof(double) * nat, hipMemcpyHostToDevice, s));
    if (wpa) AGPU_HIP(hipMemcpyAsync(wpa, a.wpot_per_at, sizeof(double) * 9 * nat, hipMemcpyHostToDevice, s));
    if (md) AGPU_HIP(hipMemcpyAsync(md, a.mask, sizeof(int) * nat, hipMemcpyHostToDevice, s));
    if (epb && nbond)
        AGPU_HIP(hipMemcpyAsync(epb, a.epot_per_bond, sizeof(double) * nbond, hipMemcpyHostToDevice, s));
    if (fpb && nbond)
        AGPU_HIP(hipMemcpyAsync(fpb, a.f_per_bond, sizeof(double) * 3 * nbond, hipMemcpyHostToDevice, s));
    if (wpb && nbond)
        AGPU_HIP(hipMemcpyAsync(wpb, a.wpot_per_bond, sizeof(double) * 9 * nbond, hipMemcpyHostToDevice, s));
    switch (pot->kind) {
    case AGPU_KIND_TERSOFF:
    case AGPU_KIND_KUMAGAI:
    case AGPU_KIND_BRENNER:
    case AGPU_KIND_TERSOFF_SCR:
    case AGPU_KIND_KUMAGAI_SCR:
    case AGPU_KIND_BRENNER_SCR:
    case AGPU_KIND_JUSLIN:
    case AGPU_KIND_JUSLIN_SCR: rc = tersoff_compute(pot, &d); break;
    case AGPU_KIND_EAM_ALLOY:
    case AGPU_KIND_EAM: rc = eam_compute(pot, &d); break;
    case AGPU_KIND_EWALD:
    case AGPU_KIND_PME: rc = ewald_compute(pot, &d); break;
    case AGPU_KIND_REBO2: rc = rebo2_compute(pot, &d); break;
    default: rc = set_error(AGPU_ERR_STATE, "unknown potential kind");
    }
    if (rc != AGPU_OK) return rc;
    if (a.epot) AGPU_HIP(hipMemcpyAsync(a.epot, d.epot, sizeof(double), hipMemcpyDeviceToHost, s));
    if (a.f) AGPU_HIP(hipMemcpyAsync(a.f, d.f, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, s));
    if (a.wpot) AGPU_HIP(hipMemcpyAsync(a.wpot, d.wpot, sizeof(double) * 9, hipMemcpyDeviceToHost, s));
    if (epa) AGPU_HIP(hipMemcpyAsync(a.epot_per_at, epa, sizeof(double) * nat, hipMemcpyDeviceToHost, s));
    if (wpa) AGPU_HIP(hipMemcpyAsync(a.wpot_per_at, wpa, sizeof(double) * 9 * nat, hipMemcpyDeviceToHost, s));
    if (nl->total != nbond && (epb || fpb || wpb))
        return set_error(AGPU_ERR_STATE, "energy_and_forces: the list changed size during the call");
    if (epb && nbond)
        AGPU_HIP(hipMemcpyAsync(a.epot_per_bond, epb, sizeof(double) * nbond, hipMemcpyDeviceToHost, s));
    if (fpb && nbond)
        AGPU_HIP(hipMemcpyAsync(a.f_per_bond, fpb, sizeof(double) * 3 * nbond, hipMemcpyDeviceToHost, s));
    if (wpb && nbond)
        AGPU_HIP(hipMemcpyAsync(a.wpot_per_bond, wpb, sizeof(double) * 9 * nbond, hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

int agpu_potential_energy_and_forces(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                                     const double *q, double *epot, double *f, double *wpot,
                                     const int *mask, double *epot_per_at, double *epot_per_bond,
                                     double *f_per_bond, double *wpot_per_at, double *wpot_per_bond,
                                     int flags)
{
    if (!pot || !p || !nl) return set_error(AGPU_ERR_ARG, "energy_and_forces: NULL handle");
    // per-bond properties: the classes whose reference kernels fill them
    // (features per_bond: the bond-order families and REBO2)
    const bool per_bond = pot->kind != AGPU_KIND_EAM_ALLOY && pot->kind != AGPU_KIND_EAM &&
                          pot->kind != AGPU_KIND_EWALD && pot->kind != AGPU_KIND_PME;
    if ((epot_per_bond || f_per_bond || wpot_per_bond) && !per_bond)
        return set_error(AGPU_ERR_UNSUPPORTED,
                         "energy_and_forces: per-bond properties are not supported by this class");
    // TabulatedEAM has no mask feature (tabulated_eam.f90:24)
    if (mask && pot->kind == AGPU_KIND_EAM)
        return set_error(AGPU_ERR_UNSUPPORTED, "energy_and_forces: TabulatedEAM takes no mask");
    if ((pot->kind == AGPU_KIND_EWALD || pot->kind == AGPU_KIND_PME) && !q)
        return set_error(AGPU_ERR_ARG, "energy_and_forces: Coulomb class needs charges");
    if (!epot || !f || !wpot) return set_error(AGPU_ERR_ARG, "energy_and_forces: epot, f and wpot are required");
    agpu_eval_args a;
    a.q = q;
    a.epot = epot;
    a.f = f;
    a.wpot = wpot;
    a.mask = mask;
    a.epot_per_at = epot_per_at;
    a.wpot_per_at = wpot_per_at;
    a.epot_per_bond = epot_per_bond;
    a.f_per_bond = f_per_bond;
    a.wpot_per_bond = wpot_per_bond;
    a.flags = flags;
    return stage_and_call(pot, p, nl, a);
}

int agpu_coulomb_potential(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                           const double *q, double *phi, int flags)
{
    if (!pot || !p || !nl || !q || !phi) return set_error(AGPU_ERR_ARG, "potential: bad arguments");
    if (pot->kind != AGPU_KIND_EWALD && pot->kind != AGPU_KIND_PME)
        return set_error(AGPU_ERR_UNSUPPORTED, "potential: not a Coulomb class");
    if (!pot->bound || pot->p != p || pot->nl != nl)
        return set_error(AGPU_ERR_STATE, "potential: not bound to these objects");
    if (!(flags & AGPU_HOST)) return ewald_potential(pot, q, phi, 0);
    const int nat = p->nat;
    DevBuf st;
    AGPU_CHECK(st.reserve(sizeof(double) * 2 * (size_t)(nat + 1)));
    double *qd = st.as<double>(), *pd = qd + nat + 1;
    hipStream_t s = stream();
    AGPU_HIP(hipMemcpyAsync(qd, q, sizeof(double) * nat, hipMemcpyHostToDevice, s));
    AGPU_HIP(hipMemcpyAsync(pd, phi, sizeof(double) * nat, hipMemcpyHostToDevice, s));
    AGPU_CHECK(ewald_potential(pot, qd, pd, 0));
    AGPU_HIP(hipMemcpyAsync(phi, pd, sizeof(double) * nat, hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

// LAMMPS pair-style step (PairAtomistica::FAtomistica, pair_atomistica.cpp:
// 499-561): local atoms plus ghosts in an open box, the mask marking the
// local atoms so that only terms centred on them carry energy, forces on all
// nall atoms (LAMMPS returns the ghost forces to their owners), energy added
// to *eng and the virial subtracted in Voigt order as the reference does.
// One LAMMPS step; the list is built here (list == false) or was imported.
static int lammps_step(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                       const double *x, const int *Z, const int *tag, const int64_t *seed, const int *numneigh,
                       const int *neigh, double *f, double *eng, double *virial, double *eatom,
                       double *vatom = nullptr)
{
    if (!pot || !p || !nl || nlocal < 0 || nall < nlocal || (nall > 0 && (!x || !Z || !f)) || !eng || !virial)
        return set_error(AGPU_ERR_ARG, "lammps_compute: bad arguments");
    const bool list = seed != nullptr;
    if (list && nall > 0 && (!numneigh || !neigh))
        return set_error(AGPU_ERR_ARG, "lammps_compute_list: bad neighbour arrays");
    const bool rebo2 = pot->kind == AGPU_KIND_REBO2;
    if (rebo2 && !tag)
        return set_error(AGPU_ERR_UNSUPPORTED,
                         "lammps_compute: REBO2 takes no mask and needs the atom tags (agpu_lammps_compute_tagged)");
    if (pot->kind == AGPU_KIND_EWALD || pot->kind == AGPU_KIND_PME)
        return set_error(AGPU_ERR_UNSUPPORTED, "lammps_compute: Ewald needs a periodic cell");
    if (nall == 0) return AGPU_OK;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = x[k];
    for (int i = 1; i < nall; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], x[3 * (size_t)i + k]);
            hi[k] = std::max(hi[k], x[3 * (size_t)i + k]);
        }
    const double margin = 1.0 + pot->cutoff;
    double cell[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 3; ++k) cell[4 * k] = hi[k] - lo[k] + 2 * margin;
    std::vector<double> xs(3 * (size_t)nall);
    for (int i = 0; i < nall; ++i)
        for (int k = 0; k < 3; ++k) xs[3 * (size_t)i + k] = x[3 * (size_t)i + k] - lo[k] + margin;
    const int pbc[3] = {0, 0, 0};
    if (p->nat != nall) AGPU_CHECK(agpu_particles_allocate(p, nall));
    AGPU_CHECK(agpu_particles_set_cell(p, cell, pbc));
    if (p->h_Z.size() != (size_t)nall || !std::equal(Z, Z + nall, p->h_Z.begin()))
        AGPU_CHECK(agpu_particles_set_atomic_numbers(p, Z, AGPU_HOST));
    AGPU_CHECK(agpu_particles_set_positions(p, xs.data(), AGPU_HOST));
    nl->external = false;
    if (!pot->bound || pot->p != p || pot->nl != nl) AGPU_CHECK(agpu_potential_bind_to(pot, p, nl));
    if (list) AGPU_CHECK(nl_import(nl, p, seed, numneigh, neigh));
    double epot = 0.0, wpot[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // per-atom virial: the 3x3 per atom, turned into LAMMPS' Voigt vatom with
    // the sign of SUM_VIRIAL in LAMMPS mode (macros.inc:202)
    std::vector<double> wpa(vatom ? 9 * (size_t)nall : 0, 0.0);
    double *wpa_p = vatom ? wpa.data() : nullptr;
    if (rebo2) {
        // half rule by tags over the local centres instead of a mask
        AGPU_CHECK(p->d_tag.reserve(sizeof(int) * (size_t)nall));
        AGPU_HIP(hipMemcpyAsync(p->d_tag.ptr, tag, sizeof(int) * (size_t)nall, hipMemcpyHostToDevice, stream()));
        p->natloc = nlocal;
        const int rc = agpu_potential_energy_and_forces(pot, p, nl, nullptr, &epot, f, wpot, nullptr, eatom,
                                                        nullptr, nullptr, wpa_p, nullptr, AGPU_HOST);
        p->natloc = -1;
        AGPU_CHECK(rc);
    } else {
        std::vector<int> mask(nall);
        for (int i = 0; i < nall; ++i) mask[i] = i < nlocal;
        AGPU_CHECK(agpu_potential_energy_and_forces(pot, p, nl, nullptr, &epot, f, wpot, mask.data(), eatom,
                                                    nullptr, nullptr, wpa_p, nullptr, AGPU_HOST));
    }
    if (vatom)
        for (int i = 0; i < nall; ++i) {
            const double *w = wpa_p + 9 * (size_t)i;    // W(a, b) at a + 3 b
            double *v = vatom + 6 * (size_t)i;
            v[0] -= w[0];
            v[1] -= w[4];
            v[2] -= w[8];
            v[3] -= w[1];
            v[4] -= w[2];
            v[5] -= w[5];
        }
    *eng += epot;
    virial[0] -= wpot[0];
    virial[1] -= wpot[4];
    virial[2] -= wpot[8];
    virial[3] -= 0.5 * (wpot[1] + wpot[3]);
    virial[4] -= 0.5 * (wpot[2] + wpot[6]);
    virial[5] -= 0.5 * (wpot[5] + wpot[7]);
    return AGPU_OK;
}

int agpu_lammps_compute(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                        const double *x, const int *Z, double *f, double *eng, double *virial, double *eatom)
{
    return lammps_step(pot, p, nl, nlocal, nall, x, Z, nullptr, nullptr, nullptr, nullptr, f, eng, virial, eatom);
}

int agpu_lammps_compute_tagged(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                               const double *x, const int *Z, const int *tag, const int64_t *seed,
                               const int *numneigh, const int *neigh, double *f, double *eng, double *virial,
                               double *eatom, double *vatom)
{
    return lammps_step(pot, p, nl, nlocal, nall, x, Z, tag, seed, numneigh, neigh, f, eng, virial, eatom, vatom);
}

int agpu_lammps_compute_list(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int nlocal, int nall,
                             const double *x, const int *Z, const int64_t *seed, const int *numneigh,
                             const int *neigh, double *f, double *eng, double *virial, double *eatom)
{
    if (!seed) return set_error(AGPU_ERR_ARG, "lammps_compute_list: seed is NULL");
    return lammps_step(pot, p, nl, nlocal, nall, x, Z, nullptr, seed, numneigh, neigh, f, eng, virial, eatom);
}

}  // extern "C"
