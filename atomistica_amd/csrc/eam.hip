// Tabulated alloy EAM (setfl) and tabulated single-element EAM (funcfl) on
// gfx950.
//
// Semantics: tabulated_alloy_eam.f90 -- reader :147-259, bind_to :297-350,
// kernel :423-627 -- with the natural cubic spline tables of
// simple_spline.f90:127-615; TabulatedEAM (tabulated_eam.f90: reader
// :141-214, kernel :334-511) is the same decomposition with one element and
// the pair term Z(r)^2 / r of the effective-charge table.  Work
// decomposition (this library's own):
//
//   k_eam_density  one wavefront per atom, lanes stride over its neighbour
//                  row: rho_i = sum rho_{el j}(r), wave reduction, rho clamped
//                  at 0, then F(rho_i) and F'(rho_i) (extrapolating spline).
//   k_eam_force    one wavefront per atom again, owner-computes:
//                  f_i = -sum_j [m_i (F'_i rho'_{el j}(r) + p(r))
//                               + m_j (F'_j rho'_{el i}(r) + p(r))] dr/r
//                  with p = (phi' - phi/r)/r of the half-scaled r*phi table;
//                  the reference's scatter to j becomes the second term, so no
//                  atomics are needed and every sum has a fixed order.
// Spline records are 4-wide AoS per interval {y, c1, c2, c3}: one value or
// derivative evaluation reads one 32-byte (FP64) or 16-byte (FP32) record;
// the derivative (c1 + 2 c2 B + 3 c3 B^2) / dx is formed from the same
// coefficients, so a value-and-derivative gather is two 16-byte loads.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "agpu_internal.hpp"
#include "profile.hpp"

namespace agpu {

static const char *kElementNames[] = {
    "xx", "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",
    "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga",
    "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag",
    "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu",
    "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au",
    "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu", "Am",
    "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs", "Mt", "Ds", "Rg",
    "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};
static const int kNumElements = (int)(sizeof(kElementNames) / sizeof(kElementNames[0]));

static int z_of_symbol(const std::string &s)
{
    for (int z = 1; z < kNumElements; ++z)
        if (s == kElementNames[z]) return z;
    return -1;
}

// One natural cubic spline on a uniform grid from x0 = 0 (simple_spline_init).
struct HostSpline {
    int n = 0;
    double dx = 0.0;
    std::vector<double> rec;   // SPL_W doubles per interval
};

constexpr int SPL_W = 4;       // coefficients per interval record

static void spline_build(HostSpline &s, const std::vector<double> &y, double dx, double scale)
{
    const int n = (int)y.size();
    const double sig = 0.5;
    std::vector<double> d2y(n), u(n);
    d2y[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        double p = sig * d2y[i - 1] + 2;
        d2y[i] = (sig - 1) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / dx - (y[i] - y[i - 1]) / dx) / (2 * dx) - sig * u[i - 1]) / p;
    }
    d2y[n - 1] = 0.0;   // natural end: (un - qn*u)/(qn*d2y + 1) with qn = un = 0
    for (int k = n - 2; k >= 0; --k) d2y[k] = d2y[k] * d2y[k + 1] + u[k];
    const double dx2 = dx * dx;
    s.n = n;
    s.dx = dx;
    s.rec.assign(SPL_W * (size_t)(n - 1), 0.0);
    for (int k = 0; k < n - 1; ++k) {
        double c1 = y[k + 1] - y[k] - (2 * d2y[k] + d2y[k + 1]) * dx2 / 6;
        double c2 = d2y[k] * dx2 / 2;
        double c3 = (d2y[k + 1] - d2y[k]) * dx2 / 6;
        double *r = &s.rec[SPL_W * (size_t)k];
        // scale_y_axis multiplies every coefficient (simple_spline.f90:756-773)
        r[0] = scale * y[k];
        r[1] = scale * c1;
        r[2] = scale * c2;
        r[3] = scale * c3;
    }
}

struct SplHdr {
    int n;        // grid points
    int off;      // first record
    double rdx;   // 1/dx
    double cut;   // x0 + dx*(n-1)
    double dx;
};

constexpr int EAM_MAX_EL = 8;

struct EamDev {
    int nel;
    int funcfl;                   // pair table is Z(r): r phi = Z^2 (TabulatedEAM)
    int tF[EAM_MAX_EL], tR[EAM_MAX_EL], tP[EAM_MAX_EL * EAM_MAX_EL];
    double cutoff_sq;
    SplHdr h1F, h1R, h1P;     // the three tables of a single-element potential
};

struct EamState {
    std::string fn, elements;
    bool funcfl = false;          // TabulatedEAM (funcfl file)
    int nel = 0;
    std::vector<std::string> names;
    std::vector<int> Zs;
    double cutoff = 0.0;
    std::vector<HostSpline> F, rho, phi;   // phi: nel*nel (symmetric)
    std::vector<SplHdr> hdr;
    EamDev dev{};
    int z2db[128];
    uint64_t Z_rev = ~0ull;
    DevBuf d_hdr, d_rec64, d_rec32, partial, stats, pk;
    bool all_known = false;       // every atom is of a tabulated element
    // lanes per atom and entries per lane of the single-element kernels
    // (chosen by eam_autotune; the defaults otherwise)
    int tg = 16, tu = 4;
};

static void eam_free(void *p) { delete static_cast<EamState *>(p); }

// list-directed number stream over the whole file (Fortran read(f, *))
struct Tok {
    std::vector<std::string> lines;
    size_t line = 0;
    std::vector<std::string> pending;
    bool next_line(std::string &out)
    {
        if (line >= lines.size()) return false;
        out = lines[line++];
        pending.clear();
        return true;
    }
    bool number(double &v)
    {
        while (pending.empty()) {
            if (line >= lines.size()) return false;
            std::string l = lines[line++];
            for (char &c : l)
                if (c == ',') c = ' ';
            std::istringstream is(l);
            std::string t;
            std::vector<std::string> toks;
            while (is >> t) toks.push_back(t);
            pending.assign(toks.rbegin(), toks.rend());
        }
        std::string t = pending.back();
        pending.pop_back();
        for (char &c : t)
            if (c == 'd' || c == 'D') c = 'e';
        char *end = nullptr;
        v = std::strtod(t.c_str(), &end);
        return end && *end == '\0';
    }
};

static int read_setfl(EamState &st)
{
    std::ifstream in(st.fn);
    if (!in) return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: cannot open '" + st.fn + "'");
    Tok tk;
    std::string l;
    while (std::getline(in, l)) tk.lines.push_back(l);
    if (tk.lines.size() < 6) return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: file too short");
    std::string dummy;
    for (int k = 0; k < 3; ++k) tk.next_line(dummy);
    tk.next_line(l);
    {
        std::istringstream is(l);
        is >> st.nel;
        if (!is || st.nel < 1 || st.nel > EAM_MAX_EL)
            return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: bad element count line");
        st.names.resize(st.nel);
        for (int e = 0; e < st.nel; ++e) is >> st.names[e];
    }
    tk.next_line(l);
    int nF = 0, nr = 0;
    double dF = 0, dr = 0, cutoff = 0;
    {
        for (char &c : l)
            if (c == ',') c = ' ';
        std::istringstream is(l);
        is >> nF >> dF >> nr >> dr >> cutoff;
        if (!is || nF < 3 || nr < 3) return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: bad grid line");
    }
    st.cutoff = cutoff;
    st.Zs.resize(st.nel);
    st.F.resize(st.nel);
    st.rho.resize(st.nel);
    st.phi.resize((size_t)st.nel * st.nel);
    for (int e = 0; e < st.nel; ++e) {
        tk.next_line(l);
        std::istringstream is(l);
        int Z = 0;
        is >> Z;
        if (Z <= 0 || Z >= kNumElements || st.names[e] != kElementNames[Z])
            return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: atomic number " + std::to_string(Z) +
                                              " does not match element name " + st.names[e]);
        st.Zs[e] = Z;
        std::vector<double> y(nF);
        for (int k = 0; k < nF; ++k)
            if (!tk.number(y[k])) return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: truncated F table");
        tk.pending.clear();   // one list-directed read per table: rest of its last line is skipped
        spline_build(st.F[e], y, dF, 1.0);
        std::vector<double> yr(nr + 2, 0.0);   // padded with two zeros (:222)
        for (int k = 0; k < nr; ++k)
            if (!tk.number(yr[k])) return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: truncated rho table");
        spline_build(st.rho[e], yr, dr, 1.0);
        tk.pending.clear();
    }
    for (int a = 0; a < st.nel; ++a)
        for (int b = 0; b <= a; ++b) {
            std::vector<double> yp(nr + 2, 0.0);
            for (int k = 0; k < nr; ++k)
                if (!tk.number(yp[k])) return set_error(AGPU_ERR_IO, "TabulatedAlloyEAM: truncated phi table");
            tk.pending.clear();
            spline_build(st.phi[a * st.nel + b], yp, dr, 0.5);
            st.phi[b * st.nel + a] = st.phi[a * st.nel + b];
        }
    return AGPU_OK;
}

// Units of the reference (src/support/Units.f90:74,76)
constexpr double HARTREE_REF = 27.2113961;
constexpr double BOHR_REF = 0.529177249;

// funcfl reader, tabulated_eam_init (tabulated_eam.f90:141-214): a comment
// line, "Z mass a0 lattice", "nF dF nr dr cutoff", then list-directed reads
// of F(rho) (nF from 0), Z(r) and rho(r) (nr each, from 0).  Z is scaled by
// sqrt(Hartree Bohr / 2) (:178), so the pair energy of a directed pair is
// Z^2 / r.
static int read_funcfl(EamState &st)
{
    std::ifstream in(st.fn);
    if (!in) return set_error(AGPU_ERR_IO, "TabulatedEAM: cannot open '" + st.fn + "'");
    Tok tk;
    std::string l;
    while (std::getline(in, l)) tk.lines.push_back(l);
    if (tk.lines.size() < 4) return set_error(AGPU_ERR_IO, "TabulatedEAM: file too short");
    std::string dummy;
    tk.next_line(dummy);
    tk.next_line(l);
    int Z = 0;
    {
        std::istringstream is(l);
        is >> Z;
        if (!is || Z <= 0 || Z >= kNumElements) return set_error(AGPU_ERR_IO, "TabulatedEAM: bad element line");
    }
    tk.next_line(l);
    int nF = 0, nr = 0;
    double dF = 0, dr = 0, cutoff = 0;
    {
        for (char &c : l)
            if (c == ',') c = ' ';
        std::istringstream is(l);
        is >> nF >> dF >> nr >> dr >> cutoff;
        if (!is || nF < 3 || nr < 3) return set_error(AGPU_ERR_IO, "TabulatedEAM: bad grid line");
    }
    st.funcfl = true;
    st.nel = 1;
    st.names = {kElementNames[Z]};
    st.Zs = {Z};
    st.cutoff = cutoff;
    st.F.resize(1);
    st.rho.resize(1);
    st.phi.resize(1);
    std::vector<double> y(nF), z(nr), rho(nr);
    for (int k = 0; k < nF; ++k)
        if (!tk.number(y[k])) return set_error(AGPU_ERR_IO, "TabulatedEAM: truncated F table");
    for (int k = 0; k < nr; ++k)
        if (!tk.number(z[k])) return set_error(AGPU_ERR_IO, "TabulatedEAM: truncated Z table");
    for (int k = 0; k < nr; ++k)
        if (!tk.number(rho[k])) return set_error(AGPU_ERR_IO, "TabulatedEAM: truncated rho table");
    spline_build(st.F[0], y, dF, 1.0);
    spline_build(st.phi[0], z, dr, std::sqrt(0.5 * HARTREE_REF * BOHR_REF));
    spline_build(st.rho[0], rho, dr, 1.0);
    return AGPU_OK;
}

static int upload_tables(EamState &st)
{
    std::vector<double> rec64;
    st.hdr.clear();
    auto add = [&](const HostSpline &s) {
        SplHdr h;
        h.n = s.n;
        h.off = (int)(rec64.size() / SPL_W);
        h.dx = s.dx;
        h.rdx = 1.0 / s.dx;
        h.cut = s.dx * (s.n - 1);
        rec64.insert(rec64.end(), s.rec.begin(), s.rec.end());
        st.hdr.push_back(h);
        return (int)st.hdr.size() - 1;
    };
    for (int e = 0; e < st.nel; ++e) st.dev.tF[e] = add(st.F[e]);
    for (int e = 0; e < st.nel; ++e) st.dev.tR[e] = add(st.rho[e]);
    for (int a = 0; a < st.nel; ++a)
        for (int b = 0; b < st.nel; ++b) st.dev.tP[a * EAM_MAX_EL + b] = add(st.phi[a * st.nel + b]);
    st.dev.nel = st.nel;
    st.dev.funcfl = st.funcfl ? 1 : 0;
    st.dev.cutoff_sq = st.cutoff * st.cutoff;
    st.dev.h1F = st.hdr[st.dev.tF[0]];
    st.dev.h1R = st.hdr[st.dev.tR[0]];
    st.dev.h1P = st.hdr[st.dev.tP[0]];
    std::vector<float> rec32(rec64.begin(), rec64.end());
    AGPU_CHECK(st.d_hdr.reserve(sizeof(SplHdr) * st.hdr.size()));
    AGPU_CHECK(st.d_rec64.reserve(sizeof(double) * rec64.size()));
    AGPU_CHECK(st.d_rec32.reserve(sizeof(float) * rec32.size()));
    hipStream_t s = stream();
    AGPU_HIP(hipMemcpyAsync(st.d_hdr.ptr, st.hdr.data(), sizeof(SplHdr) * st.hdr.size(), hipMemcpyHostToDevice, s));
    AGPU_HIP(hipMemcpyAsync(st.d_rec64.ptr, rec64.data(), sizeof(double) * rec64.size(), hipMemcpyHostToDevice, s));
    AGPU_HIP(hipMemcpyAsync(st.d_rec32.ptr, rec32.data(), sizeof(float) * rec32.size(), hipMemcpyHostToDevice, s));
    AGPU_HIP(hipStreamSynchronize(s));
    return AGPU_OK;
}

// --------------------------------------------------------------------------
// device spline evaluation
// --------------------------------------------------------------------------
template <typename R> struct Rec8 { R v[8]; };

// simple_spline_f / dfunc: no extrapolation, error outside [0, cut]
// interval index of xf, or 0 when outside [1, n) (NaN included)
__device__ __forceinline__ int spl_index(const SplHdr &h, double x, double &xf)
{
    if (x == h.cut) {
        xf = h.n;
        return h.n - 1;
    }
    xf = x * h.rdx + 1.0;
    double fl = floor(xf);
    if (!(fl >= 1.0 && fl < (double)h.n)) return 0;
    return (int)fl;
}

template <typename R>
__device__ __forceinline__ R spl_f(const SplHdr &h, const R *__restrict__ rec, double x, int &err)
{
    double xf;
    int i = spl_index(h, x, xf);
    if (i < 1) {
        err |= DERR_SPLINE_RANGE;
        return (R)1;
    }
    R B = (R)(xf - i);
    const R *c = rec + SPL_W * (size_t)(h.off + i - 1);
    return c[0] + B * (c[1] + B * (c[2] + B * c[3]));
}

template <typename R>
__device__ __forceinline__ R spl_df(const SplHdr &h, const R *__restrict__ rec, double x, int &err)
{
    double xf;
    int i = spl_index(h, x, xf);
    if (i < 1) {
        err |= DERR_SPLINE_RANGE;
        return (R)1;
    }
    R B = (R)(xf - i);
    const R *c = rec + SPL_W * (size_t)(h.off + i - 1);
    return (c[1] + B * ((R)2 * c[2] + B * ((R)3 * c[3]))) * (R)h.rdx;
}

// f_and_df(extrapolate=.true.)
template <typename R>
__device__ __forceinline__ void spl_fdf_x(const SplHdr &h, const R *__restrict__ rec, double x, R &f, R &df)
{
    double xf = x * h.rdx + 1.0;
    double fl = floor(xf);
    // clamp to the first / last interval; NaN goes to the first
    int i = !(fl >= 1.0) ? 1 : (!(fl < (double)h.n) ? h.n - 1 : (int)fl);
    R B = (R)(xf - i);
    const R *c = rec + SPL_W * (size_t)(h.off + i - 1);
    f = c[0] + B * (c[1] + B * (c[2] + B * c[3]));
    df = (c[1] + B * ((R)2 * c[2] + B * ((R)3 * c[3]))) * (R)h.rdx;
}

// Lanes per atom: rows of ~50 neighbours are split over G lanes that each
// walk a strided share, so several independent gathers are in flight per lane.
// This is the default; the launch autotuner may pick another G (and U below)
// for the single-element force path.
#ifndef AGPU_EAM_G
#define AGPU_EAM_G 16
#endif
constexpr int EAM_G = AGPU_EAM_G;

// Table header for (table kind, element) -- from the kernel arguments when the
// potential has a single element (ONE), else from the device header array.
template <bool ONE>
__device__ __forceinline__ SplHdr hdr_of(const EamDev &P, const SplHdr *__restrict__ hdr, int kind, int a, int b)
{
    if (ONE) return kind == 0 ? P.h1F : (kind == 1 ? P.h1R : P.h1P);
    int t = kind == 0 ? P.tF[a] : (kind == 1 ? P.tR[a] : P.tP[a * EAM_MAX_EL + b]);
    return hdr[t];
}

// Entries per lane and batch: every lane first issues the row loads of its
// EAM_U entries, then their position gathers, then the table gathers, so one
// latency is paid per batch instead of per neighbour.
#ifndef AGPU_EAM_U
#define AGPU_EAM_U 4
#endif
constexpr int EAM_U = AGPU_EAM_U;

// Packed per-slot records of 32 bytes in the list's bin order (slot p holds
// caller atom order[p]), so that one neighbour gather reads one 32-byte
// sector near the home atom's own: q1 = {x, y, z, element index} for the
// density pass, q2 = {x, y, z, F'(rho)} written by the density pass for the
// force pass.  Also clears the call's four error/statistics words.
__global__ void k_eam_pack(int nat, const SortedAtom *__restrict__ atoms, const int *__restrict__ order,
                           const int *__restrict__ el, double4 *__restrict__ q1, int *__restrict__ flags)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < 4) flags[p] = 0;
    if (p >= nat) return;
    const SortedAtom a = atoms[p];
    q1[p] = make_double4(a.x, a.y, a.z, (double)el[order[p]]);
}

// KN: one element and every atom tabulated, so no element is read per pair.
template <typename R, int G, int U, bool ONE, bool KN>
__global__ void __launch_bounds__(256)
k_eam_density(EamDev P, const SplHdr *__restrict__ hdr, const R *__restrict__ rec, int nat,
              const int *__restrict__ order, const double4 *__restrict__ q1, CellArg C,
              NlRows L, const int *__restrict__ mask, double4 *__restrict__ q2, double *__restrict__ epa,
              double *__restrict__ partial, int *__restrict__ flags, const int *__restrict__ nlflags)
{
    constexpr int APB = 256 / G;
    __shared__ double s_e[APB];
    // a list build marked unusable: nothing is read from its rows and
    // nothing written but a zero record (the call is redone or refused)
    if (*nlflags & DERR_NL_UNUSABLE) {
        if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
        return;
    }
    const int lane = threadIdx.x % G, grp = threadIdx.x / G;
    const int ai = xcd_block() * APB + grp;
    double Fi = 0.0;
    int err = 0;
    if (ai < nat) {
        const int i = order[ai];
        const double4 qi = q1[ai];
        const int eli = KN ? 0 : (int)qi.w;
        const bool on = eli >= 0 && !(mask && mask[i] == 0);
        double rho = 0.0;
        if (on) {
            const double ri[3] = {qi.x, qi.y, qi.z};
            int n;
            const int2 *row = nl_row(L, ai, n);
            for (int s0 = lane; s0 < n; s0 += U * G) {
                int2 e[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int s = s0 + u * G;
                    e[u] = s < n ? row[s] : make_int2(-1, 0);
                }
                double4 qj[U];
#pragma unroll
                for (int u = 0; u < U; ++u) qj[u] = q1[e[u].x < 0 ? ai : e[u].x];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int elj = e[u].x < 0 ? -1 : (KN ? 0 : (int)qj[u].w);
                    if (elj < 0) continue;
                    int a, b, c;
                    dc_unpack(e[u].y, a, b, c);
                    const double rj[3] = {qj[u].x, qj[u].y, qj[u].z};
                    double d[3];
                    delta_r(C, ri, rj, a, b, c, d);
                    const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (d2 < P.cutoff_sq)
                        rho += (double)spl_f<R>(hdr_of<ONE>(P, hdr, 1, elj, 0), rec, sqrt(d2), err);
                }
            }
        }
        rho = group_sum<G>(rho);
        double dFi = 0.0;
        if (on) {
            if (rho < 0.0) rho = 0.0;
            R f, df;
            spl_fdf_x<R>(hdr_of<ONE>(P, hdr, 0, eli, 0), rec, rho, f, df);
            Fi = (double)f;
            dFi = (double)df;
        }
        if (lane == 0) {
            q2[ai] = make_double4(qi.x, qi.y, qi.z, dFi);
            if (epa && on) epa[i] += Fi;
        }
    }
    if (err) atomicOr(flags, err);
    if (lane == 0) s_e[grp] = Fi;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < APB; ++k) t += s_e[k];
        partial[blockIdx.x] = t;
    }
}

// GEN = false: no mask and no per-atom virial requested (the common call),
// so neither is loaded nor accumulated.
template <typename R, int G, int U, bool ONE, bool GEN, bool KN>
__global__ void __launch_bounds__(256)
k_eam_force(EamDev P, const SplHdr *__restrict__ hdr, const R *__restrict__ rec, int nat,
            const int *__restrict__ order, const double4 *__restrict__ q1, const double4 *__restrict__ q2,
            CellArg C, NlRows L, const int *__restrict__ mask, double *__restrict__ f, double *__restrict__ epa,
            double *__restrict__ wpa, const double *__restrict__ dens_e, double *__restrict__ partial,
            int *__restrict__ flags, const int *__restrict__ nlflags)
{
    constexpr int APB = 256 / G;
    __shared__ double s_red[APB][7];
    if (*nlflags & DERR_NL_UNUSABLE) {
        if (threadIdx.x < 10) partial[(size_t)blockIdx.x * 10 + threadIdx.x] = 0.0;
        return;
    }
    const int lane = threadIdx.x % G, grp = threadIdx.x / G;
    const int ai = xcd_block() * APB + grp;
    double acc[3] = {0.0, 0.0, 0.0}, e = 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};      // own-loop virial (xx yy zz yz xz xy)
    double wp[6] = {0, 0, 0, 0, 0, 0};     // per-atom virial
    int err = 0;
    int i = -1;
    if (ai < nat) {
        i = order[ai];
        const int eli = KN ? 0 : (int)q1[ai].w;
        if (eli >= 0) {
            const double mi = (GEN && mask && mask[i] == 0) ? 0.0 : 1.0;
            const double4 qi = q2[ai];
            const R dFi = (R)qi.w;
            const double ri[3] = {qi.x, qi.y, qi.z};
            int n;
            const int2 *row = nl_row(L, ai, n);
            const SplHdr hri = hdr_of<ONE>(P, hdr, 1, eli, 0);
            for (int s0 = lane; s0 < n; s0 += U * G) {
                int2 en[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int s = s0 + u * G;
                    en[u] = s < n ? row[s] : make_int2(-1, 0);
                }
                double4 qj[U];
                int elj[U];
                double mj[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int j = en[u].x < 0 ? ai : en[u].x;
                    qj[u] = q2[j];
                    elj[u] = en[u].x < 0 ? -1 : (KN ? 0 : (int)q1[j].w);
                    mj[u] = (GEN && mask && mask[order[j]] == 0) ? 0.0 : 1.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (elj[u] < 0) continue;
                    int a, b, c;
                    dc_unpack(en[u].y, a, b, c);
                    const double rj[3] = {qj[u].x, qj[u].y, qj[u].z};
                    double d[3];
                    delta_r(C, ri, rj, a, b, c, d);
                    const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (!(d2 < P.cutoff_sq)) continue;
                    const double rrd = rsqrt(d2);
                    const double ad = d2 * rrd;
                    R phi, dphi;
                    spl_fdf_x<R>(hdr_of<ONE>(P, hdr, 2, eli, elj[u]), rec, ad, phi, dphi);
                    if (P.funcfl) {
                        // effective charge Z(r): r phi = Z^2 (tabulated_eam.f90:446-450)
                        dphi = (R)2 * phi * dphi;
                        phi = phi * phi;
                    }
                    const R rr = (R)rrd;
                    const R pp = (dphi - phi * rr) * rr;
                    const R drho_j = spl_df<R>(hdr_of<ONE>(P, hdr, 1, elj[u], 0), rec, ad, err);
                    // one element: rho'_i(r) is the same table at the same r
                    const R drho_i = ONE ? drho_j : spl_df<R>(hri, rec, ad, err);
                    const double ci = mi * (double)(dFi * drho_j + pp);
                    const double cj = mj[u] * (double)((R)qj[u].w * drho_i + pp);
                    e += mi * (double)(phi * rr);
                    const double fs = -(ci + cj) * rrd;
                    acc[0] += fs * d[0];
                    acc[1] += fs * d[1];
                    acc[2] += fs * d[2];
                    const double wo = ci * rrd, wq = 0.5 * (ci + cj) * rrd;
                    const double dd[6] = {d[0] * d[0], d[1] * d[1], d[2] * d[2], d[1] * d[2], d[0] * d[2], d[0] * d[1]};
#pragma unroll
                    for (int q = 0; q < 6; ++q) {
                        w[q] += wo * dd[q];
                        if (GEN) wp[q] += wq * dd[q];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) acc[x] = group_sum<G>(acc[x]);
    e = group_sum<G>(e);
#pragma unroll
    for (int q = 0; q < 6; ++q) w[q] = group_sum<G>(w[q]);
    if (GEN && wpa)
        for (int q = 0; q < 6; ++q) wp[q] = group_sum<G>(wp[q]);
    if (i >= 0 && lane == 0) {
        f[3 * (size_t)i] += acc[0];
        f[3 * (size_t)i + 1] += acc[1];
        f[3 * (size_t)i + 2] += acc[2];
        if (epa) epa[i] += e;
        if (GEN && wpa) {
            double *o = wpa + 9 * (size_t)i;
            // column-major 3x3: [a + 3b]
            o[0] += wp[0];
            o[4] += wp[1];
            o[8] += wp[2];
            o[5] += wp[3];
            o[7] += wp[3];
            o[2] += wp[4];
            o[6] += wp[4];
            o[1] += wp[5];
            o[3] += wp[5];
        }
    }
    if (err) atomicOr(flags, err);
    if (lane == 0) {
        s_red[grp][0] = e;
        for (int q = 0; q < 6; ++q) s_red[grp][1 + q] = w[q];
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int q = threadIdx.x;
        double v = 0.0;
        for (int k = 0; k < APB; ++k) v += s_red[k][q];
        // record layout [e, W column-major 3x3]; e includes the embedding
        // energy of the same block's atoms from the density pass
        double *rec10 = partial + (size_t)blockIdx.x * 10;
        if (q == 0) rec10[0] = v + dens_e[blockIdx.x];
        else if (q == 1) rec10[1 + 0] = v;
        else if (q == 2) rec10[1 + 4] = v;
        else if (q == 3) rec10[1 + 8] = v;
        else if (q == 4) { rec10[1 + 5] = v; rec10[1 + 7] = v; }
        else if (q == 5) { rec10[1 + 2] = v; rec10[1 + 6] = v; }
        else { rec10[1 + 1] = v; rec10[1 + 3] = v; }
    }
}

int eam_bind(agpu_potential *pot)
{
    EamState *st = static_cast<EamState *>(pot->impl);
    agpu_particles *p = pot->p;
    for (int z = 0; z < 128; ++z) st->z2db[z] = -1;
    // element filter (filter_from_string): "*" or empty selects everything
    std::vector<int> allowed;
    bool all = st->elements.empty() || st->elements == "*";
    if (!all) {
        std::string tok;
        std::string s = st->elements + ",";
        for (char c : s) {
            if (c == ',' || c == ' ') {
                if (!tok.empty()) {
                    int z = z_of_symbol(tok);
                    if (z < 0) return set_error(AGPU_ERR_ELEMENT, "TabulatedAlloyEAM: unknown element '" + tok + "'");
                    allowed.push_back(z);
                }
                tok.clear();
            } else {
                tok += c;
            }
        }
    }
    for (int e = 0; e < st->nel; ++e) {
        int z = st->Zs[e];
        bool ok = all;
        for (int a : allowed) ok = ok || a == z;
        if (ok) st->z2db[z] = e;
    }
    AGPU_CHECK(agpu_neighbors_request_interaction_range(pot->nl, st->cutoff));
    AGPU_CHECK(map_elements(pot, st->z2db, 128));
    st->all_known = true;
    for (int z : p->h_Z)
        if (z <= 0 || z >= 128 || st->z2db[z] < 0) st->all_known = false;
    st->Z_rev = p->Z_rev;
    return AGPU_OK;
}

template <typename R, int G, int U, bool ONE, bool GEN, bool KN>
static int eam_passes(EamState *st, agpu_potential *pot, const agpu_eval_args *a, const R *rec, int nblk)
{
    agpu_particles *p = pot->p;
    agpu_neighbors *nl = pot->nl;
    hipStream_t s = stream();
    CellArg C;
    std::memcpy(C.A, p->cell.A, sizeof C.A);
    const int *order = nl->idx_sorted.as<int>();
    double *part = st->partial.as<double>();
    int *flags = st->stats.as<int>();
    const bool mixed = sizeof(R) == 4;
    const int nat = p->nat;
    double4 *q1 = st->pk.as<double4>(), *q2 = q1 + nat;
    const NlRows L = nl_rows(nl);
    AGPU_LAUNCH(mixed ? "eam_density_mixed" : "eam_density_fp64", (k_eam_density<R, G, U, ONE, KN>), dim3(nblk),
                dim3(256), 0, s, st->dev, st->d_hdr.as<SplHdr>(), rec, nat, order, q1, C,
                L, a->mask, q2, a->epot_per_at, part, flags, nl_flags(nl));
    AGPU_LAUNCH(mixed ? "eam_force_mixed" : "eam_force_fp64", (k_eam_force<R, G, U, ONE, GEN, KN>), dim3(nblk),
                dim3(256), 0, s, st->dev, st->d_hdr.as<SplHdr>(), rec, nat, order, q1, q2, C,
                L, a->mask, a->f, a->epot_per_at, a->wpot_per_at, part, part + nblk, flags, nl_flags(nl));
    return AGPU_OK;
}

// the (lanes per atom, entries per lane) variants the autotuner may pick;
// EAM_GMAX (fewest atoms per block, so the most blocks) sizes the per-block
// partial records
constexpr int EAM_TUNE[][2] = {{8, 4}, {16, 2}, {16, 4}, {16, 8}, {32, 2}, {32, 4}};
constexpr int EAM_GMAX = 32;
constexpr bool eam_tune_fits()
{
    for (const auto &v : EAM_TUNE)
        if (v[0] > EAM_GMAX) return false;
    return EAM_G <= EAM_GMAX;
}
static_assert(eam_tune_fits(), "a lane group is wider than EAM_GMAX");

template <typename R, bool ONE, bool KN>
static int eam_launch(EamState *st, agpu_potential *pot, const agpu_eval_args *a, const R *rec, int *nblk_out)
{
    const int nat = pot->p->nat;
    const bool gen = a->mask || a->wpot_per_at;
    int G = EAM_G, U = EAM_U;
    if (ONE && KN && !gen) {
        G = st->tg;
        U = st->tu;
    }
    const int nblk = (nat + 256 / G - 1) / (256 / G);
    if (G > EAM_GMAX) return set_error(AGPU_ERR_STATE, "EAM: lane group wider than the partial records allow");
    *nblk_out = nblk;
    const int *el = pot->d_el.as<int>();
    const agpu_neighbors *nl = pot->nl;
    AGPU_LAUNCH("eam_pack", k_eam_pack, dim3((nat + 255) / 256), dim3(256), 0, stream(), nat,
                nl_rows(nl).atoms, nl->idx_sorted.as<int>(), el, st->pk.as<double4>(), st->stats.as<int>());
    if constexpr (ONE && KN) {
        if (!gen) {
            switch (G * 100 + U) {
            case 804: return eam_passes<R, 8, 4, ONE, false, KN>(st, pot, a, rec, nblk);
            case 1602: return eam_passes<R, 16, 2, ONE, false, KN>(st, pot, a, rec, nblk);
            case 1608: return eam_passes<R, 16, 8, ONE, false, KN>(st, pot, a, rec, nblk);
            case 3202: return eam_passes<R, 32, 2, ONE, false, KN>(st, pot, a, rec, nblk);
            case 3204: return eam_passes<R, 32, 4, ONE, false, KN>(st, pot, a, rec, nblk);
            default: break;
            }
        }
    }
    if (gen) return eam_passes<R, EAM_G, EAM_U, ONE, true, KN>(st, pot, a, rec, nblk);
    return eam_passes<R, EAM_G, EAM_U, ONE, false, KN>(st, pot, a, rec, nblk);
}

template <typename R>
static int eam_dispatch(EamState *st, agpu_potential *pot, const agpu_eval_args *a, const R *rec, int *nblk)
{
    if (st->nel != 1) return eam_launch<R, false, false>(st, pot, a, rec, nblk);
    if (st->all_known) return eam_launch<R, true, true>(st, pot, a, rec, nblk);
    return eam_launch<R, true, false>(st, pot, a, rec, nblk);
}

int eam_variant(agpu_potential *pot, int k, int *g, int *u)
{
    constexpr int n = (int)(sizeof EAM_TUNE / sizeof EAM_TUNE[0]);
    EamState *st = static_cast<EamState *>(pot->impl);
    if (k >= 0 && k < n) {
        st->tg = EAM_TUNE[k][0];
        st->tu = EAM_TUNE[k][1];
    }
    if (g) *g = st->tg;
    if (u) *u = st->tu;
    // only the single-element, all-tabulated force path has variants
    return st->nel == 1 && st->all_known ? n : 0;
}

int eam_compute(agpu_potential *pot, const agpu_eval_args *a)
{
    EamState *st = static_cast<EamState *>(pot->impl);
    agpu_particles *p = pot->p;
    if (st->Z_rev != p->Z_rev) AGPU_CHECK(eam_bind(pot));
    // the list's flags are checked with this call's own (one host round trip)
    AGPU_CHECK(nl_update(pot->nl, p, 0, false, true));
    const int nat = p->nat;
    hipStream_t s = stream();
    if (nat == 0) return zero_dev(a->epot, sizeof(double));
    // blocks of the widest lane group (EAM_GMAX lanes per atom): the most
    // blocks any variant launches
    const int nblk_max = (nat + 256 / EAM_GMAX - 1) / (256 / EAM_GMAX);
    // layout: nblk energy scalars of the density pass, then nblk 10-records
    AGPU_CHECK(st->partial.reserve(sizeof(double) * 11 * (size_t)nblk_max));
    AGPU_CHECK(st->stats.reserve(sizeof(int) * 4));
    AGPU_CHECK(pot->stats.init());
    AGPU_CHECK(st->pk.reserve(sizeof(double4) * 2 * (size_t)(nat + 1)));
    int nblk = 0;
    if (pot->precision == AGPU_PREC_MIXED)
        AGPU_CHECK(eam_dispatch<float>(st, pot, a, st->d_rec32.as<float>(), &nblk));
    else
        AGPU_CHECK(eam_dispatch<double>(st, pot, a, st->d_rec64.as<double>(), &nblk));
    // epot := sum of the block records (embedding + pair energy), wpot +=
    AGPU_CHECK(reduce_ev(st->partial.as<double>() + nblk, nblk, 10, a->epot, a->wpot, nullptr, true));
    if (capturing()) return AGPU_OK;     // the MD driver reads the flags after the replay
    AGPU_HIP(hipMemcpyAsync(pot->stats.h, st->stats.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    AGPU_HIP(hipStreamSynchronize(s));
    bool redo = false;
    AGPU_CHECK(nl_settle(pot->nl, &redo));
    if (redo) {
        // the build is redone synchronously (radix sort or a wider
        // staging stride), then the passes again
        AGPU_CHECK(nl_update(pot->nl, p, 1, false));
        return eam_compute(pot, a);
    }
    if (pot->stats.h[0] & DERR_SPLINE_RANGE)
        return set_error(AGPU_ERR_SPLINE, std::string(st->funcfl ? "TabulatedEAM" : "TabulatedAlloyEAM") +
                                              ": spline argument outside the defined interval");
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_eam_create(agpu_potential **pot, const char *fn, const char *elements)
{
    if (!pot || !fn) return set_error(AGPU_ERR_ARG, "eam_create: NULL argument");
    EamState *st = new EamState();
    st->fn = fn;
    st->elements = elements ? elements : "*";
    int rc = read_funcfl(*st);
    if (rc == AGPU_OK) rc = upload_tables(*st);
    if (rc != AGPU_OK) {
        delete st;
        return rc;
    }
    agpu_potential *p = new agpu_potential();
    p->kind = AGPU_KIND_EAM;
    p->cutoff = st->cutoff;
    p->impl = st;
    p->free_impl = eam_free;
    *pot = p;
    return AGPU_OK;
}

extern "C" int agpu_eam_alloy_create(agpu_potential **pot, const char *fn, const char *elements)
{
    if (!pot || !fn) return set_error(AGPU_ERR_ARG, "eam_alloy_create: NULL argument");
    EamState *st = new EamState();
    st->fn = fn;
    st->elements = elements ? elements : "*";
    int rc = read_setfl(*st);
    if (rc == AGPU_OK) rc = upload_tables(*st);
    if (rc != AGPU_OK) {
        delete st;
        return rc;
    }
    agpu_potential *p = new agpu_potential();
    p->kind = AGPU_KIND_EAM_ALLOY;
    p->cutoff = st->cutoff;
    p->impl = st;
    p->free_impl = eam_free;
    *pot = p;
    return AGPU_OK;
}
