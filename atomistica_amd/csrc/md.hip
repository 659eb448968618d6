// NVE molecular dynamics around the potential, with the step captured in a
// HIP graph.
//
// Velocity Verlet (the integrator of the reference's MDCORE driver,
// src/standalone/main.f90:429-531, restated for one potential): per step
//   k_md_drift   v += dt/2 f/m;  r += dt v;  and the largest squared
//                displacement since the last list build (atomic max of the
//                float bits, which order like the values for x >= 0)
//   f = 0, then the potential's kernels (list held)
//   k_md_kick    v += dt/2 f/m
// The list is built with a Verlet shell (skin) on top of the interaction
// range and kept while every atom stays within skin/2 of where it was at the
// build.  The reference's Python path never rebuilds with a shell (its
// accum_max_dr is never advanced, python_particles.f90:125-128), so the skin
// logic is this driver's own.
//
// Graph mode: runs of K steps (K = 32, 16, 8 or 4, the longest predicted to
// stay inside the shell) are captured once per list build and replayed as
// one graph launch.  Before a replay r, v, f and the displacement maximum are
// copied aside; after it the maximum is read.  If it passed skin/2 inside the
// replay, the state is restored and the K steps are redone one at a time,
// each checking the displacement after the drift and rebuilding the list
// before the force call when needed -- so both modes give the same
// trajectory, bitwise.
#include <cmath>
#include <cstring>
#include <vector>

#include "agpu_internal.hpp"
#include "profile.hpp"

struct agpu_md {
    agpu_potential *pot = nullptr;
    agpu_particles *p = nullptr;
    agpu_neighbors *nl = nullptr;
    int nat = 0;
    double dt = 1.0, skin = 0.0;
    agpu::DevBuf v, minv, rref, f, ck, ev, dmax;
    agpu::HostStats hs;
    bool ready = false;
    // graphs of 32, 16, 8 and 4 steps, captured when first needed after a
    // list build
    static constexpr int NG = 4;
    hipGraph_t graph[NG] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t exec[NG] = {nullptr, nullptr, nullptr, nullptr};
    int64_t builds = 0, replays = 0, rollbacks = 0, steps = 0;
    ~agpu_md()
    {
        for (int g = 0; g < NG; ++g) {
            if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
            if (graph[g]) (void)hipGraphDestroy(graph[g]);
        }
        if (nl) {
            nl->held = false;
            nl->skin = 0.0;
        }
        if (p) (void)agpu_particles_changed_positions(p);
    }
};

namespace agpu {

__global__ void __launch_bounds__(256)
k_md_drift(int n, double hdt, double dt, const double *__restrict__ minv, const double *__restrict__ f,
           double *__restrict__ v, double *__restrict__ r, const double *__restrict__ rref,
           unsigned *__restrict__ dmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (i < n) {
        const double mi = minv[i];
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const size_t k = 3 * (size_t)i + x;
            const double vk = v[k] + hdt * f[k] * mi;
            const double rk = r[k] + dt * vk;
            v[k] = vk;
            r[k] = rk;
            const double d = rk - rref[k];
            s += d * d;
        }
        // round up so that the float never understates the displacement
        d2 = __double2float_ru(s);
    }
    for (int o = 32; o >= 1; o >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, o, 64));
    if ((threadIdx.x & 63) == 0 && d2 > 0.0f) atomicMax(dmax, __float_as_uint(d2));
}

// also the largest squared speed since the last list build (rounded up)
__global__ void __launch_bounds__(256)
k_md_kick(int n, double hdt, const double *__restrict__ minv, const double *__restrict__ f, double *__restrict__ v,
          unsigned *__restrict__ vmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float v2 = 0.0f;
    if (i < n) {
        const double mi = minv[i];
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double vk = v[3 * (size_t)i + x] + hdt * f[3 * (size_t)i + x] * mi;
            v[3 * (size_t)i + x] = vk;
            s += vk * vk;
        }
        v2 = __double2float_ru(s);
    }
    for (int o = 32; o >= 1; o >>= 1) v2 = fmaxf(v2, __shfl_xor(v2, o, 64));
    if ((threadIdx.x & 63) == 0 && v2 > 0.0f) atomicMax(vmax, __float_as_uint(v2));
}

static int md_force(agpu_md *md)
{
    agpu_eval_args a{};
    double *ev = md->ev.as<double>();
    a.epot = ev;
    a.f = md->f.as<double>();
    a.wpot = ev + 1;
    AGPU_CHECK(zero_dev(md->f.ptr, sizeof(double) * 3 * md->nat));
    AGPU_CHECK(zero_dev(ev + 1, sizeof(double) * 9));
    switch (md->pot->kind) {
    case AGPU_KIND_TERSOFF:
    case AGPU_KIND_KUMAGAI:
    case AGPU_KIND_BRENNER:
    case AGPU_KIND_TERSOFF_SCR:
    case AGPU_KIND_KUMAGAI_SCR:
    case AGPU_KIND_BRENNER_SCR:
    case AGPU_KIND_JUSLIN:
    case AGPU_KIND_JUSLIN_SCR: return tersoff_compute(md->pot, &a);
    case AGPU_KIND_EAM_ALLOY:
    case AGPU_KIND_EAM: return eam_compute(md->pot, &a);
    default: return set_error(AGPU_ERR_UNSUPPORTED, "MD driver: Tersoff-family and EAM potentials only");
    }
}

static int md_rebuild(agpu_md *md)
{
    hipStream_t s = stream();
    md->nl->held = false;
    // EAM reads the staged rows; the bond-order kernels need the CSR rows
    const bool eam = md->pot->kind == AGPU_KIND_EAM_ALLOY || md->pot->kind == AGPU_KIND_EAM;
    AGPU_CHECK(nl_update(md->nl, md->p, 1, !eam));
    md->nl->held = true;
    AGPU_HIP(hipMemcpyAsync(md->rref.ptr, md->p->d_r, sizeof(double) * 3 * md->nat, hipMemcpyDeviceToDevice, s));
    AGPU_CHECK(zero_dev(md->dmax.ptr, sizeof(unsigned) * 3));   // displacement, its copy, speed
    md->builds++;
    for (int g = 0; g < agpu_md::NG; ++g) {
        if (md->exec[g]) {
            AGPU_HIP(hipGraphExecDestroy(md->exec[g]));
            md->exec[g] = nullptr;
        }
        if (md->graph[g]) {
            AGPU_HIP(hipGraphDestroy(md->graph[g]));
            md->graph[g] = nullptr;
        }
    }
    return AGPU_OK;
}

static int md_drift(agpu_md *md)
{
    const int nb = (md->nat + 255) / 256;
    AGPU_LAUNCH("md_drift", k_md_drift, dim3(nb), dim3(256), 0, stream(), md->nat, 0.5 * md->dt, md->dt,
                md->minv.as<double>(), md->f.as<double>(), md->v.as<double>(), md->p->d_r, md->rref.as<double>(),
                md->dmax.as<unsigned>());
    return AGPU_OK;
}

static int md_kick(agpu_md *md)
{
    const int nb = (md->nat + 255) / 256;
    AGPU_LAUNCH("md_kick", k_md_kick, dim3(nb), dim3(256), 0, stream(), md->nat, 0.5 * md->dt,
                md->minv.as<double>(), md->f.as<double>(), md->v.as<double>(), md->dmax.as<unsigned>() + 2);
    return AGPU_OK;
}

// largest squared displacement since the list build; *v2 (optional) the
// largest squared speed seen since then
static double md_read_dmax(agpu_md *md, double *v2 = nullptr)
{
    hipStream_t s = stream();
    if (hipMemcpyAsync(md->hs.h, md->dmax.ptr, sizeof(unsigned) * 3, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return INFINITY;
    float d2, w2;
    std::memcpy(&d2, md->hs.h, sizeof d2);
    std::memcpy(&w2, md->hs.h + 2, sizeof w2);
    if (v2) *v2 = (double)w2;
    return (double)d2;
}

// one step, checking the displacement after the drift
static int md_step_checked(agpu_md *md)
{
    AGPU_CHECK(md_drift(md));
    const double lim = 0.25 * md->skin * md->skin;
    if (md_read_dmax(md) >= lim) AGPU_CHECK(md_rebuild(md));
    AGPU_CHECK(md_force(md));
    AGPU_CHECK(md_kick(md));
    md->steps++;
    return AGPU_OK;
}

static int md_capture(agpu_md *md, int gi, int K)
{
    hipStream_t s = stream();
    AGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    set_capturing(true);
    int rc = AGPU_OK;
    for (int k = 0; k < K && rc == AGPU_OK; ++k) {
        rc = md_drift(md);
        if (rc == AGPU_OK) rc = md_force(md);
        if (rc == AGPU_OK) rc = md_kick(md);
    }
    set_capturing(false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != AGPU_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    md->graph[gi] = g;
    AGPU_HIP(hipGraphInstantiate(&md->exec[gi], g, nullptr, nullptr, 0));
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_md_create(agpu_md **out, agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl,
                              const double *masses, const double *v0, double dt, double skin)
{
    if (!out || !pot || !p || !nl || !masses || !(dt > 0.0) || !(skin >= 0.0))
        return set_error(AGPU_ERR_ARG, "md_create: bad arguments");
    if (!(pot->kind == AGPU_KIND_TERSOFF || pot->kind == AGPU_KIND_KUMAGAI || pot->kind == AGPU_KIND_BRENNER ||
          pot->kind == AGPU_KIND_TERSOFF_SCR || pot->kind == AGPU_KIND_KUMAGAI_SCR ||
          pot->kind == AGPU_KIND_BRENNER_SCR || pot->kind == AGPU_KIND_JUSLIN || pot->kind == AGPU_KIND_JUSLIN_SCR ||
          pot->kind == AGPU_KIND_EAM_ALLOY || pot->kind == AGPU_KIND_EAM))
        return set_error(AGPU_ERR_UNSUPPORTED, "MD driver: Tersoff-family and EAM potentials only");
    const int n = p->nat;
    for (int i = 0; i < n; ++i)
        if (!(masses[i] > 0.0)) return set_error(AGPU_ERR_ARG, "md_create: masses must be positive");
    agpu_md *md = new agpu_md();
    md->pot = pot;
    md->p = p;
    md->nl = nl;
    md->nat = n;
    md->dt = dt;
    md->skin = skin;
    int rc = AGPU_OK;
    std::vector<double> minv(n + 1, 0.0);
    for (int i = 0; i < n; ++i) minv[i] = 1.0 / masses[i];
    hipStream_t s = stream();
    if ((rc = md->v.reserve(sizeof(double) * 3 * (n + 1))) || (rc = md->minv.reserve(sizeof(double) * (n + 1))) ||
        (rc = md->rref.reserve(sizeof(double) * 3 * (n + 1))) || (rc = md->f.reserve(sizeof(double) * 3 * (n + 1))) ||
        (rc = md->ck.reserve(sizeof(double) * 9 * (n + 1))) || (rc = md->ev.reserve(sizeof(double) * 16)) ||
        (rc = md->dmax.reserve(sizeof(unsigned) * 4)) || (rc = md->hs.init())) {
        delete md;
        return rc;
    }
    if (n) {
        hipError_t e = hipMemcpyAsync(md->minv.ptr, minv.data(), sizeof(double) * n, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            if (v0) e = hipMemcpyAsync(md->v.ptr, v0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s);
            else e = hipMemsetAsync(md->v.ptr, 0, sizeof(double) * 3 * n, s);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            delete md;
            return hip_fail(e, "md_create", __FILE__, __LINE__);
        }
    }
    nl->skin = skin;
    *out = md;
    return AGPU_OK;
}

extern "C" int agpu_md_destroy(agpu_md *md)
{
    delete md;
    return AGPU_OK;
}

// nsteps velocity-Verlet steps.  graph_steps > 0 replays captured graphs of
// up to that many steps (32, 16, 8 or 4: the longest that the displacement so
// far and the largest speed predict to stay inside the shell); other steps
// run one by one.  epot (host, optional) receives the potential energy of
// the last step.
extern "C" int agpu_md_run(agpu_md *md, int nsteps, int graph_steps, double *epot)
{
    if (!md || nsteps < 0 || graph_steps < 0) return set_error(AGPU_ERR_ARG, "md_run: bad arguments");
    if (md->p->nat != md->nat) return set_error(AGPU_ERR_STATE, "md_run: the particles changed size");
    hipStream_t s = stream();
    if (!md->pot->bound || md->pot->p != md->p || md->pot->nl != md->nl)
        AGPU_CHECK(agpu_potential_bind_to(md->pot, md->p, md->nl));
    if (!md->ready) {
        // first force evaluation at the initial positions (also sizes every buffer)
        AGPU_CHECK(md_rebuild(md));
        AGPU_CHECK(md_force(md));
        md->ready = true;
    }
    const double lim = 0.25 * md->skin * md->skin;
    static const int KS[agpu_md::NG] = {32, 16, 8, 4};
    int done = 0;
    double v2 = 0.0;
    double d2 = md_read_dmax(md, &v2);
    while (done < nsteps) {
        // steps that surely stay inside the shell: the largest speed so far,
        // with a quarter of margin for acceleration, from the displacement
        // reached so far (the rollback below keeps this a prediction only)
        int K = 0;
        // the screened bond-order kernels size their records on the host
        // every step, so they always run step by step
        const bool graphable = md->pot->kind == AGPU_KIND_EAM_ALLOY || md->pot->kind == AGPU_KIND_EAM ||
                               !tersoff_screened(md->pot);
        if (graphable && graph_steps > 0 && md->skin > 0.0) {
            const double room = 0.5 * md->skin - std::sqrt(d2);
            const double per = 1.25 * std::sqrt(v2) * md->dt;
            const double safe = per > 0.0 ? room / per : 1e9;
            for (int g = 0; g < agpu_md::NG; ++g)
                if (KS[g] <= graph_steps && KS[g] <= nsteps - done && KS[g] <= safe) {
                    K = KS[g];
                    break;
                }
        }
        if (K > 0) {
            int gi = 0;
            while (KS[gi] != K) ++gi;
            if (!md->exec[gi]) AGPU_CHECK(md_capture(md, gi, K));
            const size_t n3 = 3 * (size_t)md->nat;
            double *ck = md->ck.as<double>();
            unsigned *dm = md->dmax.as<unsigned>();
            AGPU_HIP(hipMemcpyAsync(ck, md->p->d_r, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
            AGPU_HIP(hipMemcpyAsync(ck + n3, md->v.ptr, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
            AGPU_HIP(hipMemcpyAsync(ck + 2 * n3, md->f.ptr, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
            AGPU_HIP(hipMemcpyAsync(dm + 1, dm, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
            AGPU_HIP(hipGraphLaunch(md->exec[gi], s));
            md->replays++;
            d2 = md_read_dmax(md, &v2);
            if (d2 >= lim) {
                // the list went stale inside the replay: restore and redo step by step
                AGPU_HIP(hipMemcpyAsync(md->p->d_r, ck, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
                AGPU_HIP(hipMemcpyAsync(md->v.ptr, ck + n3, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
                AGPU_HIP(hipMemcpyAsync(md->f.ptr, ck + 2 * n3, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
                AGPU_HIP(hipMemcpyAsync(dm, dm + 1, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
                md->rollbacks++;
                for (int k = 0; k < K; ++k) AGPU_CHECK(md_step_checked(md));
                d2 = md_read_dmax(md, &v2);
            } else {
                md->steps += K;
            }
            done += K;
        } else {
            AGPU_CHECK(md_step_checked(md));
            d2 = md_read_dmax(md, &v2);
            ++done;
        }
    }
    // potential error flags of the captured steps are read here
    AGPU_HIP(hipStreamSynchronize(s));
    if (epot) {
        AGPU_HIP(hipMemcpyAsync(md->hs.h, md->ev.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
        AGPU_HIP(hipStreamSynchronize(s));
        std::memcpy(epot, md->hs.h, sizeof(double));
    }
    return AGPU_OK;
}

// Copies of the state (host or device per flags; any may be NULL).
extern "C" int agpu_md_get(agpu_md *md, double *r, double *v, double *f, double *wpot, int64_t *counters,
                           int flags)
{
    if (!md) return set_error(AGPU_ERR_ARG, "md_get: md is NULL");
    hipStream_t s = stream();
    const hipMemcpyKind k = (flags & AGPU_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    const size_t n3 = 3 * (size_t)md->nat;
    if (r && n3) AGPU_HIP(hipMemcpyAsync(r, md->p->d_r, sizeof(double) * n3, k, s));
    if (v && n3) AGPU_HIP(hipMemcpyAsync(v, md->v.ptr, sizeof(double) * n3, k, s));
    if (f && n3) AGPU_HIP(hipMemcpyAsync(f, md->f.ptr, sizeof(double) * n3, k, s));
    if (wpot) AGPU_HIP(hipMemcpyAsync(wpot, md->ev.as<double>() + 1, sizeof(double) * 9, k, s));
    AGPU_HIP(hipStreamSynchronize(s));
    if (counters) {
        counters[0] = md->steps;
        counters[1] = md->builds;
        counters[2] = md->replays;
        counters[3] = md->rollbacks;
    }
    return AGPU_OK;
}
