// Launch-shape autotuner: times the interchangeable kernel variants of a
// bound potential on the caller's configuration with HIP events and keeps
// the fastest for later calls.
//
//   - neighbour-list pair search: thread per home atom (sparse bins) vs one
//     wavefront per block of bins (dense bins).  Both emit identical rows,
//     so the choice never changes a result.
//   - single-element tabulated EAM: lanes per atom G (64/G atoms per
//     wavefront) x row entries per lane and batch U (loads in flight).  The
//     per-atom sums change their association order with G, so results move
//     at the level of double rounding only.
//   - unscreened bond-order classes: the row path (a lane per row entry,
//     reverse entries found in the neighbour's row) vs the bond-compacting
//     path (bond slots, reverse-slot pass).  The per-atom sums keep their
//     order; the virial sums differ at the level of double rounding.
//   - REBO2: lanes per half pair of the torsion pass (4, 8, 16).  The
//     torsion sums of a pair change their association order with the width.
//
// The reference has no counterpart: its kernels are CPU loops with no
// launch shape to choose.
#include "agpu_internal.hpp"

#include <cstdio>
#include <cstring>
#include <string>

namespace agpu {

namespace {

struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EvPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    int init()
    {
        AGPU_HIP(hipEventCreate(&a));
        AGPU_HIP(hipEventCreate(&b));
        return AGPU_OK;
    }
};

// mean milliseconds of reps calls of fn() on the library stream (one
// untimed call first); fn returns an AGPU status
template <typename Fn>
int time_calls(EvPair &ev, int reps, Fn fn, double *ms)
{
    hipStream_t s = stream();
    AGPU_CHECK(fn());
    AGPU_HIP(hipStreamSynchronize(s));
    AGPU_HIP(hipEventRecord(ev.a, s));
    for (int r = 0; r < reps; ++r) AGPU_CHECK(fn());
    AGPU_HIP(hipEventRecord(ev.b, s));
    AGPU_HIP(hipEventSynchronize(ev.b));
    float t = 0.0f;
    AGPU_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
    *ms = (double)t / reps;
    return AGPU_OK;
}

void append(std::string &rep, const char *fmt, double a, double b = 0.0, double c = 0.0)
{
    char buf[160];
    std::snprintf(buf, sizeof buf, fmt, a, b, c);
    rep += buf;
}

}  // namespace

static int autotune(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int reps, std::string &rep)
{
    EvPair ev;
    AGPU_CHECK(ev.init());
    AGPU_CHECK(nl_update(nl, p, 0));
    if (!nl->external && p->nat > 0) {
        double best = 0.0;
        int pick = 0;
        for (int mode = 1; mode <= 2; ++mode) {
            nl->pair_mode = mode;
            double ms = 0.0;
            AGPU_CHECK(time_calls(ev, reps, [&] { return nl_update(nl, p, 1); }, &ms));
            append(rep, mode == 1 ? "nl_pairs per_atom %.4f ms\n" : "nl_pairs bin_block %.4f ms\n", ms);
            if (pick == 0 || ms < best) {
                best = ms;
                pick = mode;
            }
        }
        nl->pair_mode = pick;
        AGPU_CHECK(nl_update(nl, p, 1));
        rep += pick == 1 ? "nl_pairs -> per_atom\n" : "nl_pairs -> bin_block\n";
    }
    const bool eam = pot->kind == AGPU_KIND_EAM_ALLOY || pot->kind == AGPU_KIND_EAM;
    if (eam && p->nat > 0) {
        const int nv = eam_variant(pot, -1, nullptr, nullptr);
        if (nv > 0) {
            DevBuf out;
            const size_t nat = (size_t)p->nat;
            AGPU_CHECK(out.reserve(sizeof(double) * (3 * nat + 10)));
            agpu_eval_args a{};
            a.epot = out.as<double>();
            a.wpot = a.epot + 1;
            a.f = a.epot + 10;
            double best = 0.0;
            int pick = -1, g = 0, u = 0;
            for (int k = 0; k < nv; ++k) {
                eam_variant(pot, k, &g, &u);
                double ms = 0.0;
                AGPU_CHECK(time_calls(ev, reps, [&] {
                    AGPU_CHECK(zero_dev(out.ptr, sizeof(double) * (3 * nat + 10)));
                    return eam_compute(pot, &a);
                }, &ms));
                append(rep, "eam lanes_per_atom %.0f entries_per_lane %.0f: %.4f ms\n", g, u, ms);
                if (pick < 0 || ms < best) {
                    best = ms;
                    pick = k;
                }
            }
            eam_variant(pot, pick, &g, &u);
            append(rep, "eam -> lanes_per_atom %.0f entries_per_lane %.0f\n", g, u);
        }
    }
    const int nb = bop_variant(pot, -1);
    if (nb > 0 && p->nat > 0) {
        DevBuf out;
        const size_t nat = (size_t)p->nat;
        AGPU_CHECK(out.reserve(sizeof(double) * (3 * nat + 10)));
        agpu_eval_args a{};
        a.epot = out.as<double>();
        a.wpot = a.epot + 1;
        a.f = a.epot + 10;
        double best = 0.0;
        int pick = -1;
        for (int k = 0; k < nb; ++k) {
            bop_variant(pot, k);
            double ms = 0.0;
            AGPU_CHECK(time_calls(ev, reps, [&] {
                AGPU_CHECK(zero_dev(out.ptr, sizeof(double) * (3 * nat + 10)));
                return tersoff_compute(pot, &a);
            }, &ms));
            append(rep, k == 0 ? "bond order rows %.4f ms\n" : "bond order compacted %.4f ms\n", ms);
            if (pick < 0 || ms < best) {
                best = ms;
                pick = k;
            }
        }
        bop_variant(pot, pick);
        rep += pick == 0 ? "bond order -> rows\n" : "bond order -> compacted\n";
    }
    if (pot->kind == AGPU_KIND_REBO2 && p->nat > 0) {
        DevBuf out;
        const size_t nat = (size_t)p->nat;
        AGPU_CHECK(out.reserve(sizeof(double) * (3 * nat + 10)));
        agpu_eval_args a{};
        a.epot = out.as<double>();
        a.wpot = a.epot + 1;
        a.f = a.epot + 10;
        const int nv = rebo2_variant(pot, -1, nullptr);
        double best = 0.0;
        int pick = -1, g = 0;
        for (int k = 0; k < nv; ++k) {
            rebo2_variant(pot, k, &g);
            double ms = 0.0;
            AGPU_CHECK(time_calls(ev, reps, [&] {
                AGPU_CHECK(zero_dev(out.ptr, sizeof(double) * (3 * nat + 10)));
                return rebo2_compute(pot, &a);
            }, &ms));
            append(rep, "rebo2 torsion lanes_per_pair %.0f: %.4f ms\n", g, ms);
            if (pick < 0 || ms < best) {
                best = ms;
                pick = k;
            }
        }
        rebo2_variant(pot, pick, &g);
        append(rep, "rebo2 -> torsion lanes_per_pair %.0f\n", g);
    }
    return AGPU_OK;
}

}  // namespace agpu

using namespace agpu;

extern "C" int agpu_potential_autotune(agpu_potential *pot, agpu_particles *p, agpu_neighbors *nl, int reps,
                                       char *report, int len)
{
    if (!pot || !p || !nl) return set_error(AGPU_ERR_ARG, "autotune: null handle");
    if (!pot->bound || pot->p != p || pot->nl != nl)
        return set_error(AGPU_ERR_STATE, "autotune: potential not bound to these objects");
    if (reps < 1) reps = 1;
    if (capturing()) return set_error(AGPU_ERR_STATE, "autotune: not during graph capture");
    std::string rep;
    const int rc = autotune(pot, p, nl, reps, rep);
    if (report && len > 0) {
        const size_t n = rep.size() < (size_t)(len - 1) ? rep.size() : (size_t)(len - 1);
        std::memcpy(report, rep.data(), n);
        report[n] = '\0';
    }
    return rc;
}

extern "C" int agpu_potential_set_launch(agpu_potential *pot, agpu_neighbors *nl, int pair_mode, int variant,
                                         int *nvariants)
{
    if (!pot || !nl) return set_error(AGPU_ERR_ARG, "set_launch: null handle");
    if (!pot->bound || pot->nl != nl)
        return set_error(AGPU_ERR_STATE, "set_launch: potential not bound to this neighbour list");
    if (pair_mode < 0 || pair_mode > 2) return set_error(AGPU_ERR_ARG, "set_launch: pair_mode must be 0, 1 or 2");
    const bool eam = pot->kind == AGPU_KIND_EAM_ALLOY || pot->kind == AGPU_KIND_EAM;
    const bool rebo2 = pot->kind == AGPU_KIND_REBO2;
    const int nv = eam ? eam_variant(pot, -1, nullptr, nullptr)
                       : (rebo2 ? rebo2_variant(pot, -1, nullptr) : bop_variant(pot, -1));
    if (variant < -1 || variant >= (nv > 0 ? nv : 1))
        return set_error(AGPU_ERR_ARG, "set_launch: no such kernel variant for this potential");
    nl->pair_mode = pair_mode;
    if (variant >= 0 && nv > 0) {
        if (eam) eam_variant(pot, variant, nullptr, nullptr);
        else if (rebo2) rebo2_variant(pot, variant, nullptr);
        else bop_variant(pot, variant);
    }
    if (nvariants) *nvariants = nv;
    return AGPU_OK;
}
