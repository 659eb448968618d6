#!/usr/bin/env python
"""Throughput of the potential-evaluation path on one MI355X.

A step is one pass of the hot path over one batch of synthetic input: the
neighbour list is rebuilt (positions are flagged as changed, as the ASE
calculator does on every call) and energy, forces and virial are evaluated,
with positions, forces and virial resident in HBM.  The default workload is
the tabulated-EAM Cu system (40^3 fcc cells, 256 000 atoms, rc = 5.5 A,
Finnis-Sinclair-form functions on 10^4-point splines) in FP64.

Prints one JSON line.  Per-kernel durations come from HIP events recorded
around every launch on the library's launch stream during a profiled pass of
the same steps right after the timed region.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPUs this process may use, taken before an OpenMP runtime starts: with
# OMP_PROC_BIND the runtime pins the initial thread to its first place
ALLOWED_CPUS = sorted(os.sched_getaffinity(0))
# OpenMP placement of the CPU-baseline threads (read when the runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "atom-evals/s (E+F+virial), FP64 plain vs FP32/FP64-acc batched at force err≤1e-5"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def build_workload(name):
    from atomistica_amd import workloads as W
    if name == "eam_cu_256k":
        pos, cell, pbc, Z = W.fcc_cu(40, sigma=0.08, seed=2)
        d = tempfile.mkdtemp()
        fn = os.path.join(d, "fs_cu.eam.alloy")
        with open(fn, "w") as f:
            f.write(W.fs_cu_setfl_text())
        desc = ("EAM Cu fcc 40x40x40 cells = 256000 atoms, a=3.615, sigma=0.08 (rng 2), "
                "periodic; Finnis-Sinclair-form rho/phi/F on 10000-point splines, rc=5.5")
        return dict(pos=pos, cell=cell, pbc=pbc, Z=Z, kind="eam", fn=fn, avgn=100, desc=desc)
    if name in ("tersoff_si_1m", "tersoff_si_4096"):
        n = 50 if name == "tersoff_si_1m" else 8
        pos, cell, pbc, Z = W.diamond_si(n, sigma=0.05, seed=1)
        desc = ("Tersoff Si (PRB 39, 5566) diamond %dx%dx%d cells = %d atoms, a=5.431, "
                "sigma=0.05 (rng 1), periodic" % (n, n, n, len(pos)))
        return dict(pos=pos, cell=cell, pbc=pbc, Z=Z, kind="tersoff", avgn=16, desc=desc)
    if name == "tersoff_scr_si_64k":
        pos, cell, pbc, Z = W.diamond_si(20, sigma=0.1, seed=1)
        desc = ("TersoffScr Si (screened PRB 39, 5566 set) diamond 20x20x20 cells = %d atoms, a=5.431, "
                "sigma=0.1 (rng 1), periodic; screening range 6.36 A" % len(pos))
        return dict(pos=pos, cell=cell, pbc=pbc, Z=Z, kind="tersoff_scr", avgn=80, desc=desc)
    if name == "ewald_nacl_216k":
        pos, cell, pbc, Z, q = W.rocksalt(30, sigma=0.1, seed=5)
        beta = W.ewald_beta(10.0, 1e-8)
        desc = ("Ewald rocksalt +-1e 30^3 cells = 216000 ions, a=5.64, sigma=0.1 (rng 5), "
                "rc=10, erfc(beta rc)=1e-8, exp(-kmax^2/4beta^2)=1e-8")
        return dict(pos=pos, cell=cell, pbc=pbc, Z=Z, q=q, kind="ewald", beta=beta,
                    kmax=W.ewald_kmax(beta, 1e-8), avgn=220, desc=desc)
    if name == "pme_nacl_216k":
        pos, cell, pbc, Z, q = W.rocksalt(30, sigma=0.1, seed=5)
        desc = ("PME rocksalt +-1e 30^3 cells = 216000 ions, a=5.64, sigma=0.1 (rng 5), rc=10 "
                "(beta^2 = ln(10) 12 / rc^2), B-spline order 8 on a 160^3 FFT grid")
        return dict(pos=pos, cell=cell, pbc=pbc, Z=Z, q=q, kind="pme", grid=(160, 160, 160), order=8,
                    avgn=220, desc=desc)
    if name == "rebo2_ch_100k":
        pos, cell, pbc, Z = W.amorphous_ch(100000, x_h=0.2, density=2.0, seed=10)
        desc = ("REBO2 amorphous hydrocarbon 100000 atoms (80% C, 20% H), cubic periodic box at "
                "2.0 g/cm^3, uniform random with rejection at C-C 1.2, C-H 1.0, H-H 0.8 A (rng 10)")
        return dict(pos=pos, cell=cell, pbc=pbc, Z=Z, kind="rebo2", dihedral=True, avgn=24,
                    desc=desc + "; P bicubic, F and T tricubic (torsion term on)")
    raise SystemExit("unknown workload " + name)


def make_potential(N, wl, precision):
    if wl["kind"] == "eam":
        pot = N.TabulatedAlloyEAM(fn=wl["fn"])
    elif wl["kind"] == "tersoff":
        pot = N.Tersoff()
    elif wl["kind"] == "tersoff_scr":
        pot = N.TersoffScr()
    elif wl["kind"] == "rebo2":
        pot = N.Rebo2(dihedral=wl.get("dihedral", False))
    elif wl["kind"] == "pme":
        pot = N.PME(cutoff=10.0, order=wl["order"], grid=wl["grid"])
    else:
        pot = N.Ewald(cutoff=10.0, beta=wl["beta"], kmax=wl["kmax"])
    pot.set_precision(precision)
    return pot


# algorithmic HBM bytes per atom of one launch: positions r (24), species (4),
# row offset (4), neighbour row (8 per entry: index + packed image), forces (24)
def algorithmic_bytes(kernel, nn):
    if kernel.startswith("eam_force"):
        return 64 + 8 * nn            # + F'(rho) of the density pass (8)
    if kernel.startswith("eam_density"):
        return 40 + 8 * nn            # writes F'(rho) (8)
    if kernel.startswith("tersoff"):
        return 56 + 8 * nn
    if kernel in ("nl_fill", "nl_pairs"):
        return 40 + 8 * nn            # list build: r, species, bin key, offset + the row written
    if kernel == "nl_count":
        return 36
    if kernel.startswith("ewald_real"):
        return 64 + 8 * nn
    if kernel in ("scr_count", "scr_fill"):
        return 36 + 16 * nn           # the row (8 per entry) and per-entry flags / cutoffs
    if kernel.startswith("tersoff_scr"):
        return 56 + 8 * nn
    if kernel == "pme_spread":
        return None                   # grid-point gather; traffic is per grid point
    if kernel.startswith("rebo2_"):
        return 64 + 8 * nn            # the whole pass's minimum per atom: r, species, row, list, forces
    return None


class Runner:
    def __init__(self, wl, precision):
        import torch
        from atomistica_amd import native as N
        self.torch = torch
        self.N = N
        N.use_torch_stream()
        dev = torch.device("cuda:0")
        self.wl = wl
        nat = len(wl["pos"])
        self.nat = nat
        self.p = N.from_atoms(wl["pos"], wl["cell"], wl["pbc"], wl["Z"])
        self.r = torch.tensor(wl["pos"], dtype=torch.float64, device=dev).contiguous()
        self.p.set_positions_device(self.r)
        self.nl = N.Neighbors(wl["avgn"])
        self.pot = make_potential(N, wl, precision)
        self.pot.bind_to(self.p, self.nl)
        self.f = torch.zeros((nat, 3), dtype=torch.float64, device=dev)
        self.epot = torch.zeros(1, dtype=torch.float64, device=dev)
        self.wpot = torch.zeros(9, dtype=torch.float64, device=dev)
        self.q = (torch.tensor(wl["q"], dtype=torch.float64, device=dev) if "q" in wl else None)

    def step(self, rebuild=True):
        if rebuild:
            self.p.changed_positions()
        self.f.zero_()
        self.wpot.zero_()
        self.pot.energy_and_forces_device(self.p, self.nl, self.f, self.epot, self.wpot, charges=self.q)

    def timed(self, steps, rebuild=True):
        torch = self.torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step(rebuild)
        torch.cuda.synchronize()
        return time.perf_counter() - t0


def host_info():
    """lscpu model, sockets, physical cores and the CPUs this process may use."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(ALLOWED_CPUS)}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        out = ""
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    info["model"] = kv.get("Model name")
    try:
        sockets = int(kv.get("Socket(s)", "1"))
        cps = int(kv.get("Core(s) per socket", "0"))
        info["sockets"] = sockets
        info["physical_cores"] = sockets * cps if cps else None
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "1"))
    except ValueError:
        info["physical_cores"] = None
    return info


def baseline_threads(info):
    """All physical cores this process may run on; OMP_NUM_THREADS, when the
    environment sets it (the CPU share of the GPU's host), caps it."""
    n = min(x for x in (info.get("physical_cores"), info["affinity_cpus"]) if x)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


_ORACLE_BUILD = None


def native_oracle():
    """The oracle compiled on this host with -O3 -march=native -fopenmp
    (-ffp-contract=off kept), loaded before its first use."""
    global _ORACLE_BUILD
    if _ORACLE_BUILD is None:
        from oracle import oracle as O
        try:
            O.use_library(O.build_native(tempfile.mkdtemp()))
            _ORACLE_BUILD = "gcc -O3 -march=native -fopenmp -ffp-contract=off"
        except Exception as e:           # keep the prebuilt -O2 oracle
            _ORACLE_BUILD = "prebuilt oracle/_build (-O2); native build failed: %s" % type(e).__name__
    return _ORACLE_BUILD


def _time_calls(one, budget_s):
    t0 = time.perf_counter()
    one()                      # warm-up (page faults, thread pool)
    warm = time.perf_counter() - t0
    n = max(1, min(10, int(budget_s / max(warm, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(n):
        one()
    return n, time.perf_counter() - t0


def cpu_baseline(wl, budget_s=8.0):
    """The CPU oracle (a C restatement of the reference loops) on the host
    cores, on all physical cores available and on one thread: neighbour-list
    rebuild + energy/forces/virial per call, the same system as the GPU."""
    import numpy as np
    build = native_oracle()
    from oracle import oracle as O
    from atomistica_amd import parameters as P
    info = host_info()
    pos, cell, pbc = wl["pos"], wl["cell"], wl["pbc"]
    nat = len(pos)
    if wl["kind"] == "eam":
        tables = O.read_setfl_tables(wl["fn"])
        el = np.zeros(nat, np.int32)

        def one(t):
            nl = O.nl_build(pos, cell, pbc, tables["cutoff"], nthreads=t)
            O.eam(tables, pos, el, cell, nl, nthreads=t)
    elif wl["kind"] == "tersoff":
        el = np.ones(nat, np.int32)

        def one(t):
            nl = O.nl_build(pos, cell, pbc, 3.0, nthreads=t)
            O.tersoff(P.Tersoff_PRB_39_5566_Si_C, pos, el, cell, nl, nthreads=t)
    elif wl["kind"] == "tersoff_scr":
        el = np.ones(nat, np.int32)
        prm = P.Tersoff_PRB_39_5566_Si_C__Scr
        rc = 6.0 * np.sqrt(9.0 / 8.0)

        def one(t):
            nl = O.nl_build(pos, cell, pbc, rc, avgn=80, nthreads=t)
            O.tersoff(prm, pos, el, cell, nl, nthreads=t)
    elif wl["kind"] == "rebo2":
        R = O.Rebo2(dihedral=wl.get("dihedral", False))

        def one(t):
            nl = O.nl_build(pos, cell, pbc, 2.0, nthreads=t)
            R(pos, wl["Z"], cell, nl, nthreads=t)
    else:
        return None
    threads = baseline_threads(info)
    n, dt = _time_calls(lambda: one(threads), budget_s)
    n1, dt1 = _time_calls(lambda: one(1), budget_s)
    what = "full steps (NL rebuild + E+F+virial) of the same %d-atom system" % nat
    return {"value": nat * n / dt, "unit": "atom-evals/s", "cores": threads, "kind": "port",
            "sample": "%d %s, OpenMP %d threads (OMP_PROC_BIND=close, OMP_PLACES=cores); C restatement of "
                      "the reference loops, not the reference binary" % (n, what, threads),
            "one_thread": {"value": nat * n1 / dt1, "cores": 1, "sample": "%d %s, 1 thread" % (n1, what)},
            "build": build, "host": info}


def valu_rates(workload, kern):
    """FP64/FP32 VALU FLOP rates of the profiled kernels: FLOPs per launch
    from the committed rocprofv3 counter summary (profiles/valu.json) over
    this run's HIP-event average duration."""
    fn = os.path.join(ROOT, "profiles", "valu.json")
    if not os.path.exists(fn):
        return None
    try:
        tab = json.load(open(fn)).get(workload, {})
    except ValueError:
        return None
    out = {}
    for k, rec in tab.items():
        if k not in kern:
            continue
        s = kern[k]["avg_us"] * 1e-6
        f64 = rec["flops_fp64"] / s / 1e12
        f32 = rec["flops_fp32"] / s / 1e12
        out[k] = {"tflops_fp64": f64, "frac_fp64_peak": f64 / FP64_PEAK_TFLOPS,
                  "tflops_fp32": f32, "frac_fp32_peak": f32 / FP32_PEAK_TFLOPS,
                  "valu_winst_per_s": rec["valu_insts"] / s}
    return out or None


def _kernel_times(N, fn, steps):
    """Per-kernel HIP-event averages over `steps` calls of fn()."""
    import torch
    N.profile_reset()
    N.profile_enable(True)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    N.profile_enable(False)
    kern = {}
    for k in N.profile_list():
        ms, cnt = N.profile_query(k)
        if cnt:
            kern[k] = {"avg_us": 1e3 * ms / cnt, "launches": cnt}
    return kern


def batch_block(args, cpu=True):
    """Batched Tersoff for fitting: 4096 periodic Si structures of 8..512
    atoms (log-uniform), random triclinic cells at 2.33 g/cm^3 +-10 %, random
    positions at minimum distance 2.0 A.  A step is one fitting step: forward
    E+F in FP32-arith/FP64-acc, then the backward pass for d(loss)/dtheta of
    a force-matching loss (the neighbour list is built once; positions do not
    move during fitting)."""
    import numpy as np
    import torch
    from atomistica_amd import native as N
    from atomistica_amd import parameters as P
    from atomistica_amd import workloads as W
    from atomistica_amd.batch import TersoffBatch, theta_from_params
    t0 = time.perf_counter()
    S = W.random_si_structures(4096, seed=4)
    t_gen = time.perf_counter() - t0
    b = TersoffBatch(S)
    th = theta_from_params(P.Tersoff_Si)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e64, f64 = b.forward(th, "fp64")              # builds the list
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    ntot = b.ntot
    gf = torch.randn_like(f64)
    ge = torch.ones_like(e64)

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    fwd64 = timed(lambda: b.forward(th, "fp64"), args.steps)
    fwd32 = timed(lambda: b.forward(th, "mixed"), args.steps)
    bwd = timed(lambda: b.backward(th, ge, gf), args.steps)
    e32, f32 = b.forward(th, "mixed")
    e64, f64 = b.forward(th, "fp64")
    df = (f32 - f64).cpu().numpy()
    de = (e32 - e64).cpu().numpy() / np.diff(b.offsets)
    step = fwd32 + bwd
    kern = _kernel_times(N, lambda: (b.forward(th, "fp64"), b.forward(th, "mixed"), b.backward(th, ge, gf)),
                         args.steps)
    npairs = b.pairs()
    nn = npairs / ntot
    roof = None
    dom = "tersoff_batch_fp64"
    if dom in kern:
        bpa = 56 + 8 * nn
        achieved = bpa * ntot / (kern[dom]["avg_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "algorithmic_bytes_per_atom": bpa, "atoms_per_launch": ntot}
    valu = valu_rates("tersoff_batch_4096", kern)
    if roof and valu and dom in valu:
        roof["valu_frac_fp64"] = valu[dom]["frac_fp64_peak"]
    out = {
        "metric": METRIC,
        "value": ntot / step,
        "unit": "atom-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-arith/f64-acc", "data": "synthetic",
        "config": {"workload": "batched Tersoff Si fitting: 4096 periodic structures, %d atoms, %d pairs "
                               "(log-uniform 8..512 atoms, triclinic, 2.33 g/cm^3 +-10%%, dmin 2.0 A, rng 4)"
                               % (ntot, npairs),
                   "step": "forward E+F (FP32-arith/FP64-acc) + backward d(loss)/dtheta, 11 parameters",
                   "inputs": "positions and neighbour list resident in HBM"},
        "forward_fp64": {"ms": 1e3 * fwd64, "value": ntot / fwd64},
        "forward_mixed": {"ms": 1e3 * fwd32, "value": ntot / fwd32, "speedup_vs_fp64": fwd64 / fwd32,
                          "force_rms_err": float(np.sqrt(np.mean(df ** 2))),
                          "force_max_err": float(np.abs(df).max()),
                          "energy_err_per_atom_max": float(np.abs(de).max())},
        "backward": {"ms": 1e3 * bwd},
        "first_forward_ms_incl_list": 1e3 * t_first,
        "structure_generation_s": t_gen,
        "kernels": kern,
        "roofline": roof,
        "valu": valu,
    }
    if cpu:
        out["cpu_baseline"] = cpu_baseline_batch(S, th)
    del b
    torch.cuda.empty_cache()
    return out


def cpu_baseline_batch(S, th, budget_s=8.0):
    """The oracle's FP64 forward (list build + E + F) structure by structure,
    on all available physical cores (one structure per worker) and on one
    thread, over as many structures as the budget allows."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    build = native_oracle()
    from oracle import oracle as O
    from atomistica_amd.batch import params_from_theta
    prm = params_from_theta(th)
    info = host_info()
    threads = baseline_threads(info)

    def one(k):
        # a pool thread inherits the pinned affinity of the initial thread
        os.sched_setaffinity(0, ALLOWED_CPUS)
        pos, cell = S[k]
        nl = O.nl_build(pos, cell, [True] * 3, th[10])
        O.tersoff(prm, pos, np.zeros(len(pos), np.int32), cell, nl)
        return len(pos)

    def run(workers):
        t0 = time.perf_counter()
        na = ns = 0
        with ThreadPoolExecutor(max_workers=workers) as ex:
            while ns < len(S) and time.perf_counter() - t0 < budget_s:
                chunk = list(range(ns, min(len(S), ns + 8 * workers)))
                na += sum(ex.map(one, chunk))
                ns += len(chunk)
        return na, ns, time.perf_counter() - t0

    na, ns, dt = run(threads)
    na1, ns1, dt1 = run(1)
    return {"value": na / dt, "unit": "atom-evals/s", "cores": threads, "kind": "port",
            "sample": "FP64 forward (list build + E + F) of the first %d of the 4096 structures (%d atoms), "
                      "%d worker threads, one structure each (the ctypes calls release the GIL); "
                      "C restatement of the reference loops" % (ns, na, threads),
            "one_thread": {"value": na1 / dt1, "cores": 1,
                           "sample": "the first %d structures (%d atoms), one thread" % (ns1, na1)},
            "build": build, "host": info}


def bench_batch(args):
    print(json.dumps(batch_block(args, cpu=not args.no_cpu_baseline)))


def bench_ewald_batch(args):
    """Batched Ewald: 256 rocksalt cells of n^3 conventional cells (n uniform
    in 2..8, 64..4096 ions), sigma = 0.1, rc = 10 A, erfc(beta rc) = 1e-8,
    exp(-kmax^2/4beta^2) = 1e-8.  A step is E + F + virial of every cell."""
    import numpy as np
    import torch
    from atomistica_amd import native as N
    from atomistica_amd import workloads as W
    from atomistica_amd.batch import Batch, ewald_batch
    S, q = W.rocksalt_cells(256, seed=5)
    b = Batch(S)
    rc = 10.0
    beta = W.ewald_beta(rc, 1e-8)
    kmax = W.ewald_kmax(beta, 1e-8)
    N.use_torch_stream()

    def timed(prec, steps):
        for _ in range(args.warmup):
            ewald_batch(b, q, rc, beta, kmax, prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            ewald_batch(b, q, rc, beta, kmax, prec)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    t64 = timed("fp64", args.steps)
    t32 = timed("mixed", args.steps)
    e64, f64, w64, _ = ewald_batch(b, q, rc, beta, kmax, "fp64")
    e32, f32, w32, _ = ewald_batch(b, q, rc, beta, kmax, "mixed")
    df = (f32 - f64).cpu().numpy()
    de = np.abs((e32 - e64).cpu().numpy()) / np.diff(b.offsets)
    kern = _kernel_times(N, lambda: ewald_batch(b, q, rc, beta, kmax, "fp64"), args.steps)
    out = {
        "metric": METRIC, "value": b.ntot / t64, "unit": "atom-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t64,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched Ewald: 256 rocksalt cells, n^3 conventional cells each, n uniform in 2..8 "
                               "(%d ions), a=5.64, sigma=0.1 (rng 5), rc=10, beta=%.6f, kmax=%.4f" % (b.ntot, beta, kmax),
                   "step": "E + F + virial of every cell (real space + self + direct FP64 S(k) sum)",
                   "inputs": "positions and charges resident in HBM"},
        "mixed": {"ms_per_step": 1e3 * t32, "value": b.ntot / t32, "speedup_vs_fp64": t64 / t32,
                  "force_rms_err": float(np.sqrt(np.mean(df ** 2))), "force_max_err": float(np.abs(df).max()),
                  "energy_err_per_atom_max": float(de.max()), "units": "e^2/A"},
        "kernels": kern,
        "roofline": None,
    }
    print(json.dumps(out))


def bench_md(args):
    """NVE molecular dynamics around the potential: Tersoff Si diamond 8^3
    cells (4096 atoms, 600 K) or EAM Cu 40^3 cells (256 000 atoms, 600 K),
    1 fs velocity Verlet, 0.5 A Verlet shell.  A step is one force
    evaluation plus the integrator; steps run as HIP-graph replays of up to
    32 steps, and the same steps step by step for comparison."""
    import numpy as np
    import torch
    from atomistica_amd import native as N
    from atomistica_amd import workloads as W
    kb = 8.617333262e-5
    if args.workload == "md_tersoff_4096":
        pos, cell, pbc, Z = W.diamond_si(8, sigma=0.02, seed=1)
        mass, pot, desc = 28.0855, N.Tersoff(), "Tersoff Si diamond 8^3 cells = 4096 atoms"
    else:
        wl = build_workload("eam_cu_256k")
        pos, cell, pbc, Z = wl["pos"], wl["cell"], wl["pbc"], wl["Z"]
        mass, pot, desc = 63.546, N.TabulatedAlloyEAM(fn=wl["fn"]), "EAM Cu fcc 40^3 cells = 256000 atoms"
    rng = np.random.default_rng(7)
    v0 = rng.normal(0.0, np.sqrt(kb * 600.0 / mass), pos.shape)
    v0 -= v0.mean(axis=0)
    dt, skin, K = 0.1, 0.5, 32

    def run(graph_steps):
        p = N.from_atoms(pos, cell, pbc, Z)
        nl = N.Neighbors(100)
        md = N.MD(pot, p, nl, mass, dt, skin=skin, v0=v0)
        md.run(args.warmup * K, graph_steps)
        torch.cuda.synchronize()
        t = time.perf_counter()
        md.run(args.steps * K, graph_steps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / (args.steps * K), md.state()[3]

    tg, cg = run(K)
    ts, cs = run(0)
    nat = len(pos)
    out = {
        "metric": METRIC, "value": nat / tg, "unit": "atom-evals/s",
        "n_gpus": 1, "steps": args.steps * K, "warmup": args.warmup * K, "ms_per_step": 1e3 * tg,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "NVE MD, %s, 600 K, dt 1.018 fs, Verlet shell 0.5 A" % desc,
                   "step": "velocity Verlet + E + F + virial per step, list rebuilt when an atom moves skin/2",
                   "inputs": "state resident in HBM"},
        "graph": {"steps_per_replay": K, "ms_per_step": 1e3 * tg, "counters": cg},
        "stepwise": {"ms_per_step": 1e3 * ts, "counters": cs, "graph_speedup": ts / tg},
        "roofline": None,
    }
    print(json.dumps(out))


def roofline(workload, kern, nn, nat, steps=None):
    """HBM roofline of the dominant kernel (largest total time): algorithmic
    bytes per launch over its measured average duration, the counter traffic
    of the committed rocprofv3 summary, and its FP64 VALU fraction.  A kernel
    launched k > 1 times per step over disjoint shares of the atoms (one
    launch per row-width class) is taken per step: all atoms over the k
    launches' summed duration."""
    if not kern:
        return None
    dom = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    bpa = algorithmic_bytes(dom, nn)
    if bpa is None:
        return None
    per_step = 1
    if steps and kern[dom]["launches"] > steps and dom.startswith("rebo2_"):
        per_step = kern[dom]["launches"] // steps
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        try:
            ent = json.load(open(tf)).get(workload, {}).get(dom)
            if ent:
                traffic = ent["bytes_per_launch"]
        except ValueError:
            traffic = None
    achieved = bpa * nat / (per_step * kern[dom]["avg_us"] * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
            "algorithmic_bytes_per_atom": bpa, "atoms_per_launch": nat if per_step == 1 else None,
            "atoms_per_step": nat, "launches_per_step": per_step}
    if per_step > 1 and traffic:
        traffic = traffic * per_step          # the summary's bytes are per launch
        roof["traffic"] = traffic
    if traffic:
        roof["traffic_over_algorithmic"] = traffic / (bpa * nat)
    valu = valu_rates(workload, kern)
    if valu and dom in valu:
        roof["valu_frac_fp64"] = valu[dom]["frac_fp64_peak"]
        roof["valu_frac_fp32"] = valu[dom]["frac_fp32_peak"]
    return roof


# algorithmic bytes of one atom-evaluation of the whole E+F+virial pass
# (positions, species, row offset, the list, forces)
def step_bytes(kind, nn):
    return (72 if kind == "eam" else 56) + 8 * nn


def potential_block(workload, args, precision="fp64", autotune=True, cpu=True, mixed=True):
    """One workload through the potential class: full steps (list rebuilt),
    force-only steps, per-kernel HIP-event times, roofline, the mixed form
    against FP64 and the CPU baseline."""
    import numpy as np
    import torch
    from atomistica_amd import native as N
    wl = build_workload(workload)
    if args.shuffle:
        perm = np.random.default_rng(12345).permutation(len(wl["pos"]))
        wl["pos"] = wl["pos"][perm].copy()
        wl["Z"] = wl["Z"][perm].copy()
        if "q" in wl:
            wl["q"] = wl["q"][perm].copy()
        wl["desc"] += "; atoms randomly permuted (rng 12345)"
    run = Runner(wl, precision)
    nat = run.nat
    for _ in range(args.warmup):
        run.step()
    tune = None
    if autotune:
        # launch-shape autotuner (warm-up, outside the timed region)
        rep = run.pot.autotune(run.p, run.nl, reps=5)
        tune = {"report": rep.strip().splitlines(),
                "chosen": [ln.strip() for ln in rep.splitlines() if "->" in ln]}
        for _ in range(2):
            run.step()
    dt = run.timed(args.steps)
    value = nat * args.steps / dt
    total, maxrow = run.nl.get_size()
    nn = total / nat
    # force pass with the list reused
    dt_force = run.timed(args.steps, rebuild=False)
    f64 = run.f.cpu().numpy().copy()
    e64 = float(run.epot.cpu()[0])
    # profiled pass: per-kernel HIP-event durations on the launch stream
    kern = _kernel_times(N, run.step, args.steps)
    B = step_bytes(wl["kind"], nn)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32-arith/f64-acc",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "atoms": nat, "neighbors_per_atom": nn,
                   "step": "neighbour-list rebuild + energy, forces, virial",
                   "inputs": "positions resident in HBM"},
        "force_only": {"value": nat * args.steps / dt_force, "ms_per_step": 1e3 * dt_force / args.steps},
        "step_hbm": {"algorithmic_bytes_per_atom": B, "roofline_rate": HBM_PEAK_GBS * 1e9 / B,
                     "frac_full_step": value * B / (HBM_PEAK_GBS * 1e9),
                     "frac_force_only": nat * args.steps / dt_force * B / (HBM_PEAK_GBS * 1e9)},
        "kernels": kern,
        "roofline": roofline(workload, kern, nn, nat, args.steps),
        "valu": valu_rates(workload, kern),
        "autotune": tune,
    }
    if mixed and precision == "fp64" and wl["kind"] in ("eam", "tersoff", "tersoff_scr", "ewald", "rebo2"):
        from atomistica_amd._lib import AgpuError
        try:
            run.pot.set_precision("mixed")
            run.step()
        except AgpuError as e:        # a class without the mixed form
            out["mixed"] = {"unsupported": str(e)}
            run.pot.set_precision("fp64")
        else:
            for _ in range(args.warmup):
                run.step()
            dtm = run.timed(args.steps)
            dtmf = run.timed(args.steps, rebuild=False)
            fm = run.f.cpu().numpy()
            em = float(run.epot.cpu()[0])
            out["mixed"] = {
                "value": nat * args.steps / dtm, "ms_per_step": 1e3 * dtm / args.steps,
                "force_only_ms": 1e3 * dtmf / args.steps,
                "dtype": "f32-arith/f64-acc",
                "speedup_vs_fp64": (nat * args.steps / dtm) / value,
                "force_only_speedup_vs_fp64": dt_force / dtmf,
                "force_rms_err": float(np.sqrt(np.mean((fm - f64) ** 2))),
                "force_max_err": float(np.abs(fm - f64).max()),
                "energy_err_per_atom": abs(em - e64) / nat,
            }
            run.pot.set_precision("fp64")
    if cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
    del run
    torch.cuda.empty_cache()
    return out


# the Tersoff 1M, REBO2 100k and batched Tersoff cases, carried in the default
# line next to the headline EAM workload
EXTRA_BLOCKS = ("tersoff_si_1m", "rebo2_ch_100k", "tersoff_batch_4096")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="eam_cu_256k")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the named workload (the default line also carries the Tersoff 10^6, "
                         "REBO2 10^5 and batched-Tersoff blocks)")
    ap.add_argument("--shuffle", action="store_true",
                    help="randomly permute the input atoms (the library's own bin ordering must "
                         "make the step independent of the caller's order)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing the variants in warm-up")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one 512 MiB device-to-device copy (reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (single-GPU boxes; no collective path "
              "is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "tersoff_batch_4096":
        bench_batch(args)
        return
    if args.workload == "ewald_batch_256":
        bench_ewald_batch(args)
        return
    if args.workload in ("md_tersoff_4096", "md_eam_256k"):
        bench_md(args)
        return

    import torch
    out = potential_block(args.workload, args, precision=args.precision, autotune=not args.no_autotune,
                          cpu=not args.no_cpu_baseline, mixed=not args.no_mixed)
    if args.workload == "eam_cu_256k" and not args.no_extra:
        for name in EXTRA_BLOCKS:
            print("bench: %s" % name, file=sys.stderr, flush=True)
            if name == "tersoff_batch_4096":
                blk = batch_block(args, cpu=not args.no_cpu_baseline)
            else:
                blk = potential_block(name, args, autotune=not args.no_autotune, cpu=not args.no_cpu_baseline,
                                      mixed=not args.no_mixed)
            for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data", "warmup"):
                blk.pop(k, None)
            out[name] = blk
    if args.copy_calibration:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    print(json.dumps(out))


if __name__ == "__main__":
    main()
