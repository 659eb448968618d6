"""Batched Tersoff (fitting path) on the GPU through the C-ABI.

Bars:
  - FP64 batched forward against the oracle, structure by structure: energy,
    forces within 1e-10 relative;
  - FP32-arithmetic / FP64-accumulate forward: force RMS <= 1e-5 eV/A;
  - dE/dtheta and d(force-matching loss)/dtheta from the hand-written
    backward kernel against central finite differences of the FP64 forward
    with the relative step h = 1e-6 |theta|, and dE/dA, dE/dB against the
    closed forms E_R / A, E_A / B (E is linear in A and in B);
  - the full 4096-structure set: one forward and one backward, net force
    zero per structure, bitwise reproducible.
"""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W

pytestmark = pytest.mark.gpu

REL = 1e-10


def _theta0():
    from atomistica_amd.batch import theta_from_params
    return theta_from_params(P.Tersoff_Si)


def _structures(n, seed=4, nmax=128):
    return W.random_si_structures(n, seed=seed, nmin=8, nmax=nmax)


def _oracle(theta, pos, cell):
    from atomistica_amd.batch import params_from_theta
    prm = params_from_theta(theta)
    el = np.zeros(len(pos), np.int32)
    nl = O.nl_build(pos, cell, [True] * 3, theta[10])
    return O.tersoff(prm, pos, el, cell, nl)


def test_batch_forward_fp64_parity():
    from atomistica_amd.batch import TersoffBatch
    th = _theta0()
    S = _structures(48)
    b = TersoffBatch(S)
    e, f = b.forward(th, "fp64")
    e = e.cpu().numpy()
    f = f.cpu().numpy()
    print("batch parity: %d structures, %d atoms, %d pairs" % (b.nstruct, b.ntot, b.pairs()), flush=True)
    for s, (pos, cell) in enumerate(S):
        re_, rf, _ = _oracle(th, pos, cell)
        a0, a1 = b.offsets[s], b.offsets[s + 1]
        assert abs(e[s] - re_) <= REL * max(1.0, abs(re_)), s
        scale = max(1.0, np.abs(rf).max())
        assert np.abs(f[a0:a1] - rf).max() <= REL * scale, s


def test_batch_forward_long_rows():
    """Rows longer than one 256-entry chunk (R = 10.5, S = 11 A: ~280
    neighbours) take the unfused per-entry kernels; same parity bar."""
    from atomistica_amd.batch import TersoffBatch, params_from_theta
    th = _theta0()
    th[9], th[10] = 10.5, 11.0
    S = _structures(3, seed=8, nmax=24)
    b = TersoffBatch(S)
    e, f = b.forward(th, "fp64")
    e = e.cpu().numpy()
    f = f.cpu().numpy()
    print("batch long rows: %d atoms, %d pairs" % (b.ntot, b.pairs()), flush=True)
    assert b.pairs() > 256 * b.ntot * 0.9
    prm = params_from_theta(th)
    for s, (pos, cell) in enumerate(S):
        nl = O.nl_build(pos, cell, [True] * 3, th[10], avgn=600)
        re_, rf, _ = O.tersoff(prm, pos, np.zeros(len(pos), np.int32), cell, nl)
        a0, a1 = b.offsets[s], b.offsets[s + 1]
        assert abs(e[s] - re_) <= REL * max(1.0, abs(re_)), s
        assert np.abs(f[a0:a1] - rf).max() <= REL * max(1.0, np.abs(rf).max()), s
    # the unfused backward: dE/dtheta against central differences
    import torch
    g = b.backward(th, torch.ones(b.nstruct, dtype=torch.float64, device="cuda"), None)
    fd = _fd_grad(lambda t: float(b.forward(t, "fp64")[0].sum().item()), th, range(11))
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def test_batch_forward_mixed_error():
    from atomistica_amd.batch import TersoffBatch
    th = _theta0()
    b = TersoffBatch(_structures(256, seed=5, nmax=512))
    e64, f64 = b.forward(th, "fp64")
    e32, f32 = b.forward(th, "mixed")
    df = (f32 - f64).cpu().numpy()
    rms = np.sqrt(np.mean(df ** 2))
    de = np.abs((e32 - e64).cpu().numpy()).max()
    print("batch mixed: F rms %.3e  F max %.3e  max |dE_s| %.3e" % (rms, np.abs(df).max(), de), flush=True)
    assert rms <= 1e-5


def _fd_grad(fun, th, idx):
    g = np.zeros(len(idx))
    for q, k in enumerate(idx):
        h = 1e-6 * abs(th[k])
        tp = th.copy()
        tp[k] += h
        tm = th.copy()
        tm[k] -= h
        g[q] = (fun(tp) - fun(tm)) / (2 * h)
    return g


def test_batch_energy_gradient_fd_and_closed_forms():
    import torch
    from atomistica_amd.batch import TersoffBatch
    th = _theta0()
    S = _structures(32, seed=6, nmax=96)
    b = TersoffBatch(S)
    ge = torch.ones(b.nstruct, dtype=torch.float64, device="cuda")
    g = b.backward(th, ge, None)

    def etot(t):
        return float(b.forward(t, "fp64")[0].sum().item())
    fd = _fd_grad(etot, th, range(11))
    print("dE/dtheta", g, "\nfd       ", fd, flush=True)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    # E is linear in A and B: dE/dA = E_R / A, dE/dB = E_A / B
    E = etot(th)
    tB = th.copy()
    tB[1] = 2.0 * th[1]
    E_A = etot(tB) - E                 # attractive part at B
    E_R = E - E_A
    assert abs(g[0] - E_R / th[0]) <= 1e-9 * abs(E_R / th[0])
    assert abs(g[1] - E_A / th[1]) <= 1e-9 * abs(E_A / th[1])


def test_batch_force_matching_loss_gradient():
    import torch
    from atomistica_amd.batch import TersoffBatch, tersoff_batch
    th0 = _theta0()
    rng = np.random.default_rng(4)
    th_ref = th0 * (1 + 0.01 * rng.standard_normal(11))
    S = _structures(24, seed=7, nmax=96)
    b = TersoffBatch(S)
    _, f_ref = b.forward(th_ref, "fp64")
    e_ref = b.forward(th_ref, "fp64")[0]

    def loss_np(t):
        e, f = b.forward(t, "fp64")
        return float(((f - f_ref) ** 2).sum().item() + 0.1 * ((e - e_ref) ** 2).sum().item())

    theta = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    e, f = tersoff_batch(theta, b, "fp64")
    L = ((f - f_ref) ** 2).sum() + 0.1 * ((e - e_ref) ** 2).sum()
    L.backward()
    g = theta.grad.numpy()
    fd = _fd_grad(loss_np, th0, range(11))
    print("dL/dtheta", g, "\nfd       ", fd, flush=True)
    assert np.allclose(g, fd, rtol=1e-4, atol=1e-6 * np.abs(fd).max())


def test_batch_full_set_4096():
    import time
    import torch
    from atomistica_amd.batch import TersoffBatch
    th = _theta0()
    S = W.random_si_structures(4096, seed=4)
    b = TersoffBatch(S)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e, f = b.forward(th, "mixed")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gf = torch.randn_like(f)
    g1 = b.backward(th, torch.ones_like(e), gf)
    t2 = time.perf_counter()
    g2 = b.backward(th, torch.ones_like(e), gf)
    print("batch 4096: %d atoms %d pairs  forward %.2f ms  backward %.2f ms" %
          (b.ntot, b.pairs(), 1e3 * (t1 - t0), 1e3 * (t2 - t1)), flush=True)
    assert np.all(np.isfinite(g1)) and np.array_equal(g1, g2)
    # net force per structure: zero to FP64 rounding in the FP64 form, to
    # FP32 rounding (forces of a few eV/A summed over up to 512 atoms) in
    # the mixed form
    f64 = b.forward(th, "fp64")[1].cpu().numpy()
    fn = f.cpu().numpy()
    for s in range(0, b.nstruct, 97):
        a0, a1 = b.offsets[s], b.offsets[s + 1]
        assert np.abs(f64[a0:a1].sum(0)).max() < 1e-9
        assert np.abs(fn[a0:a1].sum(0)).max() < 1e-4
