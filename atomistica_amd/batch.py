"""Batched Tersoff for parameter fitting, with parameter gradients.

Many small periodic single-element structures are evaluated together on the
GPU (one concatenated neighbour list, one forward launch), and a hand-written
backward kernel returns

    dL/dtheta = sum_s dL/dE_s dE_s/dtheta + sum_i dL/dF_i . dF_i/dtheta

for the 11 Tersoff parameters THETA_NAMES (the single-element form of the
reference's Tersoff class, bop_kernel.f90 / tersoff_func.f90, with xi =
omega = 1 and no h-term).  :class:`TersoffBatchFunction` exposes this as a
``torch.autograd.Function`` so that a force-matching loss written in torch
differentiates into the parameters.

:func:`ewald_batch` evaluates the Ewald Coulomb sum over the same kind of
batch.  There is no counterpart of batching in the reference, which
evaluates one structure per call.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call

THETA_NAMES = ("A", "B", "lambda", "mu", "beta", "n", "c", "d", "h", "R", "S")


def theta_from_params(params):
    """11-vector from a single-element Tersoff parameter dict (names of
    tersoff_registry.f90; pair-indexed entries take their first value)."""
    def first(k):
        v = params[k]
        return float(v[0] if np.ndim(v) else v)
    return np.array([first("A"), first("B"), first("lambda"), first("mu"), first("beta"),
                     first("n"), first("c"), first("d"), first("h"), first("r1"), first("r2")])


def params_from_theta(theta, el="Si"):
    """Inverse of :func:`theta_from_params` (xi = omega = 1, mubo = 0)."""
    t = [float(x) for x in theta]
    return {"el": [el], "A": [t[0]], "B": [t[1]], "lambda": [t[2]], "mu": [t[3]], "beta": [t[4]],
            "n": [t[5]], "c": [t[6]], "d": [t[7]], "h": [t[8]], "r1": [t[9]], "r2": [t[10]],
            "xi": [1.0], "omega": [1.0], "mubo": [0.0], "m": [1]}


class Batch:
    """A set of periodic structures [(positions (n,3), cell (3,3)), ...],
    concatenated on the GPU (atoms of structure s are offsets[s] ..
    offsets[s+1]-1)."""

    def __init__(self, structures):
        import torch
        from . import native
        native.use_torch_stream()
        h = ctypes.c_void_p()
        call("agpu_batch_create", ctypes.byref(h))
        self._h = h
        counts = [len(p) for p, _ in structures]
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.nstruct = len(structures)
        self.ntot = int(self.offsets[-1])
        cells = np.ascontiguousarray(np.array([np.asarray(c, dtype=np.float64) for _, c in structures]).reshape(-1))
        call("agpu_batch_set_structures", self._h, self.nstruct,
             self.offsets.ctypes.data_as(ctypes.c_void_p), cells.ctypes.data_as(ctypes.c_void_p))
        pos = np.concatenate([np.asarray(p, dtype=np.float64) for p, _ in structures]) if self.ntot else \
            np.zeros((0, 3))
        self.positions = torch.tensor(pos, dtype=torch.float64, device="cuda").contiguous()
        self._keep = (cells,)
        call("agpu_batch_set_positions", self._h, ctypes.c_void_p(self.positions.data_ptr()), 0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib.load().agpu_batch_destroy(h)
            self._h = None

    def pairs(self):
        n = ctypes.c_int64()
        call("agpu_batch_get_size", self._h, None, None, ctypes.byref(n))
        return n.value


class TersoffBatch(Batch):
    """Batched single-element Tersoff over a :class:`Batch` of structures."""

    def forward(self, theta, precision="fp64"):
        """(E per structure (nstruct,), F (ntot, 3)) as torch device tensors."""
        import torch
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(11))
        e = torch.zeros(self.nstruct, dtype=torch.float64, device="cuda")
        f = torch.zeros((self.ntot, 3), dtype=torch.float64, device="cuda")
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        call("agpu_tersoff_batch_forward", self._h, th.ctypes.data_as(ctypes.c_void_p), prec,
             ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(f.data_ptr()), 0)
        return e, f

    def backward(self, theta, grad_e=None, grad_f=None):
        """sum_s grad_e[s] dE_s/dtheta + sum_i grad_f[i] . dF_i/dtheta (numpy, 11)."""
        import torch
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(11))
        ge = None if grad_e is None else grad_e.detach().to("cuda", torch.float64).contiguous()
        gf = None if grad_f is None else grad_f.detach().to("cuda", torch.float64).contiguous()
        out = np.zeros(11)
        call("agpu_tersoff_batch_backward", self._h, th.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_void_p(ge.data_ptr()) if ge is not None else None,
             ctypes.c_void_p(gf.data_ptr()) if gf is not None else None,
             out.ctypes.data_as(ctypes.c_void_p), _lib.HOST)
        return out


def _function():
    import torch

    class TersoffBatchFunction(torch.autograd.Function):
        """(E_struct, F) = f(theta) for a fixed :class:`TersoffBatch`;
        backward returns dL/dtheta through the hand-written kernel."""

        @staticmethod
        def forward(ctx, theta, batch, precision="mixed"):
            th = theta.detach().cpu().numpy().astype(np.float64)
            e, f = batch.forward(th, precision)
            ctx.batch = batch
            ctx.th = th
            ctx.theta_device = theta.device
            return e, f

        @staticmethod
        def backward(ctx, grad_e, grad_f):
            g = ctx.batch.backward(ctx.th, grad_e, grad_f)
            return torch.tensor(g, dtype=torch.float64, device=ctx.theta_device), None, None

    return TersoffBatchFunction


_FN = None


def tersoff_batch(theta, batch, precision="mixed"):
    """Differentiable batched Tersoff: returns (E_struct, F)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(theta, batch, precision)


def ewald_batch(batch, q, cutoff, beta, kmax, precision="fp64", forces=True, virial=True, phi=False):
    """Ewald sum of every structure of ``batch`` with charges q (ntot,), in
    e^2/A: real space + self term (pme.f90:215-321, 386-437, 467-472) and the
    direct FP64 S(k) sum over |k| <= kmax.  Returns torch device tensors
    (E (nstruct,), F (ntot, 3) or None, W (nstruct, 3, 3) or None, phi (ntot,)
    or None); phi = dE/dq, the backward with respect to the charges."""
    import torch
    qd = torch.as_tensor(np.asarray(q, dtype=np.float64), device="cuda").contiguous()
    e = torch.zeros(batch.nstruct, dtype=torch.float64, device="cuda")
    f = torch.zeros((batch.ntot, 3), dtype=torch.float64, device="cuda") if forces else None
    w = torch.zeros((batch.nstruct, 3, 3), dtype=torch.float64, device="cuda") if virial else None
    p = torch.zeros(batch.ntot, dtype=torch.float64, device="cuda") if phi else None
    prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]

    def ptr(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else None
    call("agpu_ewald_batch_compute", batch._h, ptr(qd), float(cutoff), float(beta), float(kmax), prec,
         ptr(e), ptr(f), ptr(w), ptr(p), 0)
    return e, f, w, p
