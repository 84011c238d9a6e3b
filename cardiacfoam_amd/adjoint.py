"""Gradients of probe traces of the explicit monodomain step through the hand-written
adjoint kernels of libcardiacgpu.so, behind torch.autograd.Function.

``TissueAdjoint(dom, probe_cells, checkpoint_every)`` snapshots a ``Monodomain`` (explicit,
Euler ddt, BuenoOrovio, FP64, uniform diagonal conductivity) as the start state.
``probe_traces(adj, kappa, theta, n_steps)`` returns Vm (V) after each step at the probe
cells, [n_steps, n_probes], as a differentiable function of the per-cell conductivity
scale ``kappa`` (CUDA float64 [n_cells]) and ``theta`` = (tau_fi, tau_si) (float64 [2]).
Any torch loss of the traces can then be back-propagated: the backward pass re-runs
checkpointed segments and sweeps them in reverse on the device.  The device keeps the
checkpoints of one forward run per handle: a backward pass for an earlier run (a second
probe_traces call on the same handle before backward) raises instead of returning the
later run's gradient.
"""
import ctypes as C

import numpy as np

from ._capi import check, lib

THETA_NAMES = ("tau_fi", "tau_si")


class TissueAdjoint:
    def __init__(self, dom, probe_cells, checkpoint_every=100):
        self.dom = dom
        self.n = dom.mesh.n
        cells = np.ascontiguousarray(probe_cells, dtype=np.int64)
        self.n_probes = int(cells.size)
        a = C.c_void_p()
        check(lib().cg_adjoint_create(dom.t, self.n_probes, cells.ctypes.data_as(C.c_void_p),
                                      int(checkpoint_every), C.byref(a)))
        self.a = a
        self.probe_cells = cells

    def __del__(self):
        if getattr(self, "a", None):
            try:
                lib().cg_adjoint_destroy(self.a)
            except Exception:
                pass
            self.a = None

    generation = 0      # number of the latest forward run on the device (0: none)

    def forward(self, kappa, theta, n_steps, out):
        """runs the forward map; returns its generation number"""
        import torch
        torch.cuda.current_stream().synchronize()   # kappa written by torch's stream
        th = None
        if theta is not None:
            th_host = np.ascontiguousarray(theta.detach().cpu().numpy(), dtype=np.float64)
            th = th_host.ctypes.data_as(C.c_void_p)
        gen = C.c_int64()
        self.generation = -1                        # a failed forward leaves no valid run
        check(lib().cg_adjoint_forward(self.a, C.c_void_p(kappa.data_ptr()), th, int(n_steps),
                                       C.c_void_p(out.data_ptr()), C.byref(gen)))
        self.generation = gen.value
        return gen.value

    def check_backward(self, generation):
        """raises unless `generation` is the latest forward run of this handle"""
        if generation != self.generation:
            raise RuntimeError(
                "TissueAdjoint: backward of forward run %d, but run %d replaced its "
                "checkpoints on the device; call backward before the next probe_traces on "
                "the same handle (or use one TissueAdjoint per pending graph)"
                % (generation, self.generation))

    def backward(self, dprobes, dkappa, generation, n_steps):
        import torch
        self.check_backward(generation)
        torch.cuda.current_stream().synchronize()
        dth = np.zeros(2)
        check(lib().cg_adjoint_backward(self.a, int(generation), int(n_steps),
                                        C.c_void_p(dprobes.data_ptr()),
                                        C.c_void_p(dkappa.data_ptr()),
                                        dth.ctypes.data_as(C.c_void_p)))
        return dth


def _function():
    import torch

    class ProbeTraces(torch.autograd.Function):
        @staticmethod
        def forward(ctx, kappa, theta, adj, n_steps):
            if kappa.dtype != torch.float64 or not kappa.is_cuda or kappa.numel() != adj.n:
                raise ValueError("kappa must be a CUDA float64 tensor of n_cells values")
            kappa = kappa.contiguous()
            out = torch.empty(n_steps, adj.n_probes, dtype=torch.float64, device=kappa.device)
            ctx.generation = adj.forward(kappa, theta, n_steps, out)
            ctx.n_steps = n_steps
            ctx.adj = adj
            ctx.theta_dev = None if theta is None else theta.device
            return out

        @staticmethod
        def backward(ctx, grad_out):
            adj = ctx.adj
            adj.check_backward(ctx.generation)
            g = grad_out.contiguous().to(torch.float64)
            dkappa = torch.empty(adj.n, dtype=torch.float64, device=g.device)
            dth = adj.backward(g, dkappa, ctx.generation, ctx.n_steps)
            dtheta = None
            if ctx.theta_dev is not None:
                dtheta = torch.tensor(dth, dtype=torch.float64, device=ctx.theta_dev)
            return dkappa, dtheta, None, None

    return ProbeTraces


_FN = None


def probe_traces(adj, kappa, theta, n_steps):
    """Vm at the probes after each of n_steps steps, differentiable in kappa and theta."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(kappa, theta, adj, int(n_steps))
