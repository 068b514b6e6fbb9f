"""torch.autograd.Function over the adjoint-gradient engine.

energy(theta, circuit, H) returns E(theta) = <psi(theta)|H|psi(theta)> with
psi(theta) = U(theta)|0...0>; its backward returns dE/dtheta computed by the
HIP adjoint sweep (qsv_dev_adjoint_gradient), which runs during the forward
call (one forward simulation plus one reverse sweep give both E and the
whole gradient).
"""
from __future__ import annotations

import numpy as np

from . import engine as E
from . import observables as O


def _torch():
    import torch
    return torch


def _make_function():
    torch = _torch()

    class AdjointEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, circuit, H, n, dtype, fuse_k):
            th = theta.detach().to("cpu", torch.float64).numpy()
            cdt = torch.complex128 if dtype == "c128" else torch.complex64
            psi = E.new_state(n, cdt)
            lam = torch.empty_like(psi)
            e, g = O.adjoint_gradient(psi, lam, circuit, H, theta=th, nparams=th.size,
                                      fuse_k=fuse_k)
            ctx.save_for_backward(torch.from_numpy(g).to(theta.device, theta.dtype))
            return theta.new_tensor(e)

        @staticmethod
        def backward(ctx, g_out):
            (g,) = ctx.saved_tensors
            return g_out * g, None, None, None, None, None

    return AdjointEnergy


_FN = None


def energy(theta, circuit, H, n: int, dtype: str = "c128", fuse_k: int = E.AUTO):
    """Differentiable E(theta); `circuit` is a gate list whose parametrised
    gates carry `param` indices into theta."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(theta, circuit, H, n, dtype, fuse_k)


def finite_difference(circuit, H, theta, n: int, h: float = 1e-6, fuse_k: int = E.AUTO,
                      dtype: str = "c128", which=None):
    """Central differences [E(theta + h e_p) - E(theta - h e_p)] / (2h) on
    the GPU, each E a forward simulation from |0...0> plus an FP64
    expectation value.  The default form is the engine's fastest (AUTO:
    compiled tiled passes from 22 qubits up, fused blocks below); every
    shifted circuit has the structure of the first, so its program reuses
    the cached schedule and compiled passes and only its coefficients are
    new."""
    torch = _torch()
    theta = np.asarray(theta, dtype=np.float64)
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    psi = E.new_state(n, cdt)

    def energy_at(th):
        gates = [g.with_theta(th[g.param]) if g.param >= 0 else g for g in circuit]
        prog = E.Program(n, gates, fuse_k=fuse_k, dtype=dtype)
        E.set_basis(psi, 0)
        prog.run(psi)
        return O.expectation(psi, H)[1]

    ps = range(theta.size) if which is None else which
    out = []
    for p in ps:
        tp = theta.copy()
        tp[p] += h
        ep = energy_at(tp)
        tp[p] -= 2 * h
        em = energy_at(tp)
        out.append((ep - em) / (2 * h))
    return np.array(out)
