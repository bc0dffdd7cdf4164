"""C_ext as a differentiable function of the dipole permittivities and
positions (no reference counterpart).

``cext(eps, positions, ...)`` is a ``torch.autograd.Function``: the forward
runs the device solve A p = E, the backward returns the adjoint gradients of
``lh_pairsum_adjoint_grad`` (one extra solve, A lam = conj(E)). The adjoint
is computed together with the forward, since both need the same session.

eps: complex128 tensor (N,), one permittivity per dipole; positions: float64
tensor (N, 3) in nm. Gradients follow torch's convention for a real loss of
a complex input: d/dRe(eps) + i d/dIm(eps).
"""
from __future__ import annotations

import numpy as np
import torch

from . import inputs as I
from .solver import FP64, HipBackend, IncidentField, PairSumSession


class CextFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, positions, backend, wavelength_nm, n_env, spacing, cell_size,
                incident, tol, max_iter, precision):
        e = eps.detach().cpu().numpy().astype(np.complex128)
        pos = positions.detach().cpu().numpy().astype(np.float64)
        alphas = np.array([I.polarisability_from_epsilon(complex(x), wavelength_nm, spacing, n_env)
                           for x in e])
        dainv = np.array([I.inverse_polarisability_derivative(complex(x), spacing, n_env)
                          for x in e])
        k = I.wavenumber(wavelength_nm, n_env)
        s = PairSumSession(backend, pos, alphas, k, True, cell_size, precision=precision)
        try:
            cs, ge, gp, fi, ai = s.adjoint_grad(incident, tol, max_iter, dainv, positions=True)
        finally:
            s.close()
        ctx.save_for_backward(torch.from_numpy(ge[:, 0] + 1j * ge[:, 1]), torch.from_numpy(gp))
        ctx.infos = (fi, ai)
        return torch.tensor(cs[0], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g_eps, g_pos = ctx.saved_tensors
        return (grad_out * g_eps, grad_out * g_pos) + (None,) * 9


def cext(eps, positions, backend: HipBackend, wavelength_nm: float, n_env: float = 1.33,
         spacing: float = 2.0, cell_size: float | None = None,
         incident: IncidentField | None = None, tol: float = 1e-10, max_iter: int = 1000,
         precision: int = FP64):
    """C_ext (nm^2) of the dipole set, differentiable in eps and positions."""
    return CextFunction.apply(eps, positions, backend, wavelength_nm, n_env, spacing,
                              cell_size if cell_size is not None else spacing,
                              incident or IncidentField(), tol, max_iter, precision)
