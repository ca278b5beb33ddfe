"""Nuclear gradients behind torch.autograd.

``scf_energy(coords, species, nalpha, nbeta, ...)`` returns the SCF total
energy as a differentiable torch scalar of the (natoms, 3) coordinate tensor
(bohr).  Forward runs ``pf_scf_gradient`` (SCF + analytic gradient on the GPU
through the C-ABI) and keeps dE/dR; backward scales it by the incoming
gradient.  PyTorch is used only for the autograd bookkeeping: all integrals,
the SCF and the derivative contractions run in the HIP library.
"""
import ctypes

import numpy as np

from . import _native


def scf_gradient(species, coords_bohr, nalpha, nbeta, basis="sto-3g", options=None, want_terms=False, **kw):
    """(energy, gradient (natoms, 3), result dict) from pf_scf_gradient.

    With ``want_terms`` the result dict also holds ``D`` (converged density,
    (nd, nbf, nbf)) and ``terms`` (5, natoms, 3): nuclear repulsion, kinetic,
    nuclear attraction, overlap (-W dS) and two-electron contributions."""
    opt = options if options is not None else _native.default_options(**kw)
    m = _native._Mol(species, coords_bohr)
    g = np.zeros((m.n, 3))
    r = _native.Result()
    terms = D = None
    if want_terms:
        nbf = _native.basis_info(species, coords_bohr, basis, bool(opt.renormalize))["nbf"]
        D = np.zeros((1 if nalpha == nbeta else 2, nbf, nbf))
        terms = np.zeros((5, m.n, 3))
    _native._check(_native.load_library().pf_scf_gradient(
        *m.args(), basis.encode(), int(nalpha), int(nbeta), ctypes.byref(opt), ctypes.byref(r),
        g.ctypes.data_as(_native._pd), _native._ptr(D), _native._ptr(terms)))
    out = r.as_dict()
    if want_terms:
        out["D"], out["terms"] = D, terms
    return r.energy, g, out


def _function():
    import torch

    class SCFEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, coords, species, nalpha, nbeta, basis, options):
            # pf_scf_gradient fails (PyfaccError) when the SCF does not converge,
            # so autograd never receives a gradient of a non-stationary density
            e, g, _ = scf_gradient(species, coords.detach().cpu().double().numpy(), nalpha, nbeta, basis, options)
            ctx.save_for_backward(torch.as_tensor(g, dtype=torch.float64, device=coords.device))
            ctx.in_dtype = coords.dtype
            return torch.tensor(e, dtype=torch.float64, device=coords.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return (grad_out.to(torch.float64) * g).to(ctx.in_dtype), None, None, None, None, None

    return SCFEnergy


_FN = None


def scf_energy(coords, species, nalpha, nbeta, basis="sto-3g", options=None, **kw):
    """Differentiable SCF energy (a float64 scalar) of a (natoms, 3) torch
    tensor in bohr.  Raises if the SCF does not converge."""
    global _FN
    if _FN is None:
        _FN = _function()
    opt = options if options is not None else _native.default_options(**kw)
    return _FN.apply(coords, list(species), int(nalpha), int(nbeta), basis, opt)
