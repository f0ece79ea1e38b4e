"""Gradient-check harness for the hand-written adjoint of H.psi.

E_band(V_loc, R) = sum_n f_n <psi_n | H(V_loc, R) | psi_n> dV for fixed bands psi.
Its gradients come from the C ABI, not from autograd:
  dE_band / dV_loc(r) = rho(r) dV with rho = sum_n f_n |psi_n|^2
      (qm_band_density_*, the SetDensityOne counterpart, qumasun_SetDensityOne.h:9-41)
  dE_band / dR_a     = -F_a with F the nonlocal projector-derivative forces
      (qm_nonlocal_forces_*, ForceNonlocal_bundle, PseudoPotNonlocal_SvF2_kpoint.h:3629-3731)
BandEnergy wraps them in a torch.autograd.Function so torch's gradient
machinery can drive the finite-difference checks; it is used by the tests
only.  Moving atoms re-tabulates the projectors on the frozen voxel sets
(system.retabulate), so E_band is smooth in R and central differences apply.
"""
from __future__ import annotations

import numpy as np

from . import system as _system


def band_energy(H, psi, occ, k_index=0):
    """sum_n f_n <psi_n|H|psi_n> dV (FP64 accumulation on the host of the GPU H.psi)."""
    import torch
    hp = H(psi, k_index=k_index)
    nc = H.sys.ncomp
    prod = (psi.double() * hp.double()).sum(dim=1).cpu().numpy()
    per_band = prod.reshape(-1, nc).sum(axis=1)
    return float(np.dot(np.asarray(occ, dtype=np.float64), per_band) * H.sys.dV)


def set_state(H, base_system, vloc, positions):
    """Point the Hamiltonian at V_loc and atom positions (tables re-evaluated on
    the frozen voxel sets of base_system)."""
    H.set_vloc(vloc)
    s = _system.retabulate(base_system, positions)
    H.set_projectors(s)
    H.sys = s
    return s


def make_band_energy_function():
    import torch

    class BandEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vloc, positions, H, base_system, psi, occ, k_index):
            v = vloc.detach().cpu().numpy()
            R = positions.detach().cpu().numpy()
            set_state(H, base_system, v, R)
            e = band_energy(H, psi, occ, k_index)
            ctx.H, ctx.psi, ctx.occ, ctx.k = H, psi, occ, k_index
            return torch.tensor(e, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            H, psi, occ = ctx.H, ctx.psi, ctx.occ
            rho = H.band_density(psi, occ).cpu().numpy()
            _, forces = H.nonlocal_forces(psi, occ, k_index=ctx.k)
            g = float(grad)
            dv = torch.from_numpy(g * rho * H.sys.dV)
            dr = torch.from_numpy(-g * forces)
            return dv, dr, None, None, None, None, None

    return BandEnergy
