"""SCF loop around the GPU LOBPCG solve: the host-side mirror of
QUMASUN_KPOINT::mSCF (qumasun_kpoint.scfloop.h:96-222).

Each step, in the reference's order:
  1. LOBPCG on every k-point (mSolveLOBPCG, qumasun_kpoint.lobpcg.h:15-121):
     LOBPCG_STEP_FIRST iterations on the first SCF step, LOBPCG_STEP_PER_SCF
     after, the P block kept between calls;
  2. occupations (mSetOccupancy: aufbau, or Fermi-Dirac at kT > 0);
  3. the density from the bands (mSetDensityByPsi / SetDensityOne);
  4. mixing (m_mixer->Predict; the reference's default is DIIS with 9 fields
     and ratio 0.25), the change sum|rho_out - rho_in| dV / N_e returned;
  5. the potential V_tot = V_ion + V_H + V_xc (mSetPotentialVhart / Vxc /
     Vtot) for the next H.psi;
  6. the total energy (mGetTotalEnergy): E_kin + E_ion + E_H + E_xc + E_nl,
     E_kin from the band energies, E_kin = sum f e - <V_in rho_out> - E_nl,
     the field terms on the mixed density as the reference evaluates them;
  7. convergence on the density change and / or the energy change.
The compute is the HIP library's: qm_lobpcg_solve, qm_band_density_*,
qm_nonlocal_forces_* (E_nl), qm_mixer_*, qm_hartree_f64, qm_xc_lda_f64,
qm_grid_dot_f64, qm_occupancy.  V_ion is the system's local potential (the
synthetic workloads' Gaussian wells stand in for the pseudopotential's local
part and the ionic Hartree term).  Spin unpolarised, LDA, no core correction.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .hamiltonian import Hamiltonian
from .lobpcg import LOBPCG

MIX_MODES = {"simple": 0, "diis": 1}


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def occupancy(eigs, kweights, nelec, kbT=0.0):
    """(occupations nk x nb, mu) by qm_occupancy; eigs nk x nb ascending"""
    e = np.ascontiguousarray(np.atleast_2d(eigs), dtype=np.float64)
    w = np.ascontiguousarray(kweights, dtype=np.float64)
    occ = np.zeros_like(e)
    mu = ctypes.c_double(0.0)
    _lib.check(_lib.load().qm_occupancy(e.shape[0], e.shape[1], e.ctypes.data_as(ctypes.c_void_p),
                                        w.ctypes.data_as(ctypes.c_void_p), float(nelec), float(kbT),
                                        occ.ctypes.data_as(ctypes.c_void_p), ctypes.byref(mu)))
    return occ, mu.value


class Mixer:
    """qm_mixer_* on device fields"""

    def __init__(self, hamiltonian, mode="diis", history=9, ratio=0.25):
        self.lib = hamiltonian.lib
        self.H = hamiltonian
        h = ctypes.c_void_p()
        _lib.check(self.lib.qm_mixer_create(ctypes.byref(h), hamiltonian.ctx, MIX_MODES[mode], int(history),
                                            float(ratio)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.qm_mixer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, rho):
        import torch
        _lib.check(self.lib.qm_mixer_init(self.h, _p(rho), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def predict(self, rho):
        """rho (the density from the bands) is replaced by the mixed density;
        returns sum |rho_out - rho_in|"""
        import torch
        d = ctypes.c_double(0.0)
        _lib.check(self.lib.qm_mixer_predict(self.h, _p(rho), ctypes.byref(d),
                                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return d.value


def hartree(H, rho, out=None):
    import torch
    out = torch.empty_like(rho) if out is None else out
    _lib.check(H.lib.qm_hartree_f64(H.ctx, _p(rho), _p(out), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out


def xc_lda(H, rho, vxc=None, exc=None):
    import torch
    vxc = torch.empty_like(rho) if vxc is None else vxc
    e = ctypes.c_void_p(0) if exc is None else _p(exc)
    _lib.check(H.lib.qm_xc_lda_f64(H.ctx, _p(rho), _p(vxc), e, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return vxc


def grid_dot(H, a, b):
    import torch
    r = ctypes.c_double(0.0)
    _lib.check(H.lib.qm_grid_dot_f64(H.ctx, _p(a), _p(b), ctypes.byref(r),
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return r.value


class SCF:
    """Kohn-Sham SCF on one GPU for `system` with `nelec` electrons."""

    def __init__(self, system, nelec, *, precision=64, kbT=0.0, mixing="diis", history=9, ratio=0.25,
                 kweights=None, lobpcg_first=20, lobpcg_per_scf=4, residual_threshold=None, device=0):
        import torch
        self.torch = torch
        self.s = system
        self.nelec = float(nelec)
        self.kbT = kbT
        self.nk = system.kpoints.shape[0]
        self.kweights = np.ones(self.nk) if kweights is None else np.asarray(kweights, dtype=np.float64)
        self.steps = (lobpcg_first, lobpcg_per_scf)
        self.H = Hamiltonian(system, device)
        dt = torch.float64 if precision == 64 else torch.float32
        self.solvers = [LOBPCG(self.H, system.nbands, precision=precision, residual_threshold=residual_threshold)
                        for _ in range(self.nk)]
        x0 = system.initial_bands()
        self.psi = [torch.from_numpy(x0).to(f"cuda:{device}", dt) for _ in range(self.nk)]
        self.vion = torch.from_numpy(np.ascontiguousarray(system.vloc, dtype=np.float64)).to(f"cuda:{device}")
        self.dV = system.dV
        volume = system.ng * self.dV
        self.rho = torch.full_like(self.vion, self.nelec / volume)  # uniform start
        self.mixer = Mixer(self.H, mixing, history, ratio)
        self.mixer.init(self.rho)
        self.ones = torch.ones_like(self.rho)
        self.vh = torch.empty_like(self.rho)
        self.vxc = torch.empty_like(self.rho)
        self.exc = torch.empty_like(self.rho)
        self._set_potential()
        self.step_index = 0
        self.energy = None
        self.history = []

    def close(self):
        for L in self.solvers:
            L.close()
        self.mixer.close()
        self.H.close()

    def _set_potential(self):
        hartree(self.H, self.rho, self.vh)
        xc_lda(self.H, self.rho, self.vxc, self.exc)
        self.vtot = self.vion + self.vh + self.vxc
        self.H.set_vloc(self.vtot.cpu().numpy())

    def step(self):
        first = self.step_index == 0
        iters = self.steps[0] if first else self.steps[1]
        eigs = []
        for k, L in enumerate(self.solvers):
            if L.residual_threshold is not None:
                L.reset_locking()
            ev, _ = L.solve(self.psi[k], iters, k_index=k, first_time=first)
            eigs.append(ev)
        eigs = np.array(eigs)
        occ, mu = occupancy(eigs, self.kweights, self.nelec, self.kbT)
        rho = self.torch.zeros_like(self.rho)
        e_nl = 0.0
        for k in range(self.nk):
            self.H.band_density(self.psi[k], occ[k], rho=rho)
            e_nl += self.H.nonlocal_forces(self.psi[k], occ[k], k_index=k)[0]
        band = float((occ * eigs).sum())
        e_kin = band - grid_dot(self.H, self.vtot, rho) * self.dV - e_nl
        diff = self.mixer.predict(rho) * self.dV / self.nelec
        self.rho = rho
        self._set_potential()
        e_ion = grid_dot(self.H, self.vion, self.rho) * self.dV
        e_h = 0.5 * grid_dot(self.H, self.vh, self.rho) * self.dV
        e_xc = grid_dot(self.H, self.exc, self.ones) * self.dV
        prev = self.energy
        self.energy = e_kin + e_ion + e_h + e_xc + e_nl
        rec = {"step": self.step_index, "E_tot": self.energy, "E_kin": e_kin, "E_ion": e_ion, "E_H": e_h,
               "E_xc": e_xc, "E_nl": e_nl, "band": band, "diff_rho": diff, "mu": mu, "eigs": eigs, "occ": occ,
               "dE": (prev - self.energy) if prev is not None else math.nan}
        self.history.append(rec)
        self.step_index += 1
        return rec

    def run(self, max_steps, threshold_rho=1e-6, threshold_energy=None):
        """steps until the density change (and, if given, |dE|) is below its threshold"""
        for _ in range(max_steps):
            r = self.step()
            ok = r["diff_rho"] < threshold_rho
            if threshold_energy is not None:
                ok = ok and abs(r["dE"]) < threshold_energy
            if ok:
                break
        return self.history[-1]
