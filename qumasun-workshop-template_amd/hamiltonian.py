"""Host-side mirror of the reference's H.psi operator on the GPU.

`Hamiltonian` plays the role of QUMASUN_BASE1::mHamiltonianMatrix_ddm_bundle
(qumasun_base1.h:320, qumasun_base1.matrix.h:141-254): it owns the grid, V_tot
and the projector tables, and `apply(hpsi, psi, nb, k_index)` is the operator
slot OpeA(Ax, x, num) that LOBPCG::Eigen_z_multi_mpi calls
(lobpcg_z_multi_mpi.h:1732-1742; the caller lambda is qumasun_kpoint.lobpcg.h:97-99).
Arrays are torch tensors on the GPU in the reference's SoA band layout, shape
(nb * ncomp, Ng); the compute is the HIP library behind the C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _stream(torch, device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def subspace_force(route):
    """Route of the subspace eigensolve's tridiagonal reduction (process-wide):
    -1 automatic, 0 never the cooperative multi-workgroup reduction, 1 that
    reduction wherever it fits (qm_subspace_force)."""
    _lib.check(_lib.load().qm_subspace_force(int(route)))


def coop_spin_limit(polls=None):
    """Bound of the cooperative kernels' grid-barrier waits in polls,
    process-wide (qm_coop_set_spin_limit); None restores the default."""
    L = _lib.load()
    _lib.check(L.qm_coop_set_spin_limit(int(L.qm_coop_spin_limit_default() if polls is None else polls)))


class Hamiltonian:
    def __init__(self, system, device=0):
        self.lib = _lib.load()
        self.sys = system
        self.device = device
        ctx = ctypes.c_void_p()
        hx, hy, hz = system.h
        _lib.check(self.lib.qm_create(ctypes.byref(ctx), system.nx, system.ny, system.nz, hx, hy, hz, device))
        self.ctx = ctx
        self.set_vloc(system.vloc)
        self.set_projectors(system)
        self.set_kpoints(system.kpoints)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.qm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------
    def set_vloc(self, vloc):
        v = np.ascontiguousarray(vloc, dtype=np.float64)
        assert v.size == self.sys.ng
        _lib.check(self.lib.qm_set_vloc(self.ctx, _ptr(v)))

    def set_projectors(self, s):
        arrs = [np.ascontiguousarray(s.nproj, np.int32), np.ascontiguousarray(s.vox_off, np.int64),
                np.ascontiguousarray(s.vox_xyz, np.int32), np.ascontiguousarray(s.table, np.float64),
                np.ascontiguousarray(s.dtable, np.float64) if s.dtable is not None else None,
                np.ascontiguousarray(s.tab_off, np.int64), np.ascontiguousarray(s.energy, np.float64),
                np.ascontiguousarray(s.proj_off, np.int32)]
        ptrs = [(_ptr(a) if a is not None else None) for a in arrs]
        _lib.check(self.lib.qm_set_projectors(self.ctx, s.natoms, *ptrs))

    def set_kinetic(self, mode):
        """Kinetic term of later applies: "fd8" (8th-order finite differences,
        the default) or "fft" (spectral, the reference's default in
        mHamiltonianMatrix_ddm_bundle)."""
        _lib.check(self.lib.qm_set_kinetic(self.ctx, {"fd8": 0, "fft": 1}[mode]))

    def set_kpoints(self, kpoints):
        k = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
        _lib.check(self.lib.qm_set_kpoints(self.ctx, k.shape[0], _ptr(k)))

    # -- operator ------------------------------------------------------------
    def apply(self, hpsi, psi, nb=None, k_index=0, ncomp=None, part="full"):
        """hpsi <- H(k) psi on the current torch stream (hpsi overwritten).
        part="local": kinetic + V_loc only (overwrites); part="nonlocal": hpsi += Vnl psi."""
        import torch
        ncomp = self.sys.ncomp if ncomp is None else ncomp
        nb = psi.shape[0] // ncomp if nb is None else nb
        assert psi.is_cuda and hpsi.is_cuda and psi.is_contiguous() and hpsi.is_contiguous()
        assert psi.dtype == hpsi.dtype and psi.numel() >= nb * ncomp * self.sys.ng
        names = {"full": "qm_hpsi_apply", "local": "qm_local_apply", "nonlocal": "qm_nonlocal_add"}[part]
        fn = {torch.float64: getattr(self.lib, names + "_f64"),
              torch.float32: getattr(self.lib, names + "_f32")}[psi.dtype]
        _lib.check(fn(self.ctx, ctypes.c_void_p(hpsi.data_ptr()), ctypes.c_void_p(psi.data_ptr()), nb, ncomp,
                      k_index, _stream(torch, psi.device)))
        return hpsi

    def __call__(self, psi, k_index=0, ncomp=None):
        import torch
        out = torch.empty_like(psi)
        return self.apply(out, psi, k_index=k_index, ncomp=ncomp)

    # -- stencil tile-shape autotuner ------------------------------------------
    def stencil_candidates(self):
        """Descriptions of the stencil tile shapes the autotuner chooses from."""
        buf = ctypes.create_string_buffer(256)
        names = []
        for i in range(self.lib.qm_stencil_num_candidates()):
            _lib.check(self.lib.qm_stencil_candidate_name(i, buf, len(buf)))
            names.append(buf.value.decode())
        return names

    def stencil_force(self, candidate):
        """Use stencil candidate `candidate` for every later apply (-1: autotune)."""
        _lib.check(self.lib.qm_stencil_force(self.ctx, candidate))

    def stencil_choice(self, precision, nb, ncomp=None, kt=False, full=False):
        """Candidate index in use for this apply shape (-1 if not tuned yet);
        full=True: the shape of full applies (stencil beside the nonlocal
        projection), tuned separately from local-only applies."""
        ncomp = self.sys.ncomp if ncomp is None else ncomp
        out = ctypes.c_int()
        _lib.check(self.lib.qm_stencil_choice(self.ctx, precision, ncomp, int(kt) | (2 if full else 0), nb,
                                              ctypes.byref(out)))
        return out.value

    # -- nonlocal projection shape autotuner ---------------------------------
    def nonlocal_force(self, tiles):
        """Column tiles (1, 2, 4) per projection workgroup for later applies (0: autotune)."""
        _lib.check(self.lib.qm_nonlocal_force(self.ctx, tiles))

    def nonlocal_choice(self, precision, nb, ncomp=None, kt=False):
        """Column tiles in use for this projection shape (-1 if not tuned yet)."""
        ncomp = self.sys.ncomp if ncomp is None else ncomp
        out = ctypes.c_int()
        _lib.check(self.lib.qm_nonlocal_choice(self.ctx, precision, ncomp, int(kt), nb, ctypes.byref(out)))
        return out.value

    # -- dense products (the blas_DGEMM_t / blas_DGEMM_n sub-boundary) ------
    def gram(self, A, B):
        """C = A^T B (FP64) for column-major A (k x m) and B (k x n) given as
        torch tensors of shape (m, k) and (n, k) (one row per column)."""
        import torch
        assert A.dtype == B.dtype and A.is_contiguous() and B.is_contiguous() and A.shape[1] == B.shape[1]
        m, k = A.shape
        n = B.shape[0]
        C = torch.empty((n, m), dtype=torch.float64, device=A.device)  # column-major m x n
        fn = {torch.float64: self.lib.qm_gram_f64, torch.float32: self.lib.qm_gram_f32}[A.dtype]
        _lib.check(fn(self.ctx, ctypes.c_void_p(A.data_ptr()), m, ctypes.c_void_p(B.data_ptr()), n, k,
                      ctypes.c_void_p(C.data_ptr()), _stream(torch, A.device)))
        return C.T  # (m, n) view

    def update(self, Z, C):
        """Y = Z C for Z as (m, k) rows-of-columns and C an (m, n) matrix; returns Y as (n, k)."""
        import torch
        m, k = Z.shape
        n = C.shape[1]
        Cc = C.T.contiguous().to(Z.dtype)  # column-major m x n
        Y = torch.empty((n, k), dtype=Z.dtype, device=Z.device)
        fn = {torch.float64: self.lib.qm_update_f64, torch.float32: self.lib.qm_update_f32}[Z.dtype]
        _lib.check(fn(self.ctx, ctypes.c_void_p(Z.data_ptr()), m, ctypes.c_void_p(Cc.data_ptr()), n, k,
                      ctypes.c_void_p(Y.data_ptr()), _stream(torch, Z.device)))
        return Y

    def sygvd(self, A, B, nev=None):
        """A z = l B z (FP64, symmetric A, SPD B; upper triangles referenced).
        Returns (all eigenvalues ascending, Z of shape (n, nev) with the lowest
        nev eigenvectors, U the Cholesky factor of B in its upper triangle), as
        device tensors."""
        import torch
        n = A.shape[0]
        nev = n if nev is None else nev
        assert A.dtype == torch.float64 and B.dtype == torch.float64 and A.shape == B.shape == (n, n)
        # the row-major buffer of A^T is A in column-major order
        At = A.T.contiguous()
        Bt = B.T.contiguous()
        w = torch.empty(n, dtype=torch.float64, device=A.device)
        Zt = torch.empty((nev, n), dtype=torch.float64, device=A.device)
        Ut = torch.empty((n, n), dtype=torch.float64, device=A.device)
        _lib.check(self.lib.qm_sygvd_f64(self.ctx, n, ctypes.c_void_p(At.data_ptr()), ctypes.c_void_p(Bt.data_ptr()),
                                         ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(Zt.data_ptr()),
                                         ctypes.c_void_p(Ut.data_ptr()), nev, _stream(torch, A.device)))
        return w, Zt.T, Ut.T

    def hegvd(self, A, B, nev=None):
        """A z = l B z in complex FP64 (Hermitian A, HPD B; upper triangles
        referenced), qm_hegvd_z64.  Returns (all eigenvalues ascending, Z of
        shape (n, nev), U the Cholesky factor of B in its upper triangle)."""
        import torch
        n = A.shape[0]
        nev = n if nev is None else nev
        assert A.dtype == torch.complex128 and B.dtype == torch.complex128 and A.shape == B.shape == (n, n)
        At = A.T.contiguous()
        Bt = B.T.contiguous()
        w = torch.empty(n, dtype=torch.float64, device=A.device)
        Zt = torch.empty((nev, n), dtype=torch.complex128, device=A.device)
        Ut = torch.empty((n, n), dtype=torch.complex128, device=A.device)
        _lib.check(self.lib.qm_hegvd_z64(self.ctx, n, ctypes.c_void_p(At.data_ptr()), ctypes.c_void_p(Bt.data_ptr()),
                                         ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(Zt.data_ptr()),
                                         ctypes.c_void_p(Ut.data_ptr()), nev, _stream(torch, A.device)))
        return w, Zt.T, Ut.T

    # -- adjoint -------------------------------------------------------------
    def band_density(self, psi, occ, ncomp=None, rho=None):
        """rho += sum_n occ_n |psi_n|^2 (FP64, device); dE_band/dV_loc = rho * dV."""
        import torch
        ncomp = self.sys.ncomp if ncomp is None else ncomp
        nb = psi.shape[0] // ncomp
        if rho is None:
            rho = torch.zeros(self.sys.ng, dtype=torch.float64, device=psi.device)
        o = np.ascontiguousarray(occ, dtype=np.float64)
        assert o.size == nb
        fn = {torch.float64: self.lib.qm_band_density_f64, torch.float32: self.lib.qm_band_density_f32}[psi.dtype]
        _lib.check(fn(self.ctx, ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(psi.data_ptr()), _ptr(o), nb,
                      ncomp, _stream(torch, psi.device)))
        return rho

    def nonlocal_forces(self, psi, occ, k_index=0, ncomp=None):
        """(E_nl, forces (natoms, 3)) with forces = -dE_nl/dR."""
        import torch
        ncomp = self.sys.ncomp if ncomp is None else ncomp
        nb = psi.shape[0] // ncomp
        o = np.ascontiguousarray(occ, dtype=np.float64)
        f = np.zeros((self.sys.natoms, 3))
        e = np.zeros(1)
        fn = {torch.float64: self.lib.qm_nonlocal_forces_f64,
              torch.float32: self.lib.qm_nonlocal_forces_f32}[psi.dtype]
        _lib.check(fn(self.ctx, ctypes.c_void_p(psi.data_ptr()), _ptr(o), nb, ncomp, k_index, _ptr(f), _ptr(e),
                      _stream(torch, psi.device)))
        return float(e[0]), f
