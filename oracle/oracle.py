"""ctypes front end of the CPU FP64 oracle (oracle/qm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by the cpu_baseline leg of bench.py.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libqm_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        I, D, P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.oracle_hpsi.argtypes = [I, I, I, D, D, D, P, I, I, P, P, D, D, D, I, P, P, P, P, P, P, P]
        L.oracle_kinetic_add.argtypes = [I, I, I, D, D, D, I, I, P, P, D, D, D]
        L.oracle_nonlocal_add.argtypes = [I, I, I, D, D, D, I, I, P, P, D, D, D, I, P, P, P, P, P, P, P]
        L.oracle_ylm.argtypes = [I, I, D, D, D]
        L.oracle_ylm.restype = D
        L.oracle_kb_projector.argtypes = [I, I, D, D, D, D, D]
        L.oracle_kb_projector.restype = D
        L.oracle_kb_projector_radial.argtypes = [I, I, I, D, D, D, D, D]
        L.oracle_kb_projector_radial.restype = D
        L.oracle_projector_table.argtypes = [I, P, P, P, I, D, D, D, D, P, P, P, P]
        L.oracle_band_density.argtypes = [ctypes.c_int64, I, I, P, P, P]
        L.oracle_nonlocal_forces.argtypes = [I, I, I, D, D, D, I, I, P, P, D, D, D, I, P, P, P, P, P,
                                             ctypes.c_int64, P, P, P, P]
        L.oracle_nonlocal_forces.restype = D
        L.oracle_num_threads.restype = I
        L.oracle_set_threads.argtypes = [I]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _proj_args(s):
    keep = [_c(s.nproj, np.int32), _c(s.vox_off, np.int64), _c(s.vox_xyz, np.int32),
            _c(s.table, np.float64), _c(s.tab_off, np.int64), _c(s.energy, np.float64),
            _c(s.proj_off, np.int32)]
    return keep, [_p(a) for a in keep]


def set_threads(n):
    lib().oracle_set_threads(int(n))


def num_threads():
    return lib().oracle_num_threads()


def hpsi(s, psi, k=(0.0, 0.0, 0.0), ncomp=None, vloc=None, nonlocal_=True):
    """H.psi for a band block psi of shape (nb*ncomp, Ng) (SoA layout)."""
    ncomp = s.ncomp if ncomp is None else ncomp
    psi = _c(psi, np.float64)
    nb = psi.shape[0] // ncomp
    out = np.empty_like(psi)
    v = _c(s.vloc if vloc is None else vloc, np.float64)
    keep, pa = _proj_args(s)
    hx, hy, hz = s.h
    lib().oracle_hpsi(s.nx, s.ny, s.nz, hx, hy, hz, _p(v), nb, ncomp, _p(psi), _p(out),
                      float(k[0]), float(k[1]), float(k[2]), s.natoms if nonlocal_ else 0, *pa)
    return out


def kinetic(s, psi, k=(0.0, 0.0, 0.0), ncomp=None):
    ncomp = s.ncomp if ncomp is None else ncomp
    psi = _c(psi, np.float64)
    nb = psi.shape[0] // ncomp
    out = np.zeros_like(psi)
    hx, hy, hz = s.h
    lib().oracle_kinetic_add(s.nx, s.ny, s.nz, hx, hy, hz, nb, ncomp, _p(psi), _p(out),
                             float(k[0]), float(k[1]), float(k[2]))
    return out


def nonlocal_(s, psi, k=(0.0, 0.0, 0.0), ncomp=None):
    ncomp = s.ncomp if ncomp is None else ncomp
    psi = _c(psi, np.float64)
    nb = psi.shape[0] // ncomp
    out = np.zeros_like(psi)
    keep, pa = _proj_args(s)
    hx, hy, hz = s.h
    lib().oracle_nonlocal_add(s.nx, s.ny, s.nz, hx, hy, hz, nb, ncomp, _p(psi), _p(out),
                              float(k[0]), float(k[1]), float(k[2]), s.natoms, *pa)
    return out


def ylm(l, m, x, y, z):
    return lib().oracle_ylm(l, m, x, y, z)


def kb_projector(l, m, sigma, rc, x, y, z, irad=0):
    return lib().oracle_kb_projector_radial(l, m, irad, sigma, rc, x, y, z)


def projector_table(s, positions=None):
    """The oracle's own evaluation of the KB tables at the system's voxels."""
    pos = _c(s.positions if positions is None else positions, np.float64)
    lmax = _c(s.lmax, np.int32)
    rc = _c(s.rc, np.float64)
    vo, vx, to = _c(s.vox_off, np.int64), _c(s.vox_xyz, np.int32), _c(s.tab_off, np.int64)
    out = np.zeros(int(s.tab_off[-1]))
    sigma = 0.7  # KB radial width of the synthetic projectors
    hx, hy, hz = s.h
    lib().oracle_projector_table(s.natoms, _p(pos), _p(lmax), _p(rc), int(getattr(s, "nradial", 1)), sigma,
                                 hx, hy, hz,
                                 _p(vo), _p(vx), _p(to), _p(out))
    return out


def band_density(s, psi, occ, ncomp=None):
    ncomp = s.ncomp if ncomp is None else ncomp
    psi = _c(psi, np.float64)
    nb = psi.shape[0] // ncomp
    occ = _c(occ, np.float64)
    rho = np.zeros(s.ng)
    lib().oracle_band_density(s.ng, nb, ncomp, _p(psi), _p(occ), _p(rho))
    return rho


def nonlocal_forces(s, psi, occ, k=(0.0, 0.0, 0.0), ncomp=None, dtable=None):
    """Returns (E_nl, forces (natoms, 3)) with F = -dE_nl/dR."""
    ncomp = s.ncomp if ncomp is None else ncomp
    psi = _c(psi, np.float64)
    nb = psi.shape[0] // ncomp
    occ = _c(occ, np.float64)
    keep, pa = _proj_args(s)
    dt = _c(s.dtable if dtable is None else dtable, np.float64)
    forces = np.zeros((s.natoms, 3))
    hx, hy, hz = s.h
    nproj, vox_off, vox_xyz, table, tab_off, energy, proj_off = pa
    e = lib().oracle_nonlocal_forces(s.nx, s.ny, s.nz, hx, hy, hz, nb, ncomp, _p(psi), _p(occ),
                                     float(k[0]), float(k[1]), float(k[2]), s.natoms,
                                     nproj, vox_off, vox_xyz, table, _p(dt), int(s.tab_off[-1]),
                                     tab_off, energy, proj_off, _p(forces))
    return e, forces


def td_taylor(s, psi, dt, order=4, op="full", k=(0.0, 0.0, 0.0), kinetic="fd8"):
    """Taylor propagator psi <- sum_{n=0..order} (-i dt Op)^n / n! psi of SoA
    complex bands (QUMASUN::TimeEvoH4_bundle, qumasun_td.time_evo_H4.h:70-160, at
    order 4; TimeEvoA4_bundle, qumasun_symplectic.time_evo_K4.h:13-96, with
    Op = the nonlocal term alone).  op: "full", "local" (kinetic + V_loc) or
    "nonlocal".  As in the reference, each term H^n psi is made from H^(n-1) psi
    and added to psi at once."""
    assert s.ncomp == 2
    x = np.array(psi, dtype=np.float64)
    z = x[0::2] + 1j * x[1::2]

    def apply(v):
        soa = np.empty((2 * v.shape[0], v.shape[1]))
        soa[0::2], soa[1::2] = v.real, v.imag
        if op == "full":
            out = hpsi(s, soa, k=k) if kinetic == "fd8" else hpsi_fft(s, soa, k=k)
        elif op == "local":
            out = hpsi(s, soa, k=k, nonlocal_=False) if kinetic == "fd8" else hpsi_fft(s, soa, k=k, nonlocal_=False)
        else:
            out = nonlocal_(s, soa, k=k)
        return out[0::2] + 1j * out[1::2]

    term = z.copy()
    coef = 1.0
    for n in range(1, order + 1):
        term = apply(term)
        coef *= dt / n
        z = z + ((-1j) ** n) * coef * term
    out = np.empty_like(x)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def td_expv(s, psi, dt):
    """psi <- e^{-i V_loc dt} psi pointwise (mProductExpToV with
    exp_v = cos(-dt V) + i sin(-dt V), qumasun_symplectic.evolution.h:80-100, 140-165)."""
    x = np.array(psi, dtype=np.float64)
    c, sn = np.cos(-dt * s.vloc), np.sin(-dt * s.vloc)
    re, im = x[0::2].copy(), x[1::2].copy()
    x[0::2] = c * re - sn * im
    x[1::2] = c * im + sn * re
    return x


def _kspace_e2(s, k=(0.0, 0.0, 0.0)):
    """|G + k|^2 / 2 on the (nz, ny, nx) FFT grid, G = 2 pi / L_d times the wrapped
    index (i > n/2 -> i - n; the Nyquist index n/2 stays +n/2), as
    OperateKineticInKspace_bundle (qumasun_base1.matrix.h:96-130)."""
    hx, hy, hz = s.h

    def axis(n, h, kk):
        i = np.arange(n)
        return 2.0 * np.pi / (h * n) * np.where(i * 2 > n, i - n, i) + kk

    qx, qy, qz = axis(s.nx, hx, k[0]), axis(s.ny, hy, k[1]), axis(s.nz, hz, k[2])
    return 0.5 * (qz[:, None, None] ** 2 + qy[None, :, None] ** 2 + qx[None, None, :] ** 2)


def fft_kinetic(s, psi, k=(0.0, 0.0, 0.0), ncomp=None):
    """Spectral kinetic term IFFT(|G + k|^2 / 2 FFT psi) of a SoA band block, FP64
    (numpy FFT; the reference's mHamiltonianMatrix_ddm_bundle FFT branch,
    qumasun_base1.matrix.h:171-228)."""
    ncomp = s.ncomp if ncomp is None else ncomp
    x = np.asarray(psi, dtype=np.float64)
    z = x if ncomp == 1 else x[0::2] + 1j * x[1::2]
    shp = (z.shape[0], s.nz, s.ny, s.nx)
    e2 = _kspace_e2(s, k)
    t = np.fft.ifftn(e2[None] * np.fft.fftn(z.reshape(shp), axes=(1, 2, 3)), axes=(1, 2, 3)).reshape(z.shape)
    if ncomp == 1:
        return np.ascontiguousarray(t.real)
    out = np.empty_like(x)
    out[0::2], out[1::2] = t.real, t.imag
    return out


def hpsi_fft(s, psi, k=(0.0, 0.0, 0.0), ncomp=None, nonlocal_=True):
    """H.psi with the spectral kinetic term: FFT kinetic + V_loc psi + nonlocal."""
    ncomp = s.ncomp if ncomp is None else ncomp
    x = np.asarray(psi, dtype=np.float64)
    out = fft_kinetic(s, x, k, ncomp) + s.vloc[None, :] * x
    if nonlocal_ and s.natoms > 0:
        out += nonlocal_add(s, x, k, ncomp)
    return out


def nonlocal_add(s, psi, k=(0.0, 0.0, 0.0), ncomp=None):
    return nonlocal_(s, psi, k=k, ncomp=ncomp)


def td_kinetic_fft(s, psi, dt, k=(0.0, 0.0, 0.0)):
    """psi <- IFFT(e^{-i dt |G + k|^2 / 2} FFT psi) of SoA complex bands
    (EvolveKineticInKspace / mEvolveK_FFT, qumasun_symplectic.kinetic_fft.h:14-60, 65-140)."""
    x = np.asarray(psi, dtype=np.float64)
    z = x[0::2] + 1j * x[1::2]
    shp = (z.shape[0], s.nz, s.ny, s.nx)
    ph = np.exp(-1j * dt * _kspace_e2(s, k))
    t = np.fft.ifftn(ph[None] * np.fft.fftn(z.reshape(shp), axes=(1, 2, 3)), axes=(1, 2, 3)).reshape(z.shape)
    out = np.empty_like(x)
    out[0::2], out[1::2] = t.real, t.imag
    return out
