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
        L.oracle_projector_table.argtypes = [I, P, P, P, D, D, D, D, P, P, P, P]
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


def kb_projector(l, m, sigma, rc, x, y, z):
    return lib().oracle_kb_projector(l, m, sigma, rc, x, y, z)


def projector_table(s, positions=None):
    """The oracle's own evaluation of the KB tables at the system's voxels."""
    pos = _c(s.positions if positions is None else positions, np.float64)
    lmax = _c(s.lmax, np.int32)
    rc = _c(s.rc, np.float64)
    vo, vx, to = _c(s.vox_off, np.int64), _c(s.vox_xyz, np.int32), _c(s.tab_off, np.int64)
    out = np.zeros(int(s.tab_off[-1]))
    sigma = 0.7  # KB radial width of the synthetic projectors
    hx, hy, hz = s.h
    lib().oracle_projector_table(s.natoms, _p(pos), _p(lmax), _p(rc), sigma, hx, hy, hz,
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
