"""UPF pseudopotentials: the host-side loader and projector-table builder of
libqumasun_hip.so (qm_upf_*), and the step that feeds their tables to a
System (and from there to qm_set_projectors).

The reference reads UPF files in LoadPseudoPotential_UPF_ONCV
(upf_oncv_loader.h:355-530) and builds its real-space projector tables in
PseudoPotNonlocal_SvF2_kpoint.h (911-1178, 1912-2516).  The library restates
the loader and the radial resampling; the tables are the point values
f_k(|r|) Y_lm(r/|r|) on the grid, without the reference's Fourier-space
filtering step.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Pseudopotential:
    """One species' nonlocal pseudopotential from a UPF v2 file."""

    def __init__(self, path):
        self.lib = _lib.load()
        self.handle = ctypes.c_void_p()
        _lib.check(self.lib.qm_upf_load(ctypes.byref(self.handle), str(path).encode()))
        nr, npj = ctypes.c_int(), ctypes.c_int()
        zv, rc = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.qm_upf_info(self.handle, ctypes.byref(nr), ctypes.byref(npj), ctypes.byref(zv),
                                        ctypes.byref(rc)))
        self.nradial, self.nproj = nr.value, npj.value
        self.z_valence, self.cutoff = zv.value, rc.value
        self.path = str(path)

    def close(self):
        if self.handle:
            self.lib.qm_upf_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def radial(self, k):
        """(l, cutoff radius, KB energy in Ha) of radial projector k"""
        l, rc, e = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.qm_upf_radial(self.handle, int(k), ctypes.byref(l), ctypes.byref(rc), ctypes.byref(e)))
        return l.value, rc.value, e.value

    def eval(self, k, r):
        """f_k(r) and df_k/dr at radii r (bohr)"""
        r = np.ascontiguousarray(r, dtype=np.float64).ravel()
        f = np.empty_like(r)
        df = np.empty_like(r)
        _lib.check(self.lib.qm_upf_eval(self.handle, int(k), r.size, _ptr(r), _ptr(f), _ptr(df)))
        return f, df

    def tables(self, positions, shape, h, gradients=True):
        """Projector tables for atoms of this species at `positions` (natoms x 3,
        bohr) on a grid of `shape` points with spacing `h`, in the
        qm_set_projectors layout (dict of arrays)."""
        pos = np.ascontiguousarray(np.atleast_2d(positions), dtype=np.float64)
        nat = pos.shape[0]
        nx, ny, nz = (int(v) for v in shape)
        hx, hy, hz = (float(v) for v in h)
        counts = np.zeros(3, dtype=np.int64)
        null = ctypes.c_void_p()
        _lib.check(self.lib.qm_upf_tables(self.handle, nat, _ptr(pos), nx, ny, nz, hx, hy, hz, _ptr(counts),
                                          null, null, null, null, null, null, null, null))
        vtot, ttot, ptot = (int(c) for c in counts)
        out = {
            "nproj": np.zeros(nat, np.int32), "vox_off": np.zeros(nat + 1, np.int64),
            "vox_xyz": np.zeros((vtot, 3), np.int32), "table": np.zeros(ttot),
            "dtable": np.zeros((3, ttot)) if gradients else None, "tab_off": np.zeros(nat + 1, np.int64),
            "energy": np.zeros(ptot), "proj_off": np.zeros(nat + 1, np.int32),
        }
        _lib.check(self.lib.qm_upf_tables(
            self.handle, nat, _ptr(pos), nx, ny, nz, hx, hy, hz, _ptr(counts), _ptr(out["nproj"]),
            _ptr(out["vox_off"]), _ptr(out["vox_xyz"]), _ptr(out["table"]),
            _ptr(out["dtable"]) if gradients else null, _ptr(out["tab_off"]), _ptr(out["energy"]),
            _ptr(out["proj_off"])))
        return out


def attach(system, species):
    """Replace `system`'s projector set with UPF tables.  `species` is one
    Pseudopotential for every atom, or a list with one per atom.  Returns the
    system (changed in place); Hamiltonian(system) then uploads the tables."""
    nat = system.natoms
    if isinstance(species, Pseudopotential):
        species = [species] * nat
    if len(species) != nat:
        raise ValueError(f"{len(species)} species for {nat} atoms")
    parts = [species[a].tables(system.positions[a], (system.nx, system.ny, system.nz), system.h)
             for a in range(nat)]
    cat = np.concatenate
    system.nproj = np.array([p["nproj"][0] for p in parts], dtype=np.int32)
    counts = [(p["vox_off"][1], p["tab_off"][1], p["proj_off"][1]) for p in parts]
    system.vox_off = np.concatenate([[0], np.cumsum([c[0] for c in counts])]).astype(np.int64)
    system.tab_off = np.concatenate([[0], np.cumsum([c[1] for c in counts])]).astype(np.int64)
    system.proj_off = np.concatenate([[0], np.cumsum([c[2] for c in counts])]).astype(np.int32)
    system.vox_xyz = np.ascontiguousarray(cat([p["vox_xyz"] for p in parts]) if nat else np.zeros((0, 3)),
                                          dtype=np.int32)
    system.table = np.ascontiguousarray(cat([p["table"] for p in parts]) if nat else np.zeros(0))
    system.dtable = np.ascontiguousarray(cat([p["dtable"] for p in parts], axis=1) if nat else np.zeros((3, 0)))
    system.energy = np.ascontiguousarray(cat([p["energy"] for p in parts]) if nat else np.zeros(0))
    system.lmax = np.array([max(s.radial(k)[0] for k in range(s.nradial)) if s.nradial else 0 for s in species],
                           dtype=np.int32)
    system.rc = np.array([s.cutoff for s in species])
    return system
