"""Synthetic QUMASUN systems: grid, V_loc, KB projector spheres, k-points, initial bands.

The reference builds these inputs from UPF pseudopotentials (SvF-interpolated
projector tables, PseudoPotNonlocal_SvF2_kpoint.h:911-1178, 2213-2516) and its
own V_local (PseudoPotLocal_SvF2.h).  The H.psi operator only consumes the
resulting arrays, so this module produces arrays of the same meaning from
closed forms:

* V_loc(r) = sum_a sum_images -14 exp(-|r - R_a|^2 / (2 * 0.8^2))
* KB projector p_{lm}(r) = r^l exp(-r^2 / (2 * 0.7^2)) Y_lm(r_hat) for |r| < r_c,
  with real Y_lm in the reference convention (RealSphericalHarmonics.h:148-162:
  no Condon-Shortley sign; order m = -l..l, SetBlockYlm.h:666-740), written here
  as solid-harmonic polynomials S_lm(x, y, z) = r^l Y_lm, for l <= 3.
* With nradial = 2 each l has a second radial projector
  r^l (r^2 / 0.7^2) exp(-r^2 / (2 * 0.7^2)) Y_lm; projectors are ordered by l,
  then radial index, then m, the reference's order of radial projectors k and
  m = -l_k..l_k (PseudoPotNonlocal_SvF2_kpoint.h:940-945, 960-966).  l <= 3 with
  two radial projectors gives 32 per atom, the shape of the reference's W.upf.
* KB energies D_l = (+4, -2, +1, -0.5) Ha for l = 0..3 (first radial projector)
  and (-1.5, +0.75, -0.4, +0.25) Ha for the second.
* Projector support: the voxels of the atom's cutoff sphere, stored with
  UNWRAPPED integer coordinates (the Bloch phase uses them,
  PseudoPotNonlocal_SvF2_kpoint.h:3578-3596) and ordered z, y, x so runs along x
  are contiguous in memory.
* Tables are column-major V_a x M_a per atom (..._kpoint.h:2255-2272); the
  gradient tables d p / d r hold the same layout per direction
  (FORCE_DIFF_PROJ layout [p | dp/dx | dp/dy | dp/dz], ..._kpoint.h:2255-2267).
* Grid spacing h = L / N and k = 2 pi / L * (reduced k) (qumasun_base1.constructor.h:28-34).

All randomness comes from numpy.random.default_rng(seed) so the oracle and the
GPU path consume identical host arrays.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

VLOC_DEPTH = -14.0
VLOC_SIGMA = 0.8
KB_SIGMA = 0.7
KB_ENERGY = (4.0, -2.0, 1.0, -0.5)
KB_ENERGY2 = (-1.5, 0.75, -0.4, 0.25)  # second radial projector of each l
_VLOC_RCUT = 10.0 * VLOC_SIGMA  # exp(-50) ~ 2e-22 of the well depth is dropped


@dataclasses.dataclass
class System:
    name: str
    nx: int
    ny: int
    nz: int
    cell: tuple            # (Lx, Ly, Lz) bohr, orthorhombic
    positions: np.ndarray  # (natoms, 3) bohr
    lmax: np.ndarray       # (natoms,) int32
    rc: np.ndarray         # (natoms,) bohr
    nbands: int
    ncomp: int             # 1 = real Gamma-point bands, 2 = SoA complex bands
    kpoints: np.ndarray    # (nk, 3) 1/bohr
    seed: int
    nradial: int = 1             # radial KB projectors per l (1 or 2)
    vloc: np.ndarray = None
    nproj: np.ndarray = None     # (natoms,) int32
    proj_off: np.ndarray = None  # (natoms+1,) int32
    vox_off: np.ndarray = None   # (natoms+1,) int64
    vox_xyz: np.ndarray = None   # (Vtot, 3) int32 unwrapped grid coordinates
    tab_off: np.ndarray = None   # (natoms+1,) int64
    table: np.ndarray = None     # (Ttot,) float64
    dtable: np.ndarray = None    # (3, Ttot) float64, d p / d r_{x,y,z}
    energy: np.ndarray = None    # (Ptot,) float64

    @property
    def h(self):
        return (self.cell[0] / self.nx, self.cell[1] / self.ny, self.cell[2] / self.nz)

    @property
    def ng(self):
        return self.nx * self.ny * self.nz

    @property
    def dV(self):
        hx, hy, hz = self.h
        return hx * hy * hz

    @property
    def natoms(self):
        return int(self.positions.shape[0])

    def build(self):
        """Fill V_loc and the projector set from positions / lmax / rc."""
        self.vloc = gaussian_vloc(self)
        build_projectors(self)
        return self

    def initial_bands(self, nbands=None, seed=None, dtype=np.float64):
        """uniform[-0.5, 0.5] bands (real and imaginary parts for ncomp=2), then
        orthonormalised so that sum |psi|^2 dV = 1 (Gram-Schmidt via QR).
        Returned in the SoA band layout, shape (nbands*ncomp, Ng)."""
        nb = self.nbands if nbands is None else nbands
        rng = np.random.default_rng(self.seed if seed is None else seed)
        ng = self.ng
        if self.ncomp == 1:
            a = rng.uniform(-0.5, 0.5, size=(nb, ng))
            q, _ = np.linalg.qr(a.T)
            psi = q.T / math.sqrt(self.dV)
        else:
            re = rng.uniform(-0.5, 0.5, size=(nb, ng))
            im = rng.uniform(-0.5, 0.5, size=(nb, ng))
            q, _ = np.linalg.qr((re + 1j * im).T)
            z = q.T / math.sqrt(self.dV)
            psi = np.empty((2 * nb, ng))
            psi[0::2] = z.real
            psi[1::2] = z.imag
        return np.ascontiguousarray(psi, dtype=dtype)


# --------------------------------------------------------------------------
# V_loc

def gaussian_vloc(s: System) -> np.ndarray:
    nx, ny, nz = s.nx, s.ny, s.nz
    hx, hy, hz = s.h
    v = np.zeros(s.ng)
    for R in s.positions:
        rng_ = []
        for d, (hd, nd) in enumerate(((hx, nx), (hy, ny), (hz, nz))):
            lo = math.ceil((R[d] - _VLOC_RCUT) / hd)
            hi = math.floor((R[d] + _VLOC_RCUT) / hd)
            rng_.append(np.arange(lo, hi + 1))
        ix, iy, iz = rng_
        dx2 = (ix * hx - R[0]) ** 2
        dy2 = (iy * hy - R[1]) ** 2
        dz2 = (iz * hz - R[2]) ** 2
        r2 = dz2[:, None, None] + dy2[None, :, None] + dx2[None, None, :]
        val = VLOC_DEPTH * np.exp(-r2 / (2.0 * VLOC_SIGMA ** 2))
        val[r2 > _VLOC_RCUT ** 2] = 0.0
        g = ((iz % nz)[:, None, None] * ny + (iy % ny)[None, :, None]) * nx + (ix % nx)[None, None, :]
        v += np.bincount(g.ravel(), weights=val.ravel(), minlength=s.ng)
    return v


# --------------------------------------------------------------------------
# KB projectors

_C0 = 1.0 / math.sqrt(4.0 * math.pi)
_C1 = math.sqrt(3.0 / (4.0 * math.pi))
_C2a = math.sqrt(15.0 / (4.0 * math.pi))   # xy, yz, xz
_C2b = math.sqrt(5.0 / (16.0 * math.pi))   # 2z^2 - x^2 - y^2
_C2c = math.sqrt(15.0 / (16.0 * math.pi))  # x^2 - y^2
_C3a = 0.25 * math.sqrt(35.0 / (2.0 * math.pi))  # 3x^2 y - y^3, x^3 - 3 x y^2
_C3b = 0.5 * math.sqrt(105.0 / math.pi)          # xyz
_C3c = 0.25 * math.sqrt(21.0 / (2.0 * math.pi))  # y (4z^2 - x^2 - y^2), x (4z^2 - x^2 - y^2)
_C3d = 0.25 * math.sqrt(7.0 / math.pi)           # z (2z^2 - 3x^2 - 3y^2)
_C3e = 0.25 * math.sqrt(105.0 / math.pi)         # z (x^2 - y^2)


def solid_harmonics(lmax, x, y, z):
    """S_lm = r^l Y_lm and its gradient, for l <= lmax, m = -l..l.
    Returns (S[w], dS[w] = (dx, dy, dz))."""
    one = np.ones_like(x)
    zero = np.zeros_like(x)
    S = [_C0 * one]
    dS = [(zero, zero, zero)]
    if lmax >= 1:
        S += [_C1 * y, _C1 * z, _C1 * x]
        dS += [(zero, _C1 * one, zero), (zero, zero, _C1 * one), (_C1 * one, zero, zero)]
    if lmax >= 2:
        S += [_C2a * x * y, _C2a * y * z, _C2b * (2 * z * z - x * x - y * y), _C2a * x * z, _C2c * (x * x - y * y)]
        dS += [(_C2a * y, _C2a * x, zero),
               (zero, _C2a * z, _C2a * y),
               (-2 * _C2b * x, -2 * _C2b * y, 4 * _C2b * z),
               (_C2a * z, zero, _C2a * x),
               (2 * _C2c * x, -2 * _C2c * y, zero)]
    if lmax >= 3:
        x2, y2, z2 = x * x, y * y, z * z
        S += [_C3a * (3 * x2 * y - y2 * y),
              _C3b * x * y * z,
              _C3c * y * (4 * z2 - x2 - y2),
              _C3d * z * (2 * z2 - 3 * x2 - 3 * y2),
              _C3c * x * (4 * z2 - x2 - y2),
              _C3e * z * (x2 - y2),
              _C3a * (x2 * x - 3 * x * y2)]
        dS += [(_C3a * 6 * x * y, _C3a * (3 * x2 - 3 * y2), zero),
               (_C3b * y * z, _C3b * x * z, _C3b * x * y),
               (-2 * _C3c * x * y, _C3c * (4 * z2 - x2 - 3 * y2), 8 * _C3c * y * z),
               (-6 * _C3d * x * z, -6 * _C3d * y * z, _C3d * (6 * z2 - 3 * x2 - 3 * y2)),
               (_C3c * (4 * z2 - 3 * x2 - y2), -2 * _C3c * x * y, 8 * _C3c * x * z),
               (2 * _C3e * x * z, -2 * _C3e * y * z, _C3e * (x2 - y2)),
               (_C3a * (3 * x2 - 3 * y2), -6 * _C3a * x * y, zero)]
    if lmax > 3:
        raise ValueError("synthetic KB projectors are defined for l <= 3")
    return S, dS


def sphere_voxels(R, rc, h):
    """Unwrapped integer coordinates (V, 3) of the voxels with |r_v - R| < rc,
    ordered with x fastest, then y, then z."""
    hx, hy, hz = h
    ax = [np.arange(math.ceil((R[d] - rc) / hd), math.floor((R[d] + rc) / hd) + 1)
          for d, hd in enumerate((hx, hy, hz))]
    ix, iy, iz = ax
    r2 = ((iz * hz - R[2]) ** 2)[:, None, None] + ((iy * hy - R[1]) ** 2)[None, :, None] \
        + ((ix * hx - R[0]) ** 2)[None, None, :]
    kz, ky, kx = np.nonzero(r2 < rc * rc)
    return np.stack([ix[kx], iy[ky], iz[kz]], axis=1).astype(np.int32)


def projector_values(R, lmax, vox, h, nradial=1):
    """Table (M, V) and gradient tables (3, M, V) of the KB projectors at voxels,
    M = nradial (lmax + 1)^2, ordered l, radial index, m."""
    hx, hy, hz = h
    x = vox[:, 0] * hx - R[0]
    y = vox[:, 1] * hy - R[1]
    z = vox[:, 2] * hz - R[2]
    r2 = x * x + y * y + z * z
    inv_s2 = 1.0 / KB_SIGMA ** 2
    g = np.exp(-r2 / (2.0 * KB_SIGMA ** 2))
    S, dS = solid_harmonics(lmax, x, y, z)
    xyz = (x, y, z)
    tabs, dtabs = [], []
    for l in range(lmax + 1):
        for k in range(nradial):
            # radial factor u g with u = 1 or r^2 / sigma^2; grad(S u g) = (dS u + S du - S u r / sigma^2) g
            u = np.ones_like(r2) if k == 0 else r2 * inv_s2
            for w in range(l * l, (l + 1) * (l + 1)):
                s_, d_ = S[w], dS[w]
                tabs.append(s_ * u * g)
                dtabs.append(np.stack([(d_[c] * u + (0.0 if k == 0 else s_ * 2.0 * xyz[c] * inv_s2)
                                        - s_ * u * xyz[c] * inv_s2) * g for c in range(3)]))
    tab = np.stack(tabs)
    dtab = np.stack(dtabs, axis=1)  # (3, M, V)
    return tab, dtab


def build_projectors(s: System):
    nat = s.natoms
    h = s.h
    nr = int(s.nradial)
    nproj = np.array([nr * (l + 1) ** 2 for l in s.lmax], dtype=np.int32)
    vox_list, tabs, dtabs, energies = [], [], [], []
    vox_off = np.zeros(nat + 1, dtype=np.int64)
    tab_off = np.zeros(nat + 1, dtype=np.int64)
    proj_off = np.zeros(nat + 1, dtype=np.int32)
    for a in range(nat):
        vox = sphere_voxels(s.positions[a], float(s.rc[a]), h)
        tab, dtab = projector_values(s.positions[a], int(s.lmax[a]), vox, h, nr)
        vox_list.append(vox)
        tabs.append(tab.ravel())          # (M, V) row-major == column-major V x M
        dtabs.append(dtab.reshape(3, -1))
        energies.append(np.concatenate([[(KB_ENERGY, KB_ENERGY2)[k][l]] * (2 * l + 1)
                                        for l in range(int(s.lmax[a]) + 1) for k in range(nr)]))
        vox_off[a + 1] = vox_off[a] + vox.shape[0]
        tab_off[a + 1] = tab_off[a] + tab.size
        proj_off[a + 1] = proj_off[a] + nproj[a]
    s.nproj = nproj
    s.vox_off = vox_off
    s.tab_off = tab_off
    s.proj_off = proj_off
    s.vox_xyz = np.ascontiguousarray(np.concatenate(vox_list) if nat else np.zeros((0, 3), np.int32), dtype=np.int32)
    s.table = np.ascontiguousarray(np.concatenate(tabs) if nat else np.zeros(0))
    s.dtable = np.ascontiguousarray(np.concatenate(dtabs, axis=1) if nat else np.zeros((3, 0)))
    s.energy = np.ascontiguousarray(np.concatenate(energies) if nat else np.zeros(0))
    return s


def retabulate(s: System, positions) -> System:
    """Copy of `s` with the projector (and gradient) tables evaluated at new atom
    positions on the SAME voxel sets (sphere membership frozen), as a central
    finite difference of the projector position derivative requires."""
    t = dataclasses.replace(s, positions=np.array(positions, dtype=np.float64))
    h = s.h
    tabs, dtabs = [], []
    for a in range(s.natoms):
        vox = s.vox_xyz[s.vox_off[a]:s.vox_off[a + 1]]
        tab, dtab = projector_values(t.positions[a], int(s.lmax[a]), vox, h, int(s.nradial))
        tabs.append(tab.ravel())
        dtabs.append(dtab.reshape(3, -1))
    t.table = np.ascontiguousarray(np.concatenate(tabs)) if s.natoms else s.table
    t.dtable = np.ascontiguousarray(np.concatenate(dtabs, axis=1)) if s.natoms else s.dtable
    return t


# --------------------------------------------------------------------------
# The five benchmark systems

def _bcc(ncell, a):
    base = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
    pts = [(np.array([i, j, k]) + b) * a for i in range(ncell) for j in range(ncell)
           for k in range(ncell) for b in base]
    return np.array(pts)


def _random_positions(rng, natoms, L, min_sep):
    pos = []
    while len(pos) < natoms:
        p = rng.uniform(0.0, L, size=3)
        ok = True
        for q in pos:
            d = p - q
            d -= L * np.round(d / L)
            if np.dot(d, d) < min_sep * min_sep:
                ok = False
                break
        if ok:
            pos.append(p)
    return np.array(pos)


def config(index: int, *, nbands=None, grid=None, build=True) -> System:
    """The synthetic workloads (0..4).  `nbands` / `grid` override the band count
    or grid size for reduced parity cases; the physics is unchanged."""
    if index == 0:
        L, n, nb, seed = 6.0, 32, 24, 0
        pos = np.array([[0.0, 0.0, 0.0], [3.0, 3.0, 3.0]])
        lmax = np.full(2, 2, np.int32)
        rc = np.full(2, 2.5)
        kp = np.zeros((1, 3))
        s = System("C0 bcc2 32^3 real FP64", n, n, n, (L, L, L), pos, lmax, rc, nb, 1, kp, seed)
    elif index == 1:
        L, n, nb, seed = 12.0, 64, 128, 1
        rng = np.random.default_rng(seed)
        pos = _bcc(2, 6.0) + rng.normal(0.0, 0.05, size=(16, 3))
        s = System("C1 bcc16 64^3 real mixed", n, n, n, (L, L, L), pos, np.full(16, 2, np.int32),
                   np.full(16, 2.5), nb, 1, np.zeros((1, 3)), seed)
    elif index == 2:
        L, n, nb, seed = 18.0, 96, 256, 2
        pos = _bcc(3, 6.0)
        red = np.array([[0, 0, 0], [0.25, 0, 0], [0.25, 0.25, 0], [0.25, 0.25, 0.25]])
        kp = red * (2.0 * math.pi / L)
        s = System("C2 bcc54 96^3 complex 4k", n, n, n, (L, L, L), pos, np.full(54, 2, np.int32),
                   np.full(54, 2.5), nb, 2, kp, seed)
    elif index == 3:
        L, n, nb, seed = 24.0, 160, 512, 3
        rng = np.random.default_rng(seed)
        pos = _random_positions(rng, 128, L, 2.5)
        rc = rng.uniform(1.5, 3.0, size=128)
        lmax = rng.integers(1, 3, size=128).astype(np.int32)
        s = System("C3 rand128 160^3 real uneven", n, n, n, (L, L, L), pos, lmax, rc, nb, 1,
                   np.zeros((1, 3)), seed)
    elif index == 4:
        L, n, nb, seed = 10.0, 64, 64, 4
        rng = np.random.default_rng(seed)
        pos = _random_positions(rng, 8, L, 2.5)
        s = System("C4 rand8 64^3 adjoint", n, n, n, (L, L, L), pos, np.full(8, 2, np.int32),
                   np.full(8, 2.5), nb, 1, np.zeros((1, 3)), seed)
    else:
        raise ValueError(f"unknown config {index}")
    if nbands is not None:
        s.nbands = nbands
    if grid is not None:
        # same cell, different grid: used for reduced-size parity cases
        s.nx = s.ny = s.nz = int(grid)
    return s.build() if build else s


def small_system(n=12, L=4.0, natoms=2, lmax=2, rc=1.4, nbands=4, ncomp=1, k_red=(0.0, 0.0, 0.0), seed=7,
                 nradial=1):
    """Tiny periodic system for brute-force checks (dense matrices on n^3 grids)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, L, size=(natoms, 3))
    kp = np.array([k_red]) * (2.0 * math.pi / L)
    s = System(f"small{n}", n, n, n, (L, L, L), pos, np.full(natoms, lmax, np.int32),
               np.full(natoms, rc), nbands, ncomp, kp, seed, nradial)
    return s.build()


def wlike(index: int = 0, *, nbands=None, grid=None) -> System:
    """One of the synthetic workloads with the projector shape of the reference's
    W.upf: two radial projectors for each l <= 3 (32 per atom) on every atom."""
    s = config(index, nbands=nbands, grid=grid, build=False)
    s.name += " W-like 32 projectors"
    s.lmax = np.full(s.natoms, 3, np.int32)
    s.nradial = 2
    return s.build()
