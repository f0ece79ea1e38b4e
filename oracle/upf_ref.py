"""CPU restatement of the UPF projector chain (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
UPF loader and table builder (qm_upf_*).  The product library never loads it.

What it restates, step by step:
- parse(): the nonlocal part of LoadPseudoPotential_UPF_ONCV
  (upf_oncv_loader.h:355-447): number_of_proj and z_valence (PP_HEADER), the
  radial mesh PP_R, PP_BETA.k with angular_momentum and cutoff_radius, and
  E_k = D_kk / 2 (Ry -> Ha) from PP_DIJ.  Parsed with xml.etree, independently
  of the library's own text parser.
- beta_over_r(): upf_oncv_loader.h:497-502 (divide by r; the r = 0 value is the
  next point's for l = 0, else 0).
- uniform_value(): UniformGrid::GetValue (UniformGrid.h:40-87).
- log_grid(): the 1000-point resampling grid r_i = exp(-9 + i dxi)
  (upf_oncv_loader.h:384-390, 458-461) and the projector values on it
  (upf_oncv_loader.h:504-509).
- xi_value(): RadialGrid2::GetValueByXi (RadialGrid2.h:81-157).
- radial(): the point rule of mInitializeNonlocalProjector
  (PseudoPotNonlocal_SvF2_kpoint.h:1003-1024): the first log-grid value below
  r_0, 0 beyond the largest cutoff radius of that l, else GetValueBySquare.
- tables(): p_{k l m}(r_v - R) = f_k(|r|) Y_lm(r / |r|) over the voxels within
  the largest cutoff, in the order k, m = -l..l (…_kpoint.h:940-966), Y_lm from
  the oracle's harmonics (RealSphericalHarmonics.h:148-162).

Pins (tests/test_oracle_upf.py): polynomial exactness of both interpolants
(degree <= 4 for the quintics, <= 2 for the cubic cells, <= 1 for the linear
cell), the mirror parity through r = 0, convergence of the chain to a smooth
closed-form beta, the gradient tables against central differences of the
table function, and the W pseudopotential's header values and projector count
(32 for l = 0..3 with two projectors each) read from the file itself.
The Fourier-space filtering the reference applies to the tables afterwards
(…_kpoint.h:911-1178, 1912-2208) is not restated: parity unpinned for that
step.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

LOG_POINTS = 1000
XI_MIN = -9.0


def parse(path):
    root = ET.parse(path).getroot()
    hdr = root.find("PP_HEADER").attrib
    n = int(hdr["number_of_proj"])
    r = np.array(root.find("PP_MESH/PP_R").text.split(), dtype=np.float64)
    out = {"nradial": n, "z_valence": float(hdr.get("z_valence", "0")), "r": r, "l": [], "cutoff": [],
           "beta": [], "energy": []}
    for k in range(n):
        e = root.find(f"PP_NONLOCAL/PP_BETA.{k + 1}")
        out["l"].append(int(e.attrib["angular_momentum"]))
        out["cutoff"].append(float(e.attrib.get("cutoff_radius", "0")))
        out["beta"].append(np.array(e.text.split()[: r.size], dtype=np.float64))
    if n:
        d = np.array(root.find("PP_NONLOCAL/PP_DIJ").text.split()[: n * n], dtype=np.float64).reshape(n, n)
        out["energy"] = [d[k, k] / 2.0 for k in range(n)]
    return out


def beta_over_r(r, beta, l):
    f = np.empty_like(beta)
    f[1:] = beta[1:] / r[1:]
    f[0] = f[1] if l == 0 else 0.0
    return f


def _derivs(fm2, fm1, f0, f1, f2, f3):
    d1f0 = 2.0 / 3.0 * (f1 - fm1) - 1.0 / 12.0 * (f2 - fm2)
    d1f1 = 2.0 / 3.0 * (f2 - f0) - 1.0 / 12.0 * (f3 - fm1)
    d2f0 = 4.0 / 3.0 * (f1 + fm1) - 1.0 / 12.0 * (f2 + fm2) - 5.0 / 2.0 * f0
    d2f1 = 4.0 / 3.0 * (f2 + f0) - 1.0 / 12.0 * (f3 + fm1) - 5.0 / 2.0 * f1
    return d1f0, d1f1, d2f0, d2f1


def _quintic(x, f0, d1f0, d2f0, f1, d1f1, d2f1):
    """value and d/dx of the quintic with (value, d/dx, d2/dx2) given at x = 0 and 1"""
    # solve for c3..c5 from the three end conditions at x = 1
    e0 = f1 - (f0 + d1f0 + d2f0 / 2.0)
    e1 = d1f1 - (d1f0 + d2f0)
    e2 = d2f1 - d2f0
    A = np.array([[1.0, 1.0, 1.0], [3.0, 4.0, 5.0], [6.0, 12.0, 20.0]])
    c = np.linalg.solve(A, np.stack([e0, e1, e2]).reshape(3, -1)).reshape((3,) + np.shape(e0))
    c3, c4, c5 = c
    v = f0 + d1f0 * x + d2f0 / 2.0 * x**2 + c3 * x**3 + c4 * x**4 + c5 * x**5
    dv = d1f0 + d2f0 * x + 3 * c3 * x**2 + 4 * c4 * x**3 + 5 * c5 * x**4
    return v, dv


def uniform_value(vals, r0, dr, parity, r):
    """UniformGrid::GetValue: the mesh mirrored through its first point with
    the given parity, zero past its end."""
    vals = np.asarray(vals, dtype=np.float64)
    n = vals.size
    rel = (np.asarray(r, dtype=np.float64) - r0) / dr
    i = np.floor(rel).astype(np.int64)
    x = rel - i

    def get(j):
        out = np.where(j < 0, vals[np.clip(-j, 0, n - 1)] * parity, 0.0)
        inside = (j >= 0) & (j < n)
        return np.where(inside, vals[np.clip(j, 0, n - 1)], out)

    f = [get(i + d) for d in (-2, -1, 0, 1, 2, 3)]
    d1f0, d1f1, d2f0, d2f1 = _derivs(*f)
    v, _ = _quintic(x, f[2], d1f0, d2f0, f[3], d1f1, d2f1)
    return v


def log_grid(r_last):
    dxi = (np.log(r_last) - XI_MIN) / LOG_POINTS
    return dxi, np.exp(np.arange(LOG_POINTS) * dxi + XI_MIN)


def xi_value(v, xi_min, dxi, xi):
    """RadialGrid2::GetValueByXi: value and d/dxi."""
    v = np.asarray(v, dtype=np.float64)
    n = v.size
    xi = np.atleast_1d(np.asarray(xi, dtype=np.float64))
    rel = (xi - xi_min) / dxi
    i = np.floor(rel).astype(np.int64)
    x = rel - i
    val = np.zeros_like(xi)
    dval = np.zeros_like(xi)

    def at(j):
        return v[np.clip(j, 0, n - 1)]

    lin = i == 0
    val[lin] = v[0] * (1 - x[lin]) + v[1] * x[lin]
    dval[lin] = v[1] - v[0]
    q = (i > 1) & (i < n - 3)
    if q.any():
        iq, xq = i[q], x[q]
        d1f0, d1f1, d2f0, d2f1 = _derivs(*(at(iq + d) for d in (-2, -1, 0, 1, 2, 3)))
        val[q], dval[q] = _quintic(xq, at(iq), d1f0, d2f0, at(iq + 1), d1f1, d2f1)
    c = (i == 1) | (i == n - 3)
    if c.any():
        # cubic Hermite with central-difference slopes
        ic, xc = i[c], x[c]
        v0, v1 = at(ic), at(ic + 1)
        s0 = (v1 - at(ic - 1)) / 2.0
        s1 = (at(ic + 2) - v0) / 2.0
        h00, h10 = 2 * xc**3 - 3 * xc**2 + 1, xc**3 - 2 * xc**2 + xc
        h01, h11 = -2 * xc**3 + 3 * xc**2, xc**3 - xc**2
        val[c] = h00 * v0 + h10 * s0 + h01 * v1 + h11 * s1
        dval[c] = ((6 * xc**2 - 6 * xc) * v0 + (3 * xc**2 - 4 * xc + 1) * s0 + (-6 * xc**2 + 6 * xc) * v1
                   + (3 * xc**2 - 2 * xc) * s1)
    # i >= n - 2 (and i < 0) stay 0
    return val, dval / dxi


class Radial:
    """The two-stage radial functions of one pseudopotential."""

    def __init__(self, upf):
        r = upf["r"]
        self.l = list(upf["l"])
        self.energy = list(upf["energy"])
        self.cutoff_l = np.zeros(4)
        for l, rc in zip(upf["l"], upf["cutoff"]):
            self.cutoff_l[l] = max(self.cutoff_l[l], rc)
        self.rc_max = max(upf["cutoff"]) if upf["cutoff"] else 0.0
        self.dxi, rl = log_grid(r[-1])
        r0, dr = r[0], r[1] - r[0]
        self.flog = [uniform_value(beta_over_r(r, b, l), r0, dr, -1.0 if l & 1 else 1.0, rl)
                     for b, l in zip(upf["beta"], upf["l"])]

    def value(self, k, r):
        """f_k(r) and df/dr at radii r"""
        r = np.atleast_1d(np.asarray(r, dtype=np.float64))
        f = np.zeros_like(r)
        df = np.zeros_like(r)
        small = r < np.exp(XI_MIN)
        f[small] = self.flog[k][0]
        mid = ~small & (r <= self.cutoff_l[self.l[k]])
        if mid.any():
            a, b = xi_value(self.flog[k], XI_MIN, self.dxi, np.log(r[mid]))
            f[mid], df[mid] = a, b / r[mid]
        return f, df


def sphere_voxels(R, rc, h):
    """unwrapped integer coordinates (ordered z, y, x) of the grid points
    strictly within rc of R"""
    lo = np.ceil((np.asarray(R) - rc) / h).astype(int)
    hi = np.floor((np.asarray(R) + rc) / h).astype(int)
    iz, iy, ix = np.meshgrid(*(np.arange(lo[d], hi[d] + 1) for d in (2, 1, 0)), indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    d = pts * h - R
    return pts[(d * d).sum(axis=1) < rc * rc]


def table_values(rad, ylm, d):
    """projector values at displacements d (V x 3): (V, M) in order k, m"""
    rr = np.sqrt((d * d).sum(axis=1))
    cols = []
    for k, l in enumerate(rad.l):
        f, _ = rad.value(k, rr)
        for m in range(-l, l + 1):
            cols.append(f * ylm(l, m, d[:, 0], d[:, 1], d[:, 2]))
    return np.stack(cols, axis=1) if cols else np.zeros((d.shape[0], 0))


def tables(rad, ylm, positions, h):
    """per atom: (voxels V x 3, table V x M) for atoms at positions (bohr)"""
    h = np.asarray(h, dtype=np.float64)
    out = []
    for R in np.asarray(positions, dtype=np.float64):
        vox = sphere_voxels(R, rad.rc_max, h)
        out.append((vox, table_values(rad, ylm, vox * h - R)))
    return out
