"""CPU restatement of the SCF glue around the LOBPCG solve (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
SCF pieces (qm_xc_lda_f64, qm_hartree_f64, qm_occupancy, qm_mixer_*).  The
product library never loads it.

What it restates:
- xc_lda(): Calc_XC_LDA (Vxc.h:43-76) with calc_LSDA_ge1 / calc_LSDA_lt1
  (Vxc.h:17-29): Perdew-Zunger 1981 LDA, rs = (3 / 4 pi rho)^(1/3), exchange
  e_x = -(3/4 pi)(9 pi / 4)^(1/3) / rs, correlation for rs >= 1
  gamma / (1 + b1 sqrt(rs) + b2 rs) (gamma, b1, b2 = -0.1423, 1.0529, 0.3334),
  for rs < 1 A ln rs + B + C rs ln rs + D rs (A, B, C, D = 0.0311, -0.0480,
  0.0020, -0.0116); V = e + rho de/drho; rho < 1e-8 gives zeros.
- hartree(): SetPotentialByPoissonFFT (poisson_fft.h:16-120):
  V_H(G) = 4 pi rho(G) / |G|^2 on the folded frequencies, V_H(0) = 0.
- occupancy_zero(): mSetOccupancyZero (qumasun_base1.occupancy.h:35-67).
- occupancy_fermi(): mSetOccupancyTemperature (occupancy.h:76-175) with
  SetOccupancy (SetOccupancy.h:7-19): bisection on mu between the lowest and
  highest eigenvalue until the bracket is below 1e-9, f = c / (e^p + 1) with
  p = min((e - mu)/kT, 80); the occupations (and the returned mu) are those
  of the last evaluation.
- MixSimple / MixDIIS: MixingSimple::Predict (mixing_simple.h:45-75) and
  MixingDIIS::Initialize / Predict (mixing_DIIS.h:48-229): a ring of
  `history` (rho, R = rho_out - rho_in) pairs, the bordered matrix of R.R
  inverted by Gauss-Jordan without pivoting (InverseM, inverse_m.h:17-60), its
  bordering row as coefficients; simple mixing until the ring is full.  Predict returns sum |rho_out - rho_in| (before the dV factor).

- scf_run(): the loop of QUMASUN_KPOINT::mSCF (qumasun_kpoint.scfloop.h:96-222)
  with the oracle's H.psi and exact lowest eigenpairs (ARPACK through
  scipy) in place of LOBPCG: the fixed point (density, energy) does not
  depend on the eigensolver's path.

Pins (tests/test_oracle_scf.py): PZ81 against its closed forms (exchange
energy density, the continuity of e_c at rs = 1 the parameters were fitted
for, V = d(rho e)/drho by central differences); the Poisson solve against a
plane-wave density (V = 4 pi / G^2 rho exactly) and -laplacian(V) = 4 pi rho
by FFT differentiation; occupations summing to the electron count; DIIS on a
linear fixed-point map reaching the exact fixed point (numpy solve) and
simple mixing reproducing its formula; InverseM against numpy.linalg.inv.
"""
from __future__ import annotations

import numpy as np

PZ_GE1 = (-0.1423, 1.0529, 0.3334)
PZ_LT1 = (0.0311, -0.0480, 0.0020, -0.0116)
KB_HARTREE = 1.380649e-23 / 4.3597447222071e-18  # Hartree per kelvin


def xc_lda(rho):
    """(rho e_x, rho e_c, V_x, V_c) per point"""
    rho = np.asarray(rho, dtype=np.float64)
    out = [np.zeros_like(rho) for _ in range(4)]
    m = rho >= 1.0e-8
    r = rho[m]
    c2 = np.cbrt(3.0 / (4.0 * np.pi))
    c3 = -(3.0 / (4.0 * np.pi)) * np.cbrt(9.0 * np.pi / 4.0)
    rs = c2 / np.cbrt(r)
    ex = c3 / rs
    ec = np.empty_like(rs)
    nd = np.empty_like(rs)
    hi = rs >= 1.0
    g, b1, b2 = PZ_GE1
    sr = np.sqrt(rs[hi])
    ec[hi] = g / (1.0 + b1 * sr + b2 * rs[hi])
    nd[hi] = ec[hi] ** 2 * (b1 * 0.5 * sr + b2 * rs[hi]) / (3.0 * g)
    A, B, C, D = PZ_LT1
    lo = ~hi
    lr = np.log(rs[lo])
    ec[lo] = B + A * lr + D * rs[lo] + C * rs[lo] * lr
    nd[lo] = -(A + (D + C) * rs[lo] + C * rs[lo] * lr) / 3.0
    out[0][m] = r * ex
    out[1][m] = r * ec
    out[2][m] = ex + ex / 3.0
    out[3][m] = ec + nd
    return out


def _g2(shape, h):
    axes = []
    for n, hh in zip(shape, h):
        i = np.arange(n)
        axes.append((2.0 * np.pi / (hh * n)) * np.where(2 * i > n, i - n, i))
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    return gx * gx + gy * gy + gz * gz


def hartree(rho, shape, h):
    """rho, V_H as flat arrays in the grid order (x fastest)"""
    nx, ny, nz = shape
    r = np.asarray(rho, dtype=np.float64).reshape(nz, ny, nx).transpose(2, 1, 0)
    rk = np.fft.fftn(r)
    g2 = _g2(shape, h)
    g2[0, 0, 0] = 1.0
    vk = 4.0 * np.pi * rk / g2
    vk[0, 0, 0] = 0.0
    v = np.fft.ifftn(vk).real
    return np.ascontiguousarray(v.transpose(2, 1, 0)).ravel()


def occupancy_zero(nelec, nbands, kweights, spin=False, begin_state=0):
    """per k-point occupations, aufbau (mSetOccupancyZero)"""
    tot = float(np.sum(kweights))
    occ = np.zeros((len(kweights), nbands))
    for k, wk in enumerate(kweights):
        we = (1.0 if spin else 2.0) * wk / tot
        rem = nelec / 2.0 - begin_state
        for n in range(nbands):
            if rem >= 1.0:
                occ[k, n] = we
                rem -= 1.0
            elif rem > 0.0:
                occ[k, n] = we * rem
                rem = 0.0
    return occ


def occupancy_fermi(eigs, kweights, nelec, kbT):
    """per k-point Fermi-Dirac occupations and mu (mSetOccupancyTemperature)"""
    eigs = np.atleast_2d(np.asarray(eigs, dtype=np.float64))
    tot = float(np.sum(kweights))
    mu_min, mu_max = eigs[:, 0].min(), eigs[:, -1].max()
    mu = 0.5 * (mu_min + mu_max)
    go = abs(mu_max - mu_min) > 1e-9
    occ = np.zeros_like(eigs)

    def fill(mu):
        s = 0.0
        for k, wk in enumerate(kweights):
            c = 2.0 * wk / tot
            p = np.minimum((eigs[k] - mu) / kbT, 80.0)
            occ[k] = c / (np.exp(p) + 1.0)
            s += occ[k].sum()
        return s

    mu_used = mu
    while go:
        ne = fill(mu)
        mu_used = mu
        if ne < nelec:
            mu_min = mu
        elif ne > nelec:
            mu_max = mu
        else:
            break
        mu = 0.5 * (mu_min + mu_max)
        go = abs(mu_max - mu_min) > 1e-9
    return occ, mu_used


def inverse_m(m):
    """Gauss-Jordan inverse without pivoting (InverseM)"""
    m = np.array(m, dtype=np.float64)
    n = m.shape[0]
    inv = np.eye(n)
    for i in range(n):
        b = 1.0 / m[i, i]
        if np.isfinite(b):
            m[i, :] *= b
            inv[i, :] *= b
        for j in range(n):
            if j != i:
                b2 = m[j, i]
                m[j, :] -= m[i, :] * b2
                inv[j, :] -= inv[i, :] * b2
    return inv


class MixSimple:
    def __init__(self, ratio):
        self.ratio = ratio
        self.prev = None

    def init(self, rho):
        self.prev = np.array(rho, dtype=np.float64)

    def predict(self, rho):
        d = rho - self.prev
        out = d * self.ratio + self.prev
        self.prev = out.copy()
        return out, float(np.abs(d).sum())


class MixDIIS:
    def __init__(self, ratio, history):
        self.ratio, self.H = ratio, history
        self.step = 0
        self.rho = None

    def init(self, rho):
        n = np.asarray(rho).size
        self.rho = np.zeros((self.H, n))
        self.R = np.zeros((self.H, n))
        self.rho[0] = rho
        self.M = np.zeros((self.H + 1, self.H + 1))
        self.step = 0

    def _idx(self, m):
        return (m + self.H) % self.H

    def predict(self, rho_out):
        H = self.H
        k = self._idx(self.step)
        self.R[k] = rho_out - self.rho[k]
        diff = float(np.abs(self.R[k]).sum())
        # row / column k of R.R over the valid history, bordered by ones
        m_end = max(self.step - H + 1, 0)
        for m1 in range(self.step, m_end - 1, -1):
            m = self._idx(m1)
            self.M[m, k] = self.R[k] @ self.R[m]
        for m in range(H):
            self.M[k, m] = self.M[m, k]
            self.M[H, m] = 1.0
            self.M[m, H] = 1.0
        self.M[H, H] = 0.0
        nxt = self._idx(self.step + 1)
        if self.step + 1 >= H:
            inv = inverse_m(self.M)
            c = inv[H, :]  # the bordering row of the inverse
            rho_n = self.rho[nxt] * c[nxt]
            R_n = self.R[nxt] * c[nxt]
            for m in range(H):
                if m != nxt:
                    R_n = R_n + self.R[m] * c[m]
                    rho_n = rho_n + self.rho[m] * c[m]
            rho_n = rho_n + R_n * self.ratio
            self.rho[nxt] = rho_n
            self.R[nxt] = R_n
        else:
            self.rho[nxt] = self.rho[k] + self.R[k] * self.ratio
        self.step += 1
        return self.rho[nxt].copy(), diff


def scf_run(s, nelec, hpsi, nonlocal_energy, max_steps=80, tol=1e-10, mixing="diis", ratio=0.25, history=9):
    """Real Gamma-point SCF on system `s` (ncomp 1) with the oracle's operators:
    hpsi(s, psi, vloc=V) and nonlocal_energy(s, psi, occ).  Returns the
    history (a dict per step: E_tot and its parts, diff_rho, eigenvalues) and
    the final mixed density."""
    from scipy.sparse.linalg import LinearOperator, eigsh

    ng, dV = s.ng, s.dV
    shape, h = (s.nx, s.ny, s.nz), s.h
    vion = np.asarray(s.vloc, dtype=np.float64)
    rho = np.full(ng, nelec / (ng * dV))
    mix = MixDIIS(ratio, history) if mixing == "diis" else MixSimple(ratio)
    mix.init(rho)

    def potential(r):
        vh = hartree(r, shape, h)
        ex, ec, vx, vc = xc_lda(r)
        return vion + vh + vx + vc, vh, ex + ec

    V, vh, exc = potential(rho)
    nb = s.nbands
    hist = []
    v0 = np.random.default_rng(0).standard_normal(ng)
    prev = None
    for step in range(max_steps):
        op = LinearOperator((ng, ng), dtype=np.float64,
                            matvec=lambda x, V=V: hpsi(s, np.ascontiguousarray(x, dtype=np.float64).reshape(1, ng),
                                                       vloc=V).ravel())
        w, X = eigsh(op, k=nb, which="SA", tol=1e-13, v0=v0)
        o = np.argsort(w)
        w, X = w[o], X[:, o]
        v0 = X[:, 0].copy()
        psi = (X / np.sqrt(dV)).T.copy()
        occ = occupancy_zero(nelec, nb, [1.0])[0]
        rho_out = (occ[:, None] * psi * psi).sum(axis=0)
        e_nl = nonlocal_energy(s, psi, occ)
        band = float(occ @ w)
        e_kin = band - float(V @ rho_out) * dV - e_nl
        rho, diff = mix.predict(rho_out)
        diff *= dV / nelec
        V, vh, exc = potential(rho)
        e_ion = float(vion @ rho) * dV
        e_h = 0.5 * float(vh @ rho) * dV
        e_xc = float(exc.sum()) * dV
        E = e_kin + e_ion + e_h + e_xc + e_nl
        hist.append({"E_tot": E, "E_kin": e_kin, "E_ion": e_ion, "E_H": e_h, "E_xc": e_xc, "E_nl": e_nl,
                     "diff_rho": diff, "eigs": w, "dE": (prev - E) if prev is not None else np.nan})
        prev = E
        if diff < tol:
            break
    return hist, rho
