"""Pins the CPU FP64 oracle against mathematics it cannot fake.

The reference ships no golden data for this path (its source/test_*.cpp are
timing drivers), so each oracle function is checked against an independent
statement of the same mathematics:

* FD8 kinetic operator: the closed-form symbol on plane waves, Gamma and k.
* whole H.psi: a dense matrix built from 1-D circulant FD8 matrices (Kronecker
  sums), diag(V) and P diag(E dV) P^H with Bloch phases, on tiny grids.
* Hermiticity of H(k), Bloch covariance of the nonlocal term.
* real Y_lm: scipy.special's spherical harmonics (Condon-Shortley removed).
* projector tables: the generator's solid-harmonic polynomials (independent
  formula) and sphere membership.
* adjoint: central finite differences of E_band in V_loc and of E_nl in R_a.
"""
import math

import numpy as np
import pytest


# --------------------------------------------------------------------------
# FD8 symbol

CJ = np.array([8064.0, -1008.0, 128.0, -9.0]) / 5040.0
AJ = np.array([672.0, -168.0, 32.0, -3.0]) / 840.0


def lam2(theta, h):
    j = np.arange(1, 5)
    return (-14350.0 / 5040.0 + 2.0 * np.sum(CJ * np.cos(j * theta))) / h ** 2


def lam1(theta, h):
    j = np.arange(1, 5)
    return 2.0 * np.sum(AJ * np.sin(j * theta)) / h


def _plane_wave(s, m):
    hx, hy, hz = s.h
    ix = np.arange(s.nx)
    iy = np.arange(s.ny)
    iz = np.arange(s.nz)
    th = [2 * math.pi * m[0] / s.nx, 2 * math.pi * m[1] / s.ny, 2 * math.pi * m[2] / s.nz]
    ph = (th[2] * iz)[:, None, None] + (th[1] * iy)[None, :, None] + (th[0] * ix)[None, None, :]
    return np.exp(1j * ph).ravel(), th


@pytest.mark.parametrize("m", [(0, 0, 0), (1, 0, 0), (2, -3, 1), (5, 4, -7), (8, 8, 8)])
@pytest.mark.parametrize("kred", [(0, 0, 0), (0.25, 0, 0), (0.25, 0.25, 0.25), (0.1, -0.3, 0.45)])
def test_fd8_plane_wave_symbol(qm, orc, m, kred):
    s = qm.system.small_system(n=16, L=5.0, natoms=0, k_red=kred)
    k = s.kpoints[0]
    z, th = _plane_wave(s, m)
    psi = np.stack([z.real, z.imag])
    out = orc.kinetic(s, psi, k=k, ncomp=2)
    h = s.h
    sym = -0.5 * (sum(lam2(th[d], h[d]) for d in range(3))
                  - 2.0 * sum(k[d] * lam1(th[d], h[d]) for d in range(3)) - float(np.dot(k, k)))
    got = out[0] + 1j * out[1]
    np.testing.assert_allclose(got, sym * z, rtol=0, atol=1e-12 * max(1.0, abs(sym)))


def test_fd8_real_path_matches_complex(qm, orc):
    s = qm.system.small_system(n=12, natoms=0)
    rng = np.random.default_rng(1)
    re = rng.standard_normal((3, s.ng))
    out_r = orc.kinetic(s, re, ncomp=1)
    cpx = np.zeros((6, s.ng))
    cpx[0::2] = re
    out_c = orc.kinetic(s, cpx, ncomp=2)
    np.testing.assert_array_equal(out_r, out_c[0::2])
    np.testing.assert_array_equal(out_c[1::2], 0.0)


# --------------------------------------------------------------------------
# dense brute force

def _circulant(n, coeffs, sym):
    """n x n periodic matrix of sum_j coeffs[j] (shift(+j) +/- shift(-j))."""
    A = np.zeros((n, n))
    for j, c in enumerate(coeffs, start=1):
        for i in range(n):
            A[i, (i + j) % n] += c
            A[i, (i - j) % n] += c if sym else -c
    return A


def dense_hamiltonian(s, k, local=True, nonlocal_=True):
    """Independent dense H(k) on the grid (x fastest); local=False leaves out the
    kinetic and V_loc terms, nonlocal_=False the projectors."""
    hx, hy, hz = s.h
    n = (s.nx, s.ny, s.nz)
    h = (hx, hy, hz)
    I = [np.eye(nd) for nd in n]
    D2 = [(_circulant(n[d], CJ / h[d] ** 2, True) + np.eye(n[d]) * (-14350.0 / 5040.0) / h[d] ** 2) for d in range(3)]
    D1 = [_circulant(n[d], AJ / h[d], False) for d in range(3)]

    def kron3(ax, ay, az):  # index = ix + nx*(iy + ny*iz) -> kron(z, y, x)
        return np.kron(az, np.kron(ay, ax))

    lap = kron3(D2[0], I[1], I[2]) + kron3(I[0], D2[1], I[2]) + kron3(I[0], I[1], D2[2])
    grad = [kron3(D1[0], I[1], I[2]), kron3(I[0], D1[1], I[2]), kron3(I[0], I[1], D1[2])]
    H = -0.5 * (lap + 2j * sum(k[d] * grad[d] for d in range(3)) - float(np.dot(k, k)) * np.eye(s.ng))
    H = H + np.diag(s.vloc)
    if not local:
        H = np.zeros_like(H)
    # nonlocal: sum_a e^{-ik.r} P E dV P^T e^{ik.r}
    for a in range(s.natoms if nonlocal_ else 0):
        v0, v1 = s.vox_off[a], s.vox_off[a + 1]
        V = v1 - v0
        M = s.nproj[a]
        xyz = s.vox_xyz[v0:v1]
        g = (xyz[:, 0] % s.nx) + s.nx * ((xyz[:, 1] % s.ny) + s.ny * (xyz[:, 2] % s.nz))
        ph = np.exp(1j * (xyz[:, 0] * hx * k[0] + xyz[:, 1] * hy * k[1] + xyz[:, 2] * hz * k[2]))
        P = s.table[s.tab_off[a]:s.tab_off[a + 1]].reshape(M, V).T  # V x M
        E = s.energy[s.proj_off[a]:s.proj_off[a + 1]]
        B = np.zeros((s.ng, M), dtype=complex)
        np.add.at(B, g, np.conj(ph)[:, None] * P)  # column w: e^{-ik.r} p_w on the grid
        H = H + (B * (E * s.dV)) @ B.conj().T
    return H


@pytest.mark.parametrize("kred,natoms,lmax,rc", [((0, 0, 0), 2, 2, 1.4), ((0.25, 0.1, -0.3), 2, 2, 1.4),
                                                  ((0.5, 0.5, 0.5), 3, 1, 2.6)])
def test_hpsi_matches_dense_matrix(qm, orc, kred, natoms, lmax, rc):
    # rc = 2.6 > L/2 wraps spheres around the cell and overlaps them
    s = qm.system.small_system(n=8, L=4.0, natoms=natoms, lmax=lmax, rc=rc, k_red=kred, seed=11)
    k = s.kpoints[0]
    H = dense_hamiltonian(s, k)
    rng = np.random.default_rng(3)
    z = rng.standard_normal((2, s.ng)) + 1j * rng.standard_normal((2, s.ng))
    psi = np.empty((4, s.ng))
    psi[0::2] = z.real
    psi[1::2] = z.imag
    out = orc.hpsi(s, psi, k=k, ncomp=2)
    got = out[0::2] + 1j * out[1::2]
    want = z @ H.T
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_hermitian_with_k(qm, orc):
    s = qm.system.small_system(n=14, L=4.5, natoms=3, rc=1.8, k_red=(0.3, -0.2, 0.1), seed=5)
    k = s.kpoints[0]
    rng = np.random.default_rng(9)
    u = rng.standard_normal((2, s.ng))
    v = rng.standard_normal((2, s.ng))
    Hu = orc.hpsi(s, u, k=k, ncomp=2)
    Hv = orc.hpsi(s, v, k=k, ncomp=2)
    cu, cv = u[0] + 1j * u[1], v[0] + 1j * v[1]
    cHu, cHv = Hu[0] + 1j * Hu[1], Hv[0] + 1j * Hv[1]
    a = np.vdot(cu, cHv)
    b = np.vdot(cHu, cv)
    assert abs(a - b) < 1e-11 * abs(a)


def _herm_defect(x, y, c):
    """max |G - G^H| over the largest off-diagonal |G|, G = <x_i | y_j> over a
    SoA band block: the Hermiticity check the full-size GPU tests apply to
    y = H x."""
    nb = x.shape[0] // c
    xs, ys = x.reshape(nb, c, -1), y.reshape(nb, c, -1)
    G = [[xs[:, p] @ ys[:, q].T for q in range(c)] for p in range(c)]
    gre, gim = (G[0][0], np.zeros_like(G[0][0])) if c == 1 else (G[0][0] + G[1][1], G[0][1] - G[1][0])
    off = ~np.eye(nb, dtype=bool)
    scale = max(np.abs(gre[off]).max(), np.abs(gim[off]).max())
    return max(np.abs(gre - gre.T).max(), np.abs(gim + gim.T).max()) / scale


def test_block_hermiticity_detects_planted_mistakes(qm, orc):
    """The block Hermiticity check holds for the oracle's H at a k-point and
    fails, far above the GPU tests' FP32 tolerance (1e-5), for two planted
    non-Hermitian mistakes: the k-gradient term without its factor i, and
    V_loc multiplying the neighbouring grid point's value."""
    s = qm.system.small_system(n=14, L=4.5, natoms=3, rc=1.8, nbands=6, ncomp=2, k_red=(0.3, -0.2, 0.1), seed=5)
    k = s.kpoints[0]
    x = np.random.default_rng(1).uniform(-0.5, 0.5, size=(12, s.ng))
    y = orc.hpsi(s, x, k=k)
    assert _herm_defect(x, y, 2) <= 1e-12
    # g = -i k.grad x (SoA rows), from the kinetic term at k and at k = 0
    g = orc.kinetic(s, x, k=k) - orc.kinetic(s, x) - 0.5 * float(np.dot(k, k)) * x
    ig = np.empty_like(g)  # i g = k.grad x, an anti-Hermitian term
    ig[0::2], ig[1::2] = -g[1::2], g[0::2]
    assert _herm_defect(x, y - g + ig, 2) > 1e-3
    shifted = y + s.vloc * (np.roll(x, 1, axis=1) - x)
    assert _herm_defect(x, shifted, 2) > 1e-3
    # real Gamma-point block, same check
    r = qm.system.small_system(n=14, L=4.5, natoms=3, rc=1.8, nbands=6, seed=5)
    xr = np.random.default_rng(2).uniform(-0.5, 0.5, size=(6, r.ng))
    yr = orc.hpsi(r, xr)
    assert _herm_defect(xr, yr, 1) <= 1e-12
    assert _herm_defect(xr, yr + r.vloc * (np.roll(xr, 1, axis=1) - xr), 1) > 1e-3


def test_nonlocal_bloch_covariance(qm, orc):
    # k = reciprocal lattice vector: e^{ik.r} is periodic on the grid, so
    # Vnl(k) u = e^{-ik.r} Vnl(0) (e^{ik.r} u) exactly.
    s = qm.system.small_system(n=12, L=4.0, natoms=2, rc=1.6, seed=2)
    G = np.array([1.0, -2.0, 1.0]) * 2 * math.pi / 4.0
    hx, hy, hz = s.h
    ix = np.arange(s.nx)
    ph = ((G[2] * hz * ix)[:, None, None] + (G[1] * hy * ix)[None, :, None] + (G[0] * hx * ix)[None, None, :]).ravel()
    e = np.exp(1j * ph)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(s.ng) + 1j * rng.standard_normal(s.ng)
    a = orc.nonlocal_(s, np.stack([u.real, u.imag]), k=G, ncomp=2)
    w = e * u
    b = orc.nonlocal_(s, np.stack([w.real, w.imag]), k=(0, 0, 0), ncomp=2)
    b = np.conj(e) * (b[0] + 1j * b[1])
    np.testing.assert_allclose(a[0] + 1j * a[1], b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_nonlocal_rank(qm, orc):
    s = qm.system.small_system(n=8, L=4.0, natoms=2, lmax=1, rc=1.5, seed=6)
    cols = np.stack([orc.nonlocal_(s, np.eye(s.ng)[i][None], ncomp=1)[0] for i in range(s.ng)], axis=1)
    r = np.linalg.matrix_rank(cols, tol=1e-10 * np.abs(cols).max())
    assert r <= int(s.nproj.sum())
    assert r == int(s.nproj.sum())  # generic positions: the projectors are independent


# --------------------------------------------------------------------------
# spherical harmonics and projector tables

def _real_ylm_scipy(l, m, x, y, z):
    from scipy.special import sph_harm_y
    r = math.sqrt(x * x + y * y + z * z)
    theta = math.acos(z / r)
    phi = math.atan2(y, x)
    if m == 0:
        return sph_harm_y(l, 0, theta, phi).real
    Y = sph_harm_y(l, abs(m), theta, phi)
    cs = (-1) ** abs(m)  # remove Condon-Shortley
    return math.sqrt(2.0) * cs * (Y.real if m > 0 else Y.imag)


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_ylm_matches_scipy(orc, l):
    rng = np.random.default_rng(l)
    for _ in range(20):
        x, y, z = rng.standard_normal(3)
        for m in range(-l, l + 1):
            assert abs(orc.ylm(l, m, x, y, z) - _real_ylm_scipy(l, m, x, y, z)) < 1e-13


def test_ylm_orthonormal(orc):
    # Gauss-Legendre in cos(theta) x uniform phi integrates products up to l = 3 exactly
    xg, wg = np.polynomial.legendre.leggauss(12)
    phis = np.arange(24) * 2 * math.pi / 24
    lm = [(l, m) for l in range(4) for m in range(-l, l + 1)]
    G = np.zeros((16, 16))
    for ct, w in zip(xg, wg):
        st = math.sqrt(1 - ct * ct)
        for ph in phis:
            v = np.array([orc.ylm(l, m, st * math.cos(ph), st * math.sin(ph), ct) for l, m in lm])
            G += w * (2 * math.pi / 24) * np.outer(v, v)
    np.testing.assert_allclose(G, np.eye(16), atol=1e-13)


def test_ylm_origin(orc):
    assert orc.ylm(0, 0, 0, 0, 0) == pytest.approx(1 / math.sqrt(4 * math.pi))
    assert orc.ylm(1, 1, 0, 0, 0) == 0.0
    assert orc.ylm(2, -2, 0, 0, 0) == 0.0


@pytest.mark.parametrize("cfg", [0, 3, "wlike"])
def test_projector_tables_match_oracle(qm, orc, cfg):
    """Generator tables (solid-harmonic polynomials) against the oracle's
    (Y_lm from the associated Legendre form, pinned to scipy above), including
    the W-like set: l <= 3 with two radial projectors, 32 per atom, in the
    order l, radial index, m."""
    if cfg == "wlike":
        s = qm.system.wlike(0, nbands=2)
        assert (s.nproj == 32).all()
    else:
        s = qm.system.config(cfg, grid=48 if cfg == 3 else None)
    t = orc.projector_table(s)
    np.testing.assert_allclose(s.table, t, rtol=0, atol=1e-14)


@pytest.mark.parametrize("which", ["C4", "wlike"])
def test_projector_gradient_tables(qm, orc, which):
    # analytic d p / d r of the generator vs central differences of the oracle
    # tables in the atom position (d p(r - R)/d R = -d p/d r), same voxels.
    s = qm.system.config(4, grid=32) if which == "C4" else qm.system.wlike(4, nbands=2, grid=32)
    h = 1e-5
    for d in range(3):
        dp = s.positions.copy()
        dm = s.positions.copy()
        dp[:, d] += h
        dm[:, d] -= h
        fd = -(orc.projector_table(s, dp) - orc.projector_table(s, dm)) / (2 * h)
        # the sphere cut makes voxels at r ~ rc discontinuous; skip them
        ok = np.abs(fd) < 50.0
        np.testing.assert_allclose(s.dtable[d][ok], fd[ok], rtol=0, atol=2e-8)


def test_sphere_membership(qm):
    s = qm.system.config(0)
    hx, hy, hz = s.h
    for a in range(s.natoms):
        xyz = s.vox_xyz[s.vox_off[a]:s.vox_off[a + 1]] * np.array([hx, hy, hz])
        r = np.linalg.norm(xyz - s.positions[a], axis=1)
        assert (r < s.rc[a]).all()
        # count against brute force over a bounding box
        R = s.positions[a]
        rng_ = [np.arange(math.floor((R[d] - 3) / hx), math.ceil((R[d] + 3) / hx) + 1) for d in range(3)]
        X, Y, Z = np.meshgrid(*rng_, indexing="ij")
        rr = (X * hx - R[0]) ** 2 + (Y * hy - R[1]) ** 2 + (Z * hz - R[2]) ** 2
        assert int((rr < s.rc[a] ** 2).sum()) == len(r)


# --------------------------------------------------------------------------
# adjoint

def test_band_density_is_dE_dV(qm, orc):
    s = qm.system.small_system(n=10, natoms=2, ncomp=2, k_red=(0.2, 0, 0.1))
    k = s.kpoints[0]
    rng = np.random.default_rng(8)
    psi = rng.standard_normal((6, s.ng))
    occ = np.array([2.0, 1.5, 0.5])

    def eband(v):
        hp = orc.hpsi(s, psi, k=k, ncomp=2, vloc=v)
        return sum(occ[n] * (psi[2 * n] @ hp[2 * n] + psi[2 * n + 1] @ hp[2 * n + 1]) for n in range(3)) * s.dV

    rho = orc.band_density(s, psi, occ, ncomp=2)
    for i in [0, 17, 333, s.ng - 1]:
        dv = np.zeros(s.ng)
        dv[i] = 0.5  # E_band is linear in V_loc: central differences are exact up to rounding
        fd = (eband(s.vloc + dv) - eband(s.vloc - dv)) / 1.0
        assert abs(fd - rho[i] * s.dV) < 1e-9 * max(1.0, abs(fd))


def test_nonlocal_energy_matches_apply(qm, orc):
    s = qm.system.small_system(n=12, natoms=3, rc=1.5, ncomp=2, k_red=(0.25, 0.25, 0), seed=3)
    k = s.kpoints[0]
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((4, s.ng))
    occ = np.array([1.0, 0.5])
    e, _ = orc.nonlocal_forces(s, psi, occ, k=k, ncomp=2)
    vp = orc.nonlocal_(s, psi, k=k, ncomp=2)
    want = sum(occ[n] * (psi[2 * n] @ vp[2 * n] + psi[2 * n + 1] @ vp[2 * n + 1]) for n in range(2)) * s.dV
    assert abs(e - want) < 1e-12 * abs(want)


@pytest.mark.parametrize("ncomp,kred", [(1, (0, 0, 0)), (2, (0.25, 0.1, 0))])
def test_nonlocal_forces_finite_difference(qm, orc, ncomp, kred):
    s = qm.system.small_system(n=16, L=5.0, natoms=2, rc=1.8, ncomp=ncomp, k_red=kred, seed=12)
    k = s.kpoints[0]
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((2 * ncomp, s.ng))
    occ = np.array([2.0, 1.0])
    e0, f = orc.nonlocal_forces(s, psi, occ, k=k, ncomp=ncomp)
    h = 1e-4
    for a in range(s.natoms):
        for d in range(3):
            vals = []
            for sgn in (+1, -1):
                pos = s.positions.copy()
                pos[a, d] += sgn * h
                t = orc.projector_table(s, pos)
                s2 = qm.system.System(**{f.name: getattr(s, f.name) for f in __import__("dataclasses").fields(s)})
                s2.table = t
                vals.append(orc.nonlocal_forces(s2, psi, occ, k=k, ncomp=ncomp)[0])
            fd = -(vals[0] - vals[1]) / (2 * h)
            assert abs(fd - f[a, d]) < 1e-6 * max(1.0, np.abs(f).max()), (a, d, fd, f[a, d])


# --------------------------------------------------------------------------
# real-time propagators

def _soa(z):
    out = np.empty((2 * z.shape[0], z.shape[1]))
    out[0::2], out[1::2] = z.real, z.imag
    return out


@pytest.mark.parametrize("op", ["full", "local", "nonlocal"])
def test_td_taylor_matches_exponential(qm, orc, op):
    """The oracle's Taylor propagator against exp(-i dt Op) of the dense H(k)
    (scipy expm) on an 8^3 k-point system: the error is the truncated tail,
    O(dt^5) for order 4, so halving dt divides it by ~32.  A wrong sign or
    factor in any of the four terms leaves an O(dt^n) error with n <= 4
    (ratio <= 16); order 8 shrinks the error well below order 4."""
    import scipy.linalg
    s = qm.system.small_system(n=8, L=4.0, natoms=2, lmax=2, rc=1.4, nbands=2, ncomp=2,
                               k_red=(0.25, 0.1, -0.3), seed=11)
    k = s.kpoints[0]
    Hd = dense_hamiltonian(s, k, local=op != "nonlocal", nonlocal_=op != "local")
    rng = np.random.default_rng(5)
    z = rng.standard_normal((2, s.ng)) + 1j * rng.standard_normal((2, s.ng))
    dt0 = 0.2 / np.abs(np.linalg.eigvalsh(Hd)).max()
    errs = []
    for dt in (dt0, dt0 / 2):
        want = z @ scipy.linalg.expm(-1j * dt * Hd).T
        got = orc.td_taylor(s, _soa(z), dt, order=4, op=op, k=k)
        errs.append(np.abs(got[0::2] + 1j * got[1::2] - want).max() / np.abs(want).max())
    ratio = errs[0] / errs[1]
    print(op, "errors", errs, "ratio", ratio)
    assert 24.0 < ratio < 40.0
    assert errs[0] < 1e-4
    want = z @ scipy.linalg.expm(-1j * dt0 * Hd).T
    got8 = orc.td_taylor(s, _soa(z), dt0, order=8, op=op, k=k)
    e8 = np.abs(got8[0::2] + 1j * got8[1::2] - want).max() / np.abs(want).max()
    assert e8 < 1e-3 * errs[0]


def test_td_expv_matches_exponential(qm, orc):
    """e^{-i V dt} psi against the exponential of the diagonal V_loc (sign pinned)."""
    s = qm.system.small_system(n=8, L=4.0, natoms=2, nbands=2, ncomp=2, seed=3)
    rng = np.random.default_rng(6)
    z = rng.standard_normal((2, s.ng)) + 1j * rng.standard_normal((2, s.ng))
    dt = 0.37
    got = orc.td_expv(s, _soa(z), dt)
    want = z * np.exp(-1j * dt * s.vloc)[None, :]
    np.testing.assert_allclose(got[0::2] + 1j * got[1::2], want, rtol=0, atol=1e-14 * np.abs(want).max())


# --------------------------------------------------------------------------
# spectral (FFT) kinetic term

@pytest.mark.parametrize("dims,kred", [((8, 8, 8), (0.0, 0.0, 0.0)), ((8, 6, 10), (0.25, -0.1, 0.4)),
                                       ((7, 9, 6), (0.5, 0.5, -0.5))])
def test_fft_kinetic_plane_waves(qm, orc, dims, kred):
    """Every grid plane wave e^{2 pi i m.r / L} is an eigenvector of the
    spectral kinetic term with eigenvalue |G_m + k|^2 / 2, G_m = 2 pi m / L and m
    the wrapped index (i > n/2 -> i - n): the Nyquist index n/2 of an even
    axis stays +n/2, which matters at k != 0 (the reference's
    OperateKineticInKspace_bundle convention)."""
    nx, ny, nz = dims
    L = (3.0, 2.5, 3.5)
    s = qm.system.System("pw", nx, ny, nz, L, np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros(0), 1, 2,
                         np.array([kred]) * (2.0 * np.pi / np.array(L)), 1).build()
    k = s.kpoints[0]
    z_ax = [np.arange(n) for n in dims]
    for m in [(0, 0, 0), (1, 0, 0), (0, -1, 2), (nx // 2, ny // 2, nz // 2), (-1, ny // 2, 1)]:
        ph = [np.exp(2j * np.pi * m[d] * z_ax[d] / dims[d]) for d in range(3)]
        pw = (ph[2][:, None, None] * ph[1][None, :, None] * ph[0][None, None, :]).ravel()
        psi = np.stack([pw.real, pw.imag])
        out = orc.fft_kinetic(s, psi, k=k, ncomp=2)
        got = out[0] + 1j * out[1]
        mw = [(m[d] % dims[d]) for d in range(3)]
        mw = [mm - dims[d] if mm * 2 > dims[d] else mm for d, mm in enumerate(mw)]
        e = 0.5 * sum((2 * np.pi * mw[d] / L[d] + k[d]) ** 2 for d in range(3))
        np.testing.assert_allclose(got, e * pw, rtol=0, atol=1e-11 * max(1.0, e))


def test_fft_kinetic_real_matches_complex(qm, orc):
    s = qm.system.small_system(n=12, natoms=0)
    rng = np.random.default_rng(2)
    re = rng.standard_normal((3, s.ng))
    out_r = orc.fft_kinetic(s, re, ncomp=1)
    cpx = np.zeros((6, s.ng))
    cpx[0::2] = re
    out_c = orc.fft_kinetic(s, cpx, ncomp=2)
    np.testing.assert_allclose(out_r, out_c[0::2], rtol=0, atol=1e-12 * np.abs(out_r).max())
    np.testing.assert_allclose(out_c[1::2], 0.0, atol=1e-12 * np.abs(out_r).max())


def test_td_kinetic_fft_matches_plane_wave_phase(qm, orc):
    """The kinetic propagator multiplies each plane wave by e^{-i dt |G + k|^2 / 2}."""
    s = qm.system.small_system(n=8, L=3.0, natoms=0, nbands=1, ncomp=2, k_red=(0.25, 0.0, -0.5))
    k = s.kpoints[0]
    n = 8
    ax = np.arange(n)
    m = (3, -2, 4)
    ph = [np.exp(2j * np.pi * m[d] * ax / n) for d in range(3)]
    pw = (ph[2][:, None, None] * ph[1][None, :, None] * ph[0][None, None, :]).ravel()
    dt = 0.7
    out = orc.td_kinetic_fft(s, np.stack([pw.real, pw.imag]), dt, k=k)
    mw = [(mm % n) - n if (mm % n) * 2 > n else (mm % n) for mm in m]
    e = 0.5 * sum((2 * np.pi * mw[d] / 3.0 + k[d]) ** 2 for d in range(3))
    np.testing.assert_allclose(out[0] + 1j * out[1], np.exp(-1j * dt * e) * pw, rtol=0, atol=1e-12)
