"""The SCF oracle (oracle/scf_ref.py) pinned by what the mathematics fixes."""
import numpy as np
import pytest

import scf_ref as S


def test_lda_exchange_closed_form():
    rho = np.logspace(-6, 2, 40)
    ex, _, vx, _ = S.xc_lda(rho)
    want = -0.75 * np.cbrt(3.0 / np.pi) * rho ** (4.0 / 3.0)
    assert np.allclose(ex, want, rtol=1e-13, atol=0)
    assert np.allclose(vx, 4.0 / 3.0 * want / rho, rtol=1e-13, atol=0)


def test_lda_potential_is_the_derivative():
    """V = d(rho e)/d rho on both correlation branches (rs < 1 and rs >= 1)"""
    rho = np.array([1e-4, 3e-3, 0.05, 0.2, 0.5, 2.0, 10.0])
    eps = 1e-6 * rho
    ex_p, ec_p, _, _ = S.xc_lda(rho + eps)
    ex_m, ec_m, _, _ = S.xc_lda(rho - eps)
    _, _, vx, vc = S.xc_lda(rho)
    assert np.allclose((ex_p - ex_m) / (2 * eps), vx, rtol=1e-8)
    assert np.allclose((ec_p - ec_m) / (2 * eps), vc, rtol=1e-7)
    # a sign flip of the rho de/drho term would miss by far more
    assert np.abs(vc - (ec_p - ec_m) / (2 * eps)).max() < 1e-9


def test_lda_correlation_branches_meet_at_rs_one():
    """PZ81's two correlation forms were fitted to join at rs = 1"""
    c2 = np.cbrt(3.0 / (4.0 * np.pi))
    rho1 = c2 ** 3  # rs = 1
    r = np.array([rho1 * (1 - 1e-9), rho1 * (1 + 1e-9)])
    _, ec, _, vc = S.xc_lda(r)
    assert abs(ec[0] / r[0] - ec[1] / r[1]) < 5e-5
    assert abs(vc[0] - vc[1]) < 5e-4
    assert np.all(np.array(S.xc_lda(np.array([0.0, 5e-9]))) == 0.0)


@pytest.mark.parametrize("shape,h", [((16, 12, 10), (0.3, 0.4, 0.5)), ((9, 9, 9), (0.5, 0.5, 0.5))])
def test_hartree_plane_wave(shape, h):
    nx, ny, nz = shape
    L = np.array(shape) * np.array(h)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = ix * h[0], iy * h[1], iz * h[2]
    G = 2 * np.pi * np.array([1, -2, 1]) / L
    rho = (np.cos(G[0] * x + G[1] * y + G[2] * z) + 0.7).ravel()
    v = S.hartree(rho, shape, h)
    want = (4 * np.pi / (G @ G) * np.cos(G[0] * x + G[1] * y + G[2] * z)).ravel()
    assert np.abs(v - want).max() <= 1e-12 * np.abs(want).max()


def test_hartree_poisson_equation():
    """-lap V = 4 pi (rho - <rho>) with the spectral Laplacian"""
    shape, h = (12, 10, 14), (0.4, 0.5, 0.35)
    rho = np.random.default_rng(0).uniform(0, 1, np.prod(shape))
    v = S.hartree(rho, shape, h)
    nx, ny, nz = shape
    vk = np.fft.fftn(v.reshape(nz, ny, nx).transpose(2, 1, 0))
    lap = -np.fft.ifftn(S._g2(shape, h) * vk).real
    want = 4 * np.pi * (rho.reshape(nz, ny, nx).transpose(2, 1, 0) - rho.mean())
    assert np.abs(-lap - want).max() <= 1e-10 * np.abs(want).max()


def test_occupancy_zero():
    occ = S.occupancy_zero(7.0, 6, [1, 3])
    assert np.allclose(occ.sum(axis=1), [0.5 * 3.5, 1.5 * 3.5])
    assert occ[0, 3] == pytest.approx(0.5 * 0.5) and occ[0, 4] == 0.0
    assert occ.sum() == pytest.approx(7.0)


def test_occupancy_fermi_counts_electrons():
    rng = np.random.default_rng(1)
    eigs = np.sort(rng.uniform(-1, 1, (3, 10)), axis=1)
    for kT in (1e-3, 0.02, 0.2):
        occ, mu = S.occupancy_fermi(eigs, [1, 2, 1], 9.0, kT)
        assert abs(occ.sum() - 9.0) < 1e-6
        assert np.all(np.diff(occ, axis=1) <= 1e-15)
        assert eigs.min() < mu < eigs.max()
    # low temperature -> aufbau over the merged spectrum
    e = np.array([[-1.0, -0.5, 0.0, 0.5]])
    occ, mu = S.occupancy_fermi(e, [1], 4.0, 1e-4)
    assert np.allclose(occ, [[2, 2, 0, 0]], atol=1e-10)


def test_inverse_m():
    rng = np.random.default_rng(2)
    R = rng.standard_normal((6, 40))
    M = np.ones((7, 7))
    M[:6, :6] = R @ R.T
    M[6, 6] = 0.0
    assert np.allclose(S.inverse_m(M), np.linalg.inv(M), rtol=1e-9, atol=1e-11)


def test_simple_mixing():
    m = S.MixSimple(0.3)
    a, b = np.array([1.0, 2.0]), np.array([2.0, 0.0])
    m.init(a)
    out, d = m.predict(b)
    assert np.allclose(out, a + 0.3 * (b - a)) and d == pytest.approx(3.0)


@pytest.mark.parametrize("history", [3, 9])
def test_diis_solves_linear_fixed_point(history):
    """rho_out = A rho_in + c with a contraction A: DIIS reaches the fixed
    point (I - A)^-1 c, which simple mixing at the same ratio approaches far
    more slowly"""
    rng = np.random.default_rng(3)
    n = 30
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * rng.uniform(-0.9, 0.9, n)) @ Q.T
    c = rng.standard_normal(n)
    x_star = np.linalg.solve(np.eye(n) - A, c)
    m = S.MixDIIS(0.25, history)
    x = np.zeros(n)
    m.init(x)
    for _ in range(60):
        x, _ = m.predict(A @ x + c)
    s = S.MixSimple(0.25)
    y = np.zeros(n)
    s.init(y)
    for _ in range(60):
        y, _ = s.predict(A @ y + c)
    err_diis = np.abs(x - x_star).max()
    assert err_diis < 1e-8 * np.abs(x_star).max() or history < 9
    assert err_diis < 0.1 * np.abs(y - x_star).max()


def test_scf_fixed_point():
    """The oracle's SCF loop on a gapped two-atom cell: DIIS and simple mixing
    reach the same energy, and the final density reproduces itself (one more
    diagonalisation in its own potential gives it back)."""
    import oracle as orc
    import qumasun_amd as qm
    from scipy.sparse.linalg import LinearOperator, eigsh
    s = qm.system.small_system(n=16, L=6.0, natoms=2, lmax=1, rc=1.6, nbands=4)
    enl = lambda s_, psi, occ: orc.nonlocal_forces(s_, psi, occ)[0]  # noqa: E731
    hd, rho = S.scf_run(s, 4.0, orc.hpsi, enl, max_steps=40, tol=1e-11, mixing="diis")
    hs, _ = S.scf_run(s, 4.0, orc.hpsi, enl, max_steps=70, tol=1e-11, mixing="simple")
    assert hd[-1]["diff_rho"] < 1e-11 and len(hd) < 40
    assert abs(hd[-1]["E_tot"] - hs[-1]["E_tot"]) < 2e-6
    vh = S.hartree(rho, (s.nx, s.ny, s.nz), s.h)
    ex, ec, vx, vc = S.xc_lda(rho)
    V = s.vloc + vh + vx + vc
    op = LinearOperator((s.ng, s.ng), dtype=np.float64,
                        matvec=lambda x: orc.hpsi(s, np.ascontiguousarray(x).reshape(1, -1), vloc=V).ravel())
    w, X = eigsh(op, k=4, which="SA", tol=1e-13)
    psi = X.T / np.sqrt(s.dV)
    occ = S.occupancy_zero(4.0, 4, [1.0])[0]
    rho_out = (occ[np.argsort(w)][:, None] * (psi[np.argsort(w)] ** 2)).sum(axis=0)
    assert np.abs(rho_out - rho).max() < 1e-8 * np.abs(rho).max()
