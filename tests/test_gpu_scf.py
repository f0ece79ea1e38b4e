"""GPU parity of the SCF glue (HIP library through the C ABI) against the
oracle (oracle/scf_ref.py).

Tolerances: LDA and the grid integral <= 1e-13 relative (elementwise FP64 /
fixed-order sums); Hartree <= 1e-12 relative (rocFFT vs numpy FFT); the
mixers <= 1e-11 relative per step where the DIIS matrix is well conditioned,
the fixed point where it is not (see the test); the whole SCF at
convergence: E_tot within 1e-7 Ha and the density within 1e-6 relative (the
fixed point, reached through different eigensolvers)."""
import numpy as np
import pytest

import scf_ref as S

pytestmark = pytest.mark.gpu


def _sys(qm, n=16):
    return qm.system.small_system(n=n, L=6.0, natoms=2, lmax=1, rc=1.6, nbands=4)


def test_xc_lda_matches_oracle(qm, torch_cuda):
    torch = torch_cuda
    s = qm.system.small_system(n=20, L=5.0, natoms=1, nbands=1)
    H = qm.hamiltonian.Hamiltonian(s)
    rng = np.random.default_rng(0)
    rho = 10.0 ** rng.uniform(-10, 2, s.ng)  # both correlation branches and the 1e-8 cut
    rho[:5] = [0.0, 5e-9, 1e-8, 0.2387324146378430, 1.0]
    r = torch.from_numpy(rho).cuda()
    exc = torch.empty_like(r)
    v = qm.scf.xc_lda(H, r, exc=exc).cpu().numpy()
    ex, ec, vx, vc = S.xc_lda(rho)
    assert np.abs(v - (vx + vc)).max() <= 1e-13 * np.abs(vx + vc).max()
    assert np.allclose(v, vx + vc, rtol=1e-13, atol=0)
    assert np.allclose(exc.cpu().numpy(), ex + ec, rtol=1e-13, atol=0)
    H.close()


@pytest.mark.parametrize("n,L", [(16, 6.0), (20, 5.0)])
def test_hartree_matches_oracle(qm, torch_cuda, n, L):
    torch = torch_cuda
    s = qm.system.small_system(n=n, L=L, natoms=1, nbands=1)
    H = qm.hamiltonian.Hamiltonian(s)
    rho = np.random.default_rng(1).uniform(0, 1, s.ng)
    v = qm.scf.hartree(H, torch.from_numpy(rho).cuda()).cpu().numpy()
    want = S.hartree(rho, (s.nx, s.ny, s.nz), s.h)
    assert np.abs(v - want).max() <= 1e-12 * np.abs(want).max()
    H.close()


def test_grid_dot(qm, torch_cuda):
    torch = torch_cuda
    s = qm.system.config(0)
    H = qm.hamiltonian.Hamiltonian(s)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(s.ng), rng.standard_normal(s.ng)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    d1 = qm.scf.grid_dot(H, ta, tb)
    d2 = qm.scf.grid_dot(H, ta, tb)
    assert d1 == d2  # fixed summation order
    assert abs(d1 - a @ b) <= 1e-13 * np.abs(a * b).sum()
    H.close()


@pytest.mark.parametrize("mode,history", [("simple", 1), ("diis", 3), ("diis", 9)])
def test_mixer_matches_oracle(qm, torch_cuda, mode, history):
    """Steps of rho_out = F(rho_in), F affine and contractive, through the
    library's mixer and the oracle's.  Simple mixing and DIIS(3): <= 1e-11
    relative at every one of 15 steps.  DIIS(9): its ring fills with nearly
    dependent residuals, and the unpivoted Gauss-Jordan inverse then turns a
    last-bit change of one residual overlap into ~1e-4 of the step (the
    oracle itself moves that much under a 1-ulp change of a dot product), so
    the steps are compared up to the first DIIS step and the fixed point after
    40."""
    torch = torch_cuda
    s = qm.system.small_system(n=16, L=6.0, natoms=1, nbands=1)
    H = qm.hamiltonian.Hamiltonian(s)
    rng = np.random.default_rng(3)
    ng = s.ng
    scale = rng.uniform(-0.8, 0.8, ng)
    shift = rng.uniform(0.0, 1.0, ng)
    F = lambda x: scale * np.roll(x, 7) + shift  # noqa: E731
    x0 = rng.uniform(0, 1, ng)
    m = qm.scf.Mixer(H, mode, history, 0.3)
    o = S.MixSimple(0.3) if mode == "simple" else S.MixDIIS(0.3, history)
    t = torch.from_numpy(x0.copy()).cuda()
    m.init(t)
    o.init(x0)
    xg, xo = x0.copy(), x0.copy()
    steps = 40 if history == 9 else 15
    for step in range(steps):
        t = torch.from_numpy(F(xg)).cuda()
        dg = m.predict(t)
        xg = t.cpu().numpy()
        xo, do = o.predict(F(xo))
        if history < 9 or step < history - 1:
            err = np.abs(xg - xo).max() / np.abs(xo).max()
            assert err <= 1e-11, (step, err)
            assert abs(dg - do) <= 1e-11 * do, step
    if history == 9:
        # the fixed point x = F(x): x_i = scale_i x_{i-7} + shift_i, solved by numpy
        A = np.zeros((ng, ng))
        A[np.arange(ng), (np.arange(ng) - 7) % ng] = scale
        x_star = np.linalg.solve(np.eye(ng) - A, shift)
        assert np.abs(xg - x_star).max() <= 1e-8 * np.abs(x_star).max()
        assert np.abs(xo - x_star).max() <= 1e-8 * np.abs(x_star).max()
    m.close()
    H.close()


def test_scf_matches_oracle(qm, torch_cuda):
    """The whole loop on a gapped two-atom cell (4 electrons, 4 bands): the
    GPU SCF (FP64 LOBPCG, DIIS mixing) and the oracle's (exact eigenpairs,
    DIIS) reach the same fixed point."""
    s = _sys(qm)
    print("case SCF small 16^3 2 atoms 4 electrons", flush=True)
    import oracle as orc
    enl = lambda s_, psi, occ: orc.nonlocal_forces(s_, psi, occ)[0]  # noqa: E731
    ho, rho_o = S.scf_run(s, 4.0, orc.hpsi, enl, max_steps=40, tol=1e-11)
    scf = qm.scf.SCF(s, 4.0, precision=64, lobpcg_first=30, lobpcg_per_scf=6)
    r = scf.run(60, threshold_rho=1e-10)
    rho_g = scf.rho.cpu().numpy()
    print(f"GPU: {len(scf.history)} steps E_tot {r['E_tot']:.12f}; oracle: {len(ho)} steps {ho[-1]['E_tot']:.12f}")
    assert r["diff_rho"] < 1e-10
    assert abs(r["E_tot"] - ho[-1]["E_tot"]) < 1e-7
    assert np.abs(rho_g - rho_o).max() < 1e-6 * np.abs(rho_o).max()
    # occupied bands (the two lowest) tightly; the empty ones converge more
    # slowly under LOBPCG (the highest of the block last)
    de = np.abs(r["eigs"][0] - ho[-1]["eigs"])
    assert de[:2].max() < 1e-9 and de.max() < 1e-6, de
    for key in ("E_kin", "E_ion", "E_H", "E_xc", "E_nl"):
        assert abs(r[key] - ho[-1][key]) < 1e-6, key
    scf.close()


def test_scf_mixed_precision_close_to_fp64(qm, torch_cuda):
    """The same loop with FP32 vectors (FP64 subspace): the energy within
    1e-5 Ha of the FP64 loop's"""
    s = _sys(qm)
    a = qm.scf.SCF(s, 4.0, precision=64, lobpcg_first=30, lobpcg_per_scf=6)
    ra = a.run(60, threshold_rho=1e-10)
    a.close()
    b = qm.scf.SCF(s, 4.0, precision=32, lobpcg_first=30, lobpcg_per_scf=6)
    rb = b.run(60, threshold_rho=1e-6)
    b.close()
    print(f"FP64 {ra['E_tot']:.10f}  mixed {rb['E_tot']:.10f}")
    assert abs(ra["E_tot"] - rb["E_tot"]) < 1e-5
