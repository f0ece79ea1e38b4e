"""Hand-written adjoint of H.psi on the GPU, checked through torch.autograd.

dE_band/dV_loc and dE_band/dR (C ABI) against central finite differences of
E_band computed with the GPU H.psi: steps 1e-4 bohr in R and 1e-6 Ha in V_loc
(E_band is linear in V_loc, so that check is exact up to rounding).
Tolerances: forces within 1e-5 Ha/bohr of the difference quotient (FP64 path);
V_loc gradient within 1e-6 relative plus the 1e-8 rounding floor of the quotient.  Bands: 8-atom cell, 4 bands after 30
LOBPCG iterations (any fixed bands give exact differences)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(qm, torch_cuda):
    torch = torch_cuda
    s = qm.system.config(4, nbands=4, grid=40)
    H = qm.hamiltonian.Hamiltonian(s)
    L = qm.lobpcg.LOBPCG(H, s.nbands, precision=64)
    x = torch.from_numpy(s.initial_bands()).to("cuda", torch.float64)
    L.solve(x, 30, first_time=True)
    L.close()
    occ = np.array([2.0, 2.0, 1.5, 0.5])
    return s, H, x, occ


def test_force_gradient_vs_finite_differences(qm, torch_cuda, setup):
    torch = torch_cuda
    s, H, x, occ = setup
    print("case adjoint forces C4 40^3 4 bands FP64", flush=True)
    BE = qm.adjoint.make_band_energy_function()
    v = torch.from_numpy(s.vloc.copy()).requires_grad_(True)
    R = torch.from_numpy(s.positions.copy()).requires_grad_(True)
    e = BE.apply(v, R, H, s, x, occ, 0)
    e.backward()
    grad_R = R.grad.numpy()
    h = 1e-4
    worst = 0.0
    for a in range(s.natoms):
        for d in range(3):
            vals = []
            for sg in (1, -1):
                Rp = s.positions.copy()
                Rp[a, d] += sg * h
                qm.adjoint.set_state(H, s, s.vloc, Rp)
                vals.append(qm.adjoint.band_energy(H, x, occ))
            fd = (vals[0] - vals[1]) / (2 * h)
            worst = max(worst, abs(fd - grad_R[a, d]))
    qm.adjoint.set_state(H, s, s.vloc, s.positions)
    print("max |dE/dR adjoint - finite difference| (Ha/bohr):", worst)
    assert worst <= 1e-5


def test_vloc_gradient_vs_finite_differences(qm, torch_cuda, setup):
    torch = torch_cuda
    s, H, x, occ = setup
    print("case adjoint V_loc C4 40^3 4 bands FP64", flush=True)
    BE = qm.adjoint.make_band_energy_function()
    v = torch.from_numpy(s.vloc.copy()).requires_grad_(True)
    R = torch.from_numpy(s.positions.copy()).requires_grad_(True)
    BE.apply(v, R, H, s, x, occ, 0).backward()
    gv = v.grad.numpy()
    rng = np.random.default_rng(0)
    for i in rng.integers(0, s.ng, size=6):
        dv = np.zeros(s.ng)
        dv[i] = 1e-6
        qm.adjoint.set_state(H, s, s.vloc + dv, s.positions)
        ep = qm.adjoint.band_energy(H, x, occ)
        qm.adjoint.set_state(H, s, s.vloc - dv, s.positions)
        em = qm.adjoint.band_energy(H, x, occ)
        fd = (ep - em) / 2e-6
        # rounding floor of the quotient: 1e-16 |E_band| / 2e-6 ~ 1e-9
        assert abs(fd - gv[i]) <= 1e-6 * abs(gv[i]) + 1e-8, (i, fd, gv[i])
    qm.adjoint.set_state(H, s, s.vloc, s.positions)
