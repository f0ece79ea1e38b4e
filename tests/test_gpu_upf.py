"""UPF pseudopotentials through the whole path on the GPU: a UPF file (written
here, with the projector shape of the reference's W.upf: two radial
projectors for each l <= 3, 32 per atom) loaded by qm_upf_load, tabulated by
qm_upf_tables, uploaded by qm_set_projectors, then H.psi and the nonlocal
forces against the CPU FP64 oracle on the same tables.

Tolerances as the other parity tests: FP64 <= 1e-12 relative, FP32 H.psi
<= 2e-6, FP32 forces <= 1e-5."""
import numpy as np
import pytest

from upf_fixture import write_upf

pytestmark = pytest.mark.gpu

W_SHAPE = [(0, 2.2), (0, 1.9), (1, 2.2), (1, 2.0), (2, 2.1), (2, 1.8), (3, 2.2), (3, 2.0)]
W_ENERGY = [3.0, -1.2, 1.5, -0.6, 0.8, -0.3, 0.4, 0.2]


def _system(qm, case):
    if case == "C4 40^3 real":
        return qm.system.config(4, nbands=4, grid=40)
    return qm.system.config(2, nbands=4, grid=40)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("case", ["C4 40^3 real", "C2 40^3 complex 4k"])
def test_upf_tables_hpsi_and_forces(qm, orc, torch_cuda, tmp_path, case, prec):
    torch = torch_cuda
    print(f"case UPF {case} {prec}", flush=True)
    pp = qm.upf.Pseudopotential(write_upf(tmp_path / "w_like.upf", W_SHAPE, W_ENERGY))
    s = qm.upf.attach(_system(qm, case), pp)
    pp.close()
    assert int(s.nproj.max()) == 32 and s.table.size > 0
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    dtype = torch.float64 if prec == "f64" else torch.float32
    src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
    x = torch.from_numpy(psi).to(device="cuda", dtype=dtype)
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    for ik in range(s.kpoints.shape[0]):
        want = orc.hpsi(s, src, k=s.kpoints[ik])
        y = torch.full_like(x, float("nan"))
        H.apply(y, x, k_index=ik)
        torch.cuda.synchronize()
        got = y.double().cpu().numpy()
        err = float(np.abs(got - want).max() / np.abs(want).max())
        assert err <= (1e-12 if prec == "f64" else 2e-6), (case, ik, err)
        e, f = H.nonlocal_forces(x, occ, k_index=ik)
        e0, f0 = orc.nonlocal_forces(s, src, occ, k=s.kpoints[ik])
        tol = 1e-12 if prec == "f64" else 1e-5
        assert abs(e - e0) <= tol * abs(e0), (case, ik, e, e0)
        assert np.abs(f - f0).max() <= tol * np.abs(f0).max(), (case, ik)
    H.close()
