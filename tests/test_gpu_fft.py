"""Spectral (FFT) kinetic term on the GPU (qm_set_kinetic(QM_KINETIC_FFT),
rocFFT transforms straight from the SoA band layout) against the FP64 numpy
oracle (oracle.fft_kinetic, pinned to the plane-wave eigenvalues
|G + k|^2 / 2 in tests/test_oracle.py).

Tolerances: FP64 <= 1e-12 relative (max-norm of the difference over the
max-norm of the oracle result); FP32 <= 2e-6 relative."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _odd(qm):
    s = qm.system.System("odd 15x9x21", 15, 9, 21, (3.75, 2.25, 5.25), np.array([[1.0, 1.0, 2.0]]),
                         np.array([2], np.int32), np.array([1.2]), 3, 2, np.array([[0.4, -0.3, 0.2]]), 23)
    return s.build()


def _odd_real(qm):
    s = qm.system.System("odd real 15x9x21", 15, 9, 21, (3.75, 2.25, 5.25), np.array([[1.0, 1.0, 2.0]]),
                         np.array([2], np.int32), np.array([1.2]), 3, 1, np.zeros((1, 3)), 24)
    return s.build()


CASES = [
    ("C0 32^3 real", lambda qm: qm.system.config(0, nbands=6)),
    ("C2 48^3 complex 4k", lambda qm: qm.system.config(2, nbands=3, grid=48)),
    ("ragged 36x20x28 k", lambda qm: qm.system.System(
        "ragged", 36, 20, 28, (6.0, 3.5, 5.0), np.array([[1.0, 1.0, 1.0], [4.0, 2.5, 3.0]]),
        np.array([2, 1], np.int32), np.array([1.6, 1.2]), 3, 2, np.array([[0.3, -0.4, 0.2]]), 21).build()),
    ("odd 15x9x21 k", _odd),
    ("odd 15x9x21 real", _odd_real),
    ("C1 64^3 128 real bands", lambda qm: qm.system.config(1)),
]


@pytest.mark.parametrize("label,build", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("part", ["full", "local"])
def test_fft_hpsi_parity(qm, orc, torch_cuda, label, build, prec, part):
    torch = torch_cuda
    s = build(qm)
    print(f"case FFT kinetic {label} {prec} {part}", flush=True)
    H = qm.hamiltonian.Hamiltonian(s)
    H.set_kinetic("fft")
    psi = s.initial_bands()
    dtype = torch.float64 if prec == "f64" else torch.float32
    x = torch.from_numpy(psi).to("cuda", dtype)
    sel = list(range(min(s.nbands, 4)))
    rows = [b * s.ncomp + c for b in sel for c in range(s.ncomp)]
    src = psi[rows] if prec == "f64" else psi[rows].astype(np.float32).astype(np.float64)
    for ik in range(s.kpoints.shape[0]):
        y = torch.full_like(x, float("nan"))
        H.apply(y, x, k_index=ik, part=part)
        torch.cuda.synchronize()
        got = y[rows].double().cpu().numpy()
        assert np.isfinite(got).all()
        want = orc.hpsi_fft(s, src, k=s.kpoints[ik], nonlocal_=(part == "full"))
        err = _rel(got, want)
        print(f"  k {ik}: rel err {err:.2e}")
        assert err <= (1e-12 if prec == "f64" else 2e-6), (label, ik, err)
    H.close()


def test_fft_lobpcg_matches_dense(qm, orc, torch_cuda):
    """FP64 LOBPCG with the spectral kinetic term on a 10^3 complex k-point
    system: converged eigenvalues against LAPACK on the dense spectral H built
    column by column from the oracle (<= 1e-10 Ha)."""
    torch = torch_cuda
    s = qm.system.small_system(n=10, L=4.0, natoms=2, rc=1.4, nbands=6, ncomp=2, k_red=(0.25, 0.0, 0.5), seed=13)
    k = s.kpoints[0]
    n = s.ng
    eye = np.eye(n)
    cols = np.empty((n, n), dtype=complex)
    for j0 in range(0, n, 100):
        blk = eye[j0:j0 + 100]
        soa = np.zeros((2 * blk.shape[0], n))
        soa[0::2] = blk
        out = orc.hpsi_fft(s, soa, k=k, ncomp=2)
        cols[:, j0:j0 + blk.shape[0]] = (out[0::2] + 1j * out[1::2]).T
    Hd = 0.5 * (cols + cols.conj().T)
    ref = np.linalg.eigvalsh(Hd)[:s.nbands]
    H = qm.hamiltonian.Hamiltonian(s)
    H.set_kinetic("fft")
    L = qm.lobpcg.LOBPCG(H, s.nbands, precision=64)
    x = torch.from_numpy(s.initial_bands()).to("cuda", torch.float64)
    print("case FFT kinetic LOBPCG 10^3 complex k, 6 bands", flush=True)
    for c in range(12):
        ev, rn = L.solve(x, 20, first_time=(c == 0))
    err = qm.lobpcg.cluster_mean_errors(ev, ref)
    print("cluster-mean |error| (Ha):", err)
    assert err.max() <= 1e-10
    L.close()
    H.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fft_td_kinetic_and_taylor(qm, orc, torch_cuda, prec):
    """The FFT kinetic propagator step and the order-4 Taylor propagator with the
    spectral kinetic term, against the oracle."""
    torch = torch_cuda
    s = qm.system.config(2, nbands=3, grid=40)
    ik = 2
    H = qm.hamiltonian.Hamiltonian(s)
    H.set_kinetic("fft")
    psi = s.initial_bands()
    dtype = torch.float64 if prec == "f64" else torch.float32
    src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
    tol = 1e-12 if prec == "f64" else 2e-6
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print(f"case FFT kinetic step + Taylor-4 {prec}", flush=True)
    x = torch.from_numpy(psi).to("cuda", dtype)
    fn = H.lib.qm_td_kinetic_fft_f64 if prec == "f64" else H.lib.qm_td_kinetic_fft_f32
    qm._lib.check(fn(H.ctx, ctypes.c_void_p(x.data_ptr()), s.nbands, ik, 0.05, st))
    torch.cuda.synchronize()
    assert _rel(x.double().cpu().numpy(), orc.td_kinetic_fft(s, src, 0.05, k=s.kpoints[ik])) <= tol
    x = torch.from_numpy(psi).to("cuda", dtype)
    fn = H.lib.qm_td_taylor_f64 if prec == "f64" else H.lib.qm_td_taylor_f32
    qm._lib.check(fn(H.ctx, ctypes.c_void_p(x.data_ptr()), s.nbands, ik, 0.01, 4, 3, st))
    torch.cuda.synchronize()
    want = orc.td_taylor(s, src, 0.01, order=4, op="full", k=s.kpoints[ik], kinetic="fft")
    assert _rel(x.double().cpu().numpy(), want) <= tol
    H.close()
