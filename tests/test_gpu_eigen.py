"""GPU parity of the FP64 generalized symmetric eigensolvers qm_sygvd_f64 and
qm_hegvd_z64 (own Cholesky and tridiagonal reduction, rocSOLVER stedc, own
back-transformation of the wanted eigenvectors; the Cholesky route of
lapack_DSYGVD / lapack_ZHEGVD_emu) against LAPACK through scipy.

Tolerances (FP64): eigenvalues within 1e-11 * ||A||; residual
||A Z - B Z diag(w)|| and B-orthonormality ||Z^T B Z - I|| within 1e-11
(scaled), including spectra with exactly degenerate clusters.
"""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu


def _problem(n, seed, clusters=False):
    rng = np.random.default_rng(seed)
    if clusters:
        # eigenvalues with exact multiplicities and near-degenerate pairs
        lam = np.sort(np.concatenate([np.repeat(rng.uniform(-5, 5, n // 6), 3),
                                      rng.uniform(-5, 5, n - 3 * (n // 6))]))
        lam[1::7] = lam[0::7][: len(lam[1::7])] + 1e-9
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (Q * lam) @ Q.T
    else:
        M = rng.standard_normal((n, n))
        A = M + M.T
    G = rng.standard_normal((n, n)) * (0.3 / np.sqrt(n))
    B = np.eye(n) + G @ G.T
    return 0.5 * (A + A.T), 0.5 * (B + B.T)


@pytest.fixture(scope="module")
def H(qm, torch_cuda):
    s = qm.system.small_system(n=8, L=4.0, natoms=1, nbands=2)
    h = qm.hamiltonian.Hamiltonian(s)
    yield h
    h.close()


@pytest.fixture
def route(qm, request):
    qm.hamiltonian.subspace_force(request.param)
    yield request.param
    qm.hamiltonian.subspace_force(-1)


# route -1: the cooperative multi-workgroup tridiagonal reduction wherever it
# fits (every n here); route 0: the single-workgroup reduction (n <= 512) or
# rocSOLVER syevd
@pytest.mark.parametrize("route", [-1, 0], indirect=True, ids=["coop", "single"])
@pytest.mark.parametrize("n", [1, 2, 5, 72, 88, 89, 121, 128, 129, 144, 200, 384, 512, 521, 768, 1024, 1280])
@pytest.mark.parametrize("clusters", [False, True])
def test_sygvd_against_lapack(H, torch_cuda, route, n, clusters):
    torch = torch_cuda
    if clusters and n < 12:
        pytest.skip("cluster construction needs n >= 12")
    print(f"sygvd n {n} clusters {clusters} route {route}", flush=True)
    A, B = _problem(n, seed=n + (1000 if clusters else 0), clusters=clusters)
    nev = max(1, n // 3)
    w, Z, U = H.sygvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), nev)
    torch.cuda.synchronize()
    w, Z, U = w.cpu().numpy(), Z.cpu().numpy(), U.cpu().numpy()
    w_ref = scipy.linalg.eigh(A, B, eigvals_only=True)
    scale = max(1.0, np.abs(A).max() * n)
    assert np.abs(w - w_ref).max() <= 1e-11 * scale
    assert np.all(np.diff(w) >= -1e-12 * scale)
    res = A @ Z - (B @ Z) * w[:nev]
    assert np.abs(res).max() <= 1e-11 * scale
    assert np.abs(Z.T @ B @ Z - np.eye(nev)).max() <= 1e-11 * n
    Uu = np.triu(U)
    assert np.abs(Uu.T @ Uu - B).max() <= 1e-12 * n


# the back-transformation runs one workgroup per wanted eigenvector: one
# column (a single workgroup) and all n columns, on the cooperative route
@pytest.mark.parametrize("route", [-1], indirect=True, ids=["coop"])
@pytest.mark.parametrize("n", [3, 129, 384, 1040])
@pytest.mark.parametrize("which", ["one", "all"])
def test_sygvd_hegvd_nev_extremes(H, torch_cuda, route, n, which):
    torch = torch_cuda
    nev = 1 if which == "one" else n
    print(f"sygvd/hegvd n {n} nev {nev}", flush=True)
    A, B = _problem(n, seed=3 * n + nev)
    w, Z, _ = H.sygvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), nev)
    torch.cuda.synchronize()
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    w_ref = scipy.linalg.eigh(A, B, eigvals_only=True)
    scale = max(1.0, np.abs(A).max() * n)
    assert np.abs(w - w_ref).max() <= 1e-11 * scale
    assert np.abs(A @ Z - (B @ Z) * w[:nev]).max() <= 1e-11 * scale
    assert np.abs(Z.T @ B @ Z - np.eye(nev)).max() <= 1e-11 * n
    Ac, Bc = _zproblem(n, seed=5 * n + nev)
    w, Z, _ = H.hegvd(torch.from_numpy(Ac).cuda(), torch.from_numpy(Bc).cuda(), nev)
    torch.cuda.synchronize()
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    w_ref = scipy.linalg.eigh(Ac, Bc, eigvals_only=True)
    scale = max(1.0, np.abs(Ac).max() * n)
    assert np.abs(w - w_ref).max() <= 1e-11 * scale
    assert np.abs(Ac @ Z - (Bc @ Z) * w[:nev]).max() <= 1e-11 * scale
    assert np.abs(Z.conj().T @ Bc @ Z - np.eye(nev)).max() <= 1e-11 * n


# A grid-barrier wait that runs out ends the cooperative kernel and fails the
# solve with info -77 (the bound forced to 0 polls from the host, so every
# wait not over at its first poll runs out; every kernel still ends).  Route
# -1: the cooperative Cholesky fails first; route 1: the single-workgroup
# Cholesky, then the cooperative reduction fails, and its undefined
# tridiagonal is zeroed before stedc reads it.  The default bound then solves
# the same problems to LAPACK parity.
@pytest.mark.parametrize("route", [-1, 1], indirect=True, ids=["coop", "coop-reduction"])
def test_coop_barrier_timeout_is_reported(qm, H, torch_cuda, route):
    torch = torch_cuda
    n, nev = 384, 128
    A, B = _problem(n, seed=11)
    Ac, Bc = _zproblem(n, seed=12)
    print(f"barrier timeout route {route}", flush=True)
    try:
        qm.hamiltonian.coop_spin_limit(0)
        with pytest.raises(RuntimeError, match="-77"):
            H.sygvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), nev)
        with pytest.raises(RuntimeError, match="-77"):
            H.hegvd(torch.from_numpy(Ac).cuda(), torch.from_numpy(Bc).cuda(), nev)
    finally:
        qm.hamiltonian.coop_spin_limit(None)
    w, Z, _ = H.sygvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), nev)
    torch.cuda.synchronize()
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    scale = max(1.0, np.abs(A).max() * n)
    assert np.abs(w - scipy.linalg.eigh(A, B, eigvals_only=True)).max() <= 1e-11 * scale
    assert np.abs(A @ Z - (B @ Z) * w[:nev]).max() <= 1e-11 * scale
    w, Z, _ = H.hegvd(torch.from_numpy(Ac).cuda(), torch.from_numpy(Bc).cuda(), nev)
    torch.cuda.synchronize()
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    scale = max(1.0, np.abs(Ac).max() * n)
    assert np.abs(w - scipy.linalg.eigh(Ac, Bc, eigvals_only=True)).max() <= 1e-11 * scale
    assert np.abs(Ac @ Z - (Bc @ Z) * w[:nev]).max() <= 1e-11 * scale


def test_sygvd_not_positive_definite(qm, H, torch_cuda):
    torch = torch_cuda
    n = 40
    A, B = _problem(n, seed=3)
    B[5, 5] = -1.0
    with pytest.raises(qm._lib.QMError):
        H.sygvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), 4)


@pytest.mark.parametrize("n,bad", [(40, 5), (100, 33), (384, 50), (384, 200), (384, 383), (200, 191)])
def test_sygvd_reports_lapack_failing_minor(qm, H, torch_cuda, n, bad):
    """The Cholesky of B stops at the first leading minor that is not
    positive definite and reports its order, as LAPACK's dpotrf does (checked
    against scipy's dpotrf on the same matrix)."""
    torch = torch_cuda
    A, B = _problem(n, seed=n + bad)
    B[bad, bad] = -2.0
    _, info = scipy.linalg.lapack.dpotrf(B, lower=0)
    assert info > 0
    with pytest.raises(qm._lib.QMError) as e:
        H.sygvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), 4)
    assert f"leading minor {info})" in str(e.value)


def _zproblem(n, seed, clusters=False):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if clusters:
        lam = np.sort(np.concatenate([np.repeat(rng.uniform(-5, 5, n // 6), 3),
                                      rng.uniform(-5, 5, n - 3 * (n // 6))]))
        Q, _ = np.linalg.qr(M)
        A = (Q * lam) @ Q.conj().T
    else:
        A = M + M.conj().T
    G = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * (0.3 / np.sqrt(n))
    B = np.eye(n) + G @ G.conj().T
    return 0.5 * (A + A.conj().T), 0.5 * (B + B.conj().T)


# route -1: the cooperative reduction (complex, every n here); 0: rocSOLVER heevd
@pytest.mark.parametrize("route", [-1, 0], indirect=True, ids=["coop", "heevd"])
@pytest.mark.parametrize("n", [1, 2, 5, 24, 72, 129, 192, 256, 384, 768, 1040])
@pytest.mark.parametrize("clusters", [False, True])
def test_hegvd_against_lapack(H, torch_cuda, route, n, clusters):
    torch = torch_cuda
    if clusters and n < 12:
        pytest.skip("cluster construction needs n >= 12")
    print(f"hegvd n {n} clusters {clusters} route {route}", flush=True)
    A, B = _zproblem(n, seed=7 * n + (1000 if clusters else 0), clusters=clusters)
    nev = max(1, n // 3)
    w, Z, U = H.hegvd(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), nev)
    torch.cuda.synchronize()
    w, Z, U = w.cpu().numpy(), Z.cpu().numpy(), U.cpu().numpy()
    w_ref = scipy.linalg.eigh(A, B, eigvals_only=True)
    scale = max(1.0, np.abs(A).max() * n)
    assert np.abs(w - w_ref).max() <= 1e-11 * scale
    res = A @ Z - (B @ Z) * w[:nev]
    assert np.abs(res).max() <= 1e-11 * scale
    assert np.abs(Z.conj().T @ B @ Z - np.eye(nev)).max() <= 1e-11 * n
    Uu = np.triu(U)
    assert np.abs(Uu.conj().T @ Uu - B).max() <= 1e-12 * n
