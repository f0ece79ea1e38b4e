"""Pins the FP64 LOBPCG oracle: converged eigenvalues against a dense
LAPACK eigendecomposition of the dense H(k) (independent construction in
test_oracle.dense_hamiltonian), Ritz-vector properties, and the reference's
rollback branch."""
import numpy as np
import pytest

from test_oracle import dense_hamiltonian


@pytest.mark.parametrize("ncomp,kred", [(1, (0, 0, 0)), (2, (0.25, -0.1, 0.3))])
def test_oracle_lobpcg_converges_to_dense_eigh(qm, orc, ncomp, kred):
    import lobpcg_ref
    s = qm.system.small_system(n=10, L=4.0, natoms=2, rc=1.5, nbands=5, ncomp=ncomp, k_red=kred, seed=3)
    k = s.kpoints[0]
    w = np.linalg.eigvalsh(dense_hamiltonian(s, k))
    L = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a, k=k), ncomp, s.dV)
    X = s.initial_bands()
    for c in range(15):
        X, ev, rn, it = L.solve(X, 20, first_time=(c == 0))
    assert np.abs(ev[:4] - w[:4]).max() < 1e-9
    # Ritz vectors: dV-orthonormal, Rayleigh quotients equal the eigenvalues
    z = lobpcg_ref.soa_to_cplx(X, ncomp)
    G = np.conj(z) @ z.T * s.dV
    assert np.abs(G - np.eye(5)).max() < 1e-8
    hz = lobpcg_ref.soa_to_cplx(orc.hpsi(s, X, k=k), ncomp)
    rq = np.real(np.sum(np.conj(z) * hz, axis=1)) * s.dV
    assert np.abs(rq - ev).max() < 1e-8


def test_oracle_lobpcg_rollback_branch(qm, orc):
    """A Rayleigh-Ritz step whose update loses normalisation by more than 1e-3
    (here: a corrupted overlap block, so C^H S C = I no longer holds for the
    actual vectors) takes the reference's escape path: P = R, the old X
    renormalised, loop left early."""
    import lobpcg_ref

    class Corrupt(lobpcg_ref.LOBPCGRef):
        calls = 0

        def _gram(self, a, b):
            g = np.conj(a) @ b.T
            Corrupt.calls += 1
            if Corrupt.calls == 20:  # a block of the 4th iteration
                g = g + 0.05
            return g

    s = qm.system.small_system(n=8, L=4.0, natoms=1, rc=1.5, nbands=3, seed=4)
    L = Corrupt(lambda a: orc.hpsi(s, a), 1, s.dV)
    X, ev, rn, it = L.solve(s.initial_bands(), 10, first_time=True)
    assert it < 10  # escaped
    assert L.P is not None
    nrm = (X * X).sum(axis=1) * s.dV
    assert np.abs(nrm - 1).max() < 1e-12  # the kept X is renormalised


@pytest.mark.parametrize("ncomp,kred", [(1, (0, 0, 0)), (2, (0.25, -0.1, 0.3))])
def test_oracle_locking_compressed_spectrum(qm, orc, ncomp, kred):
    """Locking with a threshold every residual exceeds fixes N/2 bands on the
    first call (the cap).  The fixed bands come back unchanged, and the active
    bands stay orthogonal to them.  The active bands converge to the lowest
    eigenvalues of H compressed onto the orthogonal complement of the fixed
    bands (dense LAPACK on a null-space basis).  A wrong sign or a missing
    conjugate in V - X_f (X_f^H V) fails both checks.

    The fixed bands here are random, not eigenvectors.  The reference never
    re-orthogonalises X itself, so rounding drift along the fixed bands,
    which lowers the energy, grows over hundreds of iterations.  The checks
    therefore stop at 80 iterations, where the drift is below 1e-9."""
    import lobpcg_ref
    import scipy.linalg
    s = qm.system.small_system(n=10, L=4.0, natoms=2, rc=1.5, nbands=6, ncomp=ncomp, k_red=kred, seed=3)
    k = s.kpoints[0]
    Hd = dense_hamiltonian(s, k)
    L = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a, k=k), ncomp, s.dV, residual_threshold=1.0e6)
    X0 = s.initial_bands()
    X = X0
    ev_in = np.arange(6, dtype=np.float64) + 100.0
    for c in range(4):
        X, ev, rn, it = L.solve(X, 20, first_time=(c == 0), eigen_values=ev_in)
    f = L.num_fix
    assert f == 3
    assert np.array_equal(ev[:f], ev_in[:f])  # fixed entries left as passed
    assert np.all(np.isnan(rn[:f]))
    assert np.abs(X[:f * ncomp] - X0[:f * ncomp]).max() <= 1e-14 * np.abs(X0).max()
    z = lobpcg_ref.soa_to_cplx(X, ncomp)
    zf, za = z[:f], z[f:]
    assert np.abs(np.conj(zf) @ za.T * s.dV).max() < 1e-9
    B = scipy.linalg.null_space(np.conj(zf))
    wc = np.linalg.eigvalsh(np.conj(B).T @ Hd @ B)
    assert np.abs(ev[f:] - wc[:3]).max() < 1e-7
    # the unconstrained spectrum is a different answer (the check has teeth)
    w = np.linalg.eigvalsh(Hd)
    assert np.abs(w[:2] - wc[:2]).max() > 1e-3


def test_oracle_locking_selection_rule(qm, orc):
    """num_fix selection (lobpcg_z_multi_mpi.h:1962-1986): threshold 0 fixes
    nothing; after convergence, a threshold above every residual fixes N/2 (the
    cap); the count never decreases; locked converged bands leave the remaining
    eigenvalues on the dense spectrum."""
    import lobpcg_ref
    s = qm.system.small_system(n=10, L=4.0, natoms=2, rc=1.5, nbands=6, seed=3)
    w = np.linalg.eigvalsh(dense_hamiltonian(s, s.kpoints[0]))
    L = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a), 1, s.dV, residual_threshold=0.0)
    X = s.initial_bands()
    for c in range(10):
        X, ev, rn, it = L.solve(X, 20, first_time=(c == 0))
        assert L.num_fix == 0
    assert np.abs(ev[:5] - w[:5]).max() < 1e-9
    L.residual_threshold = 1.0e-4  # above every converged residual
    X, ev2, rn2, it = L.solve(X, 10, first_time=False, eigen_values=ev)
    assert L.num_fix == 3
    assert np.array_equal(ev2[:3], ev[:3])
    assert np.abs(ev2[3:5] - w[3:5]).max() < 1e-9
    L.residual_threshold = 0.0  # the count does not decrease
    X, ev3, rn3, it = L.solve(X, 4, first_time=False, eigen_values=ev2)
    assert L.num_fix == 3
    assert np.abs(ev3[3:5] - w[3:5]).max() < 1e-9


@pytest.mark.parametrize("ncomp,kred", [(1, (0, 0, 0)), (2, (0.25, -0.1, 0.3))])
def test_build_variant_follows_reference_iteration(qm, orc, ncomp, kred):
    """The build's LOBPCG variant (R orthonormalised against X, P normalised,
    X^H R and R^H P formed from the vectors) against a literal restatement of
    the reference's iteration (R as computed, P unnormalised, X^H R and R^H P
    from the recurrence: lobpcg_reference_variant).  Both do Rayleigh-Ritz on
    the same subspaces, so their Ritz values agree iteration by iteration, to
    rounding, until the lowest bands converge and the reference form's nearly
    dependent R columns dominate its rounding.  A wrong sign or a transposed
    block in either recurrence moves the next iteration's Ritz values far
    outside this bound."""
    import lobpcg_ref
    s = qm.system.small_system(n=10, L=4.0, natoms=2, rc=1.5, nbands=5, ncomp=ncomp, k_red=kred, seed=3)
    k = s.kpoints[0]
    apply = lambda a: orc.hpsi(s, a, k=k)  # noqa: E731
    X0 = s.initial_bands()
    L = lobpcg_ref.LOBPCGRef(apply, ncomp, s.dV)
    _, _, _, _, hist_build = L.solve(X0, 8, first_time=True, return_history=True)
    hist_ref = lobpcg_ref.lobpcg_reference_variant(apply, X0, ncomp, s.dV, 8)
    assert len(hist_build) == len(hist_ref) == 8
    dev = [float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(hist_build, hist_ref)]
    print("relative Ritz-value difference per iteration:", " ".join(f"{d:.1e}" for d in dev))
    assert max(dev) < 1e-11, dev
    # the iteration makes progress: the lowest Ritz value drops by more than the bound
    assert hist_ref[0][0] - hist_ref[-1][0] > 1e-3
