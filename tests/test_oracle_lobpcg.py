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
