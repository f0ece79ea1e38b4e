"""The library's host-side SCF step (qm_occupancy) against the oracle; no GPU."""
import numpy as np
import pytest

import scf_ref as S


def test_occupancy_zero_matches_oracle(qm):
    rng = np.random.default_rng(0)
    eigs = np.sort(rng.uniform(-1, 1, (3, 8)), axis=1)
    for nelec in (0.0, 3.0, 7.5, 16.0, 20.0):
        occ, mu = qm.scf.occupancy(eigs, [1, 2, 1], nelec, 0.0)
        assert np.array_equal(occ, S.occupancy_zero(nelec, 8, [1, 2, 1]))
        assert np.isnan(mu)


@pytest.mark.parametrize("kT", [1e-3, 0.01, 0.1])
def test_occupancy_fermi_matches_oracle(qm, kT):
    rng = np.random.default_rng(1)
    eigs = np.sort(rng.uniform(-1, 1, (4, 12)), axis=1)
    w = [1, 3, 3, 1]
    occ, mu = qm.scf.occupancy(eigs, w, 13.0, kT)
    occ_o, mu_o = S.occupancy_fermi(eigs, w, 13.0, kT)
    assert abs(mu - mu_o) <= 1e-12
    assert np.abs(occ - occ_o).max() <= 1e-12
    assert abs(occ.sum() - 13.0) < 1e-6


def test_occupancy_rejects_bad_input(qm):
    with pytest.raises(qm._lib.QMError):
        qm.scf.occupancy(np.zeros((1, 3)), [-1.0], 2.0)
