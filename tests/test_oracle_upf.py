"""The UPF oracle (oracle/upf_ref.py) pinned by what the mathematics fixes, and
the library's host-side UPF loader and table builder (qm_upf_*) checked
against it.  No GPU: the qm_upf_* entry points are host code."""
import os

import numpy as np
import pytest

import oracle as orc
import upf_ref as U
from upf_fixture import write_upf, gaussian_beta

W_UPF = "/root/reference/PP/nc-sr-05_pbe_standard_upf/W.upf"


# ---- oracle pins -------------------------------------------------------------

@pytest.mark.parametrize("parity,coef", [(1, (0.3, 0.0, -1.2, 0.0, 0.7)), (-1, (0.0, 1.1, 0.0, -0.4, 0.0))])
def test_uniform_value_reproduces_quartics(parity, coef):
    """The quintic with five-point derivative estimates reproduces any quartic;
    mirrored through r = 0 with the function's own parity it does so down to
    r = 0."""
    dr, n = 0.01, 300
    r = np.arange(n) * dr
    poly = np.polynomial.Polynomial(coef)
    x = np.random.default_rng(0).uniform(0.0, (n - 4) * dr, 200)
    got = U.uniform_value(poly(r), 0.0, dr, parity, x)
    assert np.abs(got - poly(x)).max() <= 1e-12 * np.abs(poly(r)).max()


def test_uniform_value_has_teeth():
    """a sextic is not reproduced, and the wrong mirror parity is seen near 0"""
    dr, n = 0.05, 100
    r = np.arange(n) * dr
    x = np.linspace(0.01, 2.0, 50)
    assert np.abs(U.uniform_value(r**6, 0.0, dr, 1, x) - x**6).max() > 1e-8
    assert np.abs(U.uniform_value(r**2, 0.0, dr, -1, x) - x**2).max() > 1e-6


def test_uniform_value_zero_past_end():
    dr, n = 0.01, 100
    vals = np.ones(n)
    assert np.all(U.uniform_value(vals, 0.0, dr, 1, np.array([(n + 2) * dr + 0.001, 5.0])) == 0.0)


def test_xi_value_cells():
    """GetValueByXi: interior cells reproduce quartics in the cell index, the
    second and third-to-last cells quadratics, the first cell linears; zero
    from the second-to-last cell on.  d/dxi is the exact derivative there."""
    n, xi_min, dxi = 60, -2.0, 0.1
    t = np.arange(n)
    rng = np.random.default_rng(1)
    for deg, cells in ((4, range(2, n - 3)), (2, (1, n - 3)), (1, (0,))):
        p = np.polynomial.Polynomial(rng.normal(size=deg + 1))
        for i in cells:
            u = i + rng.uniform(0, 1, 5)
            v, dv = U.xi_value(p(t), xi_min, dxi, xi_min + u * dxi)
            assert np.allclose(v, p(u), rtol=0, atol=1e-10 * np.abs(p(t)).max()), (deg, i)
            assert np.allclose(dv, p.deriv()(u) / dxi, rtol=0, atol=1e-8 * np.abs(p(t)).max() / dxi), (deg, i)
    v, _ = U.xi_value(np.ones(n), xi_min, dxi, xi_min + np.array([n - 2 + 0.5, n + 3.0]) * dxi)
    assert np.all(v == 0.0)
    # cubic cells are not quartic-exact: the branch is real
    p = np.polynomial.Polynomial([0, 0, 0, 0, 1.0])
    v, _ = U.xi_value(p(t), xi_min, dxi, xi_min + 1.5 * dxi)
    assert abs(v[0] - p(1.5)) > 1e-3


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_chain_converges_to_closed_form(l, tmp_path):
    """beta(r) = r^(l+1) exp(-r^2 / 2 s^2) through both resampling stages gives
    f(r) = r^l exp(-r^2 / 2 s^2) and its derivative to interpolation accuracy,
    and 0 beyond the cutoff."""
    path = write_upf(tmp_path / "g.upf", [(l, 3.0)], energies=[1.5])
    rad = U.Radial(U.parse(path))
    r = np.linspace(0.05, 2.95, 300)
    f, df = rad.value(0, r)
    ex, dex = gaussian_beta(l, r)
    assert np.abs(f - ex).max() <= 2e-7
    assert np.abs(df - dex).max() <= 2e-5
    f, _ = rad.value(0, np.array([3.01, 4.0]))
    assert np.all(f == 0.0)


def test_tables_gradient_matches_differences(tmp_path):
    """the gradient tables from the interpolants' derivatives agree with
    central differences of the table function"""
    path = write_upf(tmp_path / "g.upf", [(0, 2.5), (1, 2.5), (2, 2.5), (3, 2.5)], energies=[1.0, -0.5, 0.3, 0.2])
    rad = U.Radial(U.parse(path))
    ylm = np.vectorize(orc.ylm)
    d = np.random.default_rng(2).uniform(-1.4, 1.4, size=(60, 3))
    d = d[np.linalg.norm(d, axis=1) > 0.2]
    eps = 1e-5
    num = np.stack([(U.table_values(rad, ylm, d + eps * e) - U.table_values(rad, ylm, d - eps * e)) / (2 * eps)
                    for e in np.eye(3)])
    ana = np.stack(_oracle_gradient(rad, ylm, d))
    assert np.abs(num - ana).max() <= 1e-6 * np.abs(ana).max()


def _oracle_gradient(rad, ylm, d):
    """grad p = f'(r) r_hat Y + f grad Y, with grad Y by differences of the
    closed-form harmonics (independent of the library's solid-harmonic
    gradients)"""
    rr = np.linalg.norm(d, axis=1)
    out = [[], [], []]
    eps = 1e-6
    for k, l in enumerate(rad.l):
        f, df = rad.value(k, rr)
        for m in range(-l, l + 1):
            Y = ylm(l, m, d[:, 0], d[:, 1], d[:, 2])
            for c in range(3):
                e = np.eye(3)[c] * eps
                dY = (ylm(l, m, *(d + e).T) - ylm(l, m, *(d - e).T)) / (2 * eps)
                out[c].append(df * d[:, c] / rr * Y + f * dY)
    return [np.stack(o, axis=1) for o in out]


@pytest.mark.skipif(not os.path.exists(W_UPF), reason="the reference's W pseudopotential is not present")
def test_w_upf_header_and_count():
    u = U.parse(W_UPF)
    assert u["nradial"] == 8
    assert u["l"] == [0, 0, 1, 1, 2, 2, 3, 3]
    assert sum(2 * l + 1 for l in u["l"]) == 32
    assert u["z_valence"] == 14.0
    assert u["r"][0] == 0.0 and np.allclose(np.diff(u["r"]), 0.01)


# ---- library (host code) against the oracle ----------------------------------


@pytest.mark.parametrize("which", ["synthetic", "W"])
def test_library_matches_oracle(qm, which, tmp_path):
    if which == "W":
        if not os.path.exists(W_UPF):
            pytest.skip("the reference's W pseudopotential is not present")
        path = W_UPF
    else:
        path = write_upf(tmp_path / "g.upf", [(0, 2.2), (0, 1.8), (1, 2.2), (2, 2.0), (3, 2.2)],
                         energies=[2.0, -1.0, 0.5, 0.25, -0.3])
    ref = U.parse(path)
    rad = U.Radial(ref)
    pp = qm.upf.Pseudopotential(path)
    try:
        assert pp.nradial == ref["nradial"]
        assert pp.nproj == sum(2 * l + 1 for l in ref["l"])
        assert pp.z_valence == ref["z_valence"]
        for k in range(ref["nradial"]):
            l, rc, e = pp.radial(k)
            assert (l, rc) == (ref["l"][k], ref["cutoff"][k])
            assert e == ref["energy"][k]
        r = np.linspace(0.0, rad.rc_max + 0.3, 777)
        for k in range(ref["nradial"]):
            f, df = pp.eval(k, r)
            fo, dfo = rad.value(k, r)
            scale = np.abs(fo).max()
            assert np.abs(f - fo).max() <= 1e-12 * scale, k
            assert np.abs(df - dfo).max() <= 1e-10 * np.abs(dfo).max(), k
        h = np.array([0.21, 0.2, 0.19])
        shape = (40, 42, 44)
        positions = np.array([[1.03, 2.51, 3.3], [5.0, 0.2, 7.9]])
        t = pp.tables(positions, shape, h)
        ylm = np.vectorize(orc.ylm)
        for a, (vox, tab) in enumerate(U.tables(rad, ylm, positions, h)):
            v0, v1 = t["vox_off"][a], t["vox_off"][a + 1]
            assert np.array_equal(t["vox_xyz"][v0:v1], vox)
            M = int(t["nproj"][a])
            got = t["table"][t["tab_off"][a]:t["tab_off"][a + 1]].reshape(M, -1).T
            assert np.abs(got - tab).max() <= 1e-12 * np.abs(tab).max()
            d = vox * h - positions[a]
            ana = _oracle_gradient(rad, ylm, d)
            for c in range(3):
                gd = t["dtable"][c, t["tab_off"][a]:t["tab_off"][a + 1]].reshape(M, -1).T
                assert np.abs(gd - ana[c]).max() <= 2e-6 * np.abs(ana[c]).max(), c
            e = t["energy"][t["proj_off"][a]:t["proj_off"][a + 1]]
            assert np.array_equal(e, np.repeat(ref["energy"], [2 * l + 1 for l in ref["l"]]))
    finally:
        pp.close()


def test_library_rejects_bad_files(qm, tmp_path):
    bad = tmp_path / "bad.upf"
    bad.write_text("<UPF version=\"2.0.1\"><PP_INFO/></UPF>")
    with pytest.raises(RuntimeError, match="PP_HEADER"):
        qm.upf.Pseudopotential(str(bad))
    with pytest.raises(RuntimeError, match="cannot open"):
        qm.upf.Pseudopotential(str(tmp_path / "missing.upf"))
    big = write_upf(tmp_path / "big.upf", [(3, 2.0)] * 5, energies=[1.0] * 5)  # 35 projectors
    with pytest.raises(RuntimeError, match="32"):
        qm.upf.Pseudopotential(big)
