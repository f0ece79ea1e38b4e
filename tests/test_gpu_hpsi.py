"""GPU parity of H.psi, the band density and the nonlocal forces (HIP library
through the C ABI) against the CPU FP64 oracle.

Tolerances: FP64 path <= 1e-12 relative (max-norm of the difference over the
max-norm of the oracle result); FP32 path <= 2e-6 relative, the FP32 rounding
of a 25-point stencil whose diagonal dominates (c0 h^2 ~ 4.3).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 2e-6


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _gpu_apply(torch, H, psi, dtype, k_index=0, ncomp=None):
    x = torch.from_numpy(psi).to(device="cuda", dtype=dtype)
    y = torch.full_like(x, float("nan"))  # every output element must be written
    H.apply(y, x, k_index=k_index, ncomp=ncomp)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


CASES = [
    # (label, builder)
    ("C0 32^3 real", lambda qm: qm.system.config(0, nbands=6)),
    ("C2 48^3 complex 4k", lambda qm: qm.system.config(2, nbands=3, grid=48)),
    ("C3 80^3 uneven overlapping", lambda qm: qm.system.config(3, nbands=5, grid=80)),
    ("C4 40^3 real", lambda qm: qm.system.config(4, nbands=4, grid=40)),
    ("small 12^3 wrapped spheres k", lambda qm: qm.system.small_system(n=12, L=4.0, natoms=3, rc=2.6, nbands=3,
                                                                       ncomp=2, k_red=(0.5, 0.25, -0.125))),
    ("ragged 36x20x28", lambda qm: _ragged(qm)),
    # more columns than one nonlocal workgroup takes (projection 64, scatter 32
    # FP32 / 16 FP64), with a partial last group: 70 real, 74 complex columns
    ("C0 32^3 70 real bands", lambda qm: qm.system.config(0, nbands=70)),
    ("C2 40^3 37 complex bands 4k", lambda qm: qm.system.config(2, nbands=37, grid=40)),
    # enough projection workgroups (1578) for the 4-tile projection shape; the cases
    # above take the 1-tile shape of smaller launches
    ("C3 80^3 160 bands", lambda qm: qm.system.config(3, nbands=160, grid=80)),
    # edge cases: one band (real and complex at a k-point), no atoms (no nonlocal term),
    # s-only projectors (one per atom)
    ("C0 32^3 one real band", lambda qm: qm.system.config(0, nbands=1)),
    ("small 16^3 one complex band k", lambda qm: qm.system.small_system(n=16, L=4.0, natoms=2, rc=1.6, nbands=1,
                                                                         ncomp=2, k_red=(0.25, -0.5, 0.125))),
    ("small 16^3 no atoms k", lambda qm: qm.system.small_system(n=16, L=4.0, natoms=0, nbands=3, ncomp=2,
                                                                 k_red=(0.5, 0.0, 0.25))),
    ("small 20^3 s projectors only", lambda qm: qm.system.small_system(n=20, L=5.0, natoms=4, lmax=0, rc=1.8,
                                                                        nbands=5)),
    # 16 projectors per atom (the widest set on 16-padded tables)
    ("C0 32^3 16 projectors per atom", lambda qm: _pad16(qm.system.config(0, nbands=6))),
    ("small 12^3 k 16 projectors per atom", lambda qm: _pad16(qm.system.small_system(
        n=12, L=4.0, natoms=3, rc=2.6, nbands=3, ncomp=2, k_red=(0.5, 0.25, -0.125)))),
    # 32 projectors per atom (the C ABI's limit): l <= 3 with two radial projectors,
    # the shape of the reference's W.upf; tables padded to 32 (two MFMA row tiles)
    ("C0 32^3 W-like 32 projectors per atom", lambda qm: qm.system.wlike(0, nbands=6)),
    ("C2 48^3 W-like complex 4k", lambda qm: qm.system.wlike(2, nbands=3, grid=48)),
    ("small 12^3 k W-like wrapped spheres", lambda qm: qm.system.small_system(
        n=12, L=4.0, natoms=3, rc=2.6, nbands=3, ncomp=2, k_red=(0.5, 0.25, -0.125), lmax=3, nradial=2)),
    # 8 or 18 projectors per atom (l <= 1 or 2, two radial): 32-padded tables with
    # atoms below the padding; 160 bands take the 2-column-tile projection shape
    ("C3 80^3 two radial (8 or 18 per atom)", lambda qm: _two_radial(qm.system.config(3, nbands=5, grid=80,
                                                                                        build=False))),
    ("C3 80^3 W-like 160 bands", lambda qm: qm.system.wlike(3, nbands=160, grid=80)),
]


def _two_radial(s):
    s.nradial = 2
    return s.build()


def _pad16(s):
    """Every atom's projector set widened to 16 columns: the extra columns are
    smooth functions of the voxel coordinates times the atom's s-projector,
    with distinct energies (both signs)."""
    nat = s.natoms
    tabs, energies = [], []
    tab_off = np.zeros(nat + 1, np.int64)
    proj_off = np.zeros(nat + 1, np.int32)
    for a in range(nat):
        M = int(s.nproj[a])
        V = int(s.vox_off[a + 1] - s.vox_off[a])
        tab = s.table[s.tab_off[a]:s.tab_off[a + 1]].reshape(M, V)
        xyz = s.vox_xyz[s.vox_off[a]:s.vox_off[a + 1]].astype(np.float64)
        extra = [tab[0] * np.cos(0.37 * (w + 1) * xyz[:, 0] + 0.21 * w * xyz[:, 1] - 0.13 * xyz[:, 2])
                 for w in range(16 - M)]
        tabs.append(np.concatenate([tab, np.array(extra)]).ravel())
        e_extra = [0.75 * (-1) ** w * (1 + 0.1 * w) for w in range(16 - M)]
        energies.append(np.concatenate([s.energy[s.proj_off[a]:s.proj_off[a + 1]], e_extra]))
        tab_off[a + 1] = tab_off[a] + 16 * V
        proj_off[a + 1] = proj_off[a] + 16
    s.nproj = np.full(nat, 16, np.int32)
    s.tab_off, s.proj_off = tab_off, proj_off
    s.table = np.ascontiguousarray(np.concatenate(tabs))
    s.dtable = np.zeros((3, s.table.size))
    s.energy = np.ascontiguousarray(np.concatenate(energies))
    return s


def _ragged(qm):
    s = qm.system.System("ragged", 36, 20, 28, (6.0, 3.5, 5.0), np.array([[1.0, 1.0, 1.0], [4.0, 2.5, 3.0]]),
                         np.array([2, 1], np.int32), np.array([1.6, 1.2]), 3, 2,
                         np.array([[0.3, -0.4, 0.2]]), 21)
    return s.build()


@pytest.mark.parametrize("label,build", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hpsi_parity(qm, orc, torch_cuda, label, build, prec):
    torch = torch_cuda
    s = build(qm)
    print(f"case {label} prec {prec}", flush=True)
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    dtype = torch.float64 if prec == "f64" else torch.float32
    for ik in range(s.kpoints.shape[0]):
        want = orc.hpsi(s, psi if prec == "f64" else psi.astype(np.float32).astype(np.float64), k=s.kpoints[ik])
        got = _gpu_apply(torch, H, psi, dtype, k_index=ik)
        assert np.isfinite(got).all()
        err = _rel(got, want)
        assert err <= (F64_TOL if prec == "f64" else F32_TOL), (label, ik, err)
    H.close()


def test_hpsi_c1_full_size_subset(qm, orc, torch_cuda):
    """C1 at full size (64^3, 128 bands): every band applied on the GPU in one
    call, 8 of them checked against the oracle, plus Hermiticity over all."""
    torch = torch_cuda
    s = qm.system.config(1)
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    for prec, dtype, tol in (("f64", torch.float64, F64_TOL), ("f32", torch.float32, F32_TOL)):
        x = torch.from_numpy(psi).to(device="cuda", dtype=dtype)
        y = torch.full_like(x, float("nan"))
        H.apply(y, x)
        torch.cuda.synchronize()
        sel = [0, 1, 37, 64, 90, 125, 126, 127]
        src = psi[sel] if prec == "f64" else psi[sel].astype(np.float32).astype(np.float64)
        want = orc.hpsi(s, src)
        got = y[sel].double().cpu().numpy()
        assert _rel(got, want) <= tol, prec
        # <x_i | H x_j> symmetric over the whole block (size-independent property),
        # defect measured against the largest off-diagonal entry as below
        G = (x.double() @ y.double().T).cpu().numpy()
        scale = np.abs(G[~np.eye(G.shape[0], dtype=bool)]).max()
        defect = np.abs(G - G.T).max() / scale
        print(f"C1 {prec}: Hermiticity defect {defect:.2e}", flush=True)
        assert defect <= (1e-11 if prec == "f64" else 1e-5), (prec, defect)
    H.close()


@pytest.mark.parametrize("cfg", [2, 3], ids=["C2 96^3 256 complex 4k", "C3 160^3 512 real"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hpsi_full_size_properties(qm, orc, torch_cuda, cfg, prec):
    """C2 (96^3, 256 complex bands, 4 k-points) and C3 (160^3, 512 real bands,
    uneven spheres) at full size: the whole random block applied in one call
    per k-point, 4 bands checked against the oracle, and the block's Gram
    matrix <x_i | H x_j> Hermitian over all bands.

    Hermiticity is a size-independent property that non-Hermitian mistakes
    break (a k-gradient term without its factor i, a Bloch phase conjugated on
    one side of a projector, V_loc read at a shifted point;
    tests/test_oracle.py plants the first and the last in the oracle).  The defect
    max |G - G^H| is measured against the largest off-diagonal |G|: for random
    bands both grow like sqrt(Ng), so the ratio does not shrink with the grid,
    whereas the diagonal grows like Ng."""
    torch = torch_cuda
    s = qm.system.config(cfg)
    nb, c = s.nbands, s.ncomp
    print(f"full size C{cfg} prec {prec}: {s.nx}^3 x {nb} bands x {c}", flush=True)
    dtype = torch.float64 if prec == "f64" else torch.float32
    # measured defects: ~8e-14 FP64 (the FP64 Gram sums), 1.2-1.6e-7 FP32; planted
    # non-Hermitian mistakes give 0.15-0.24
    tol, herm_tol = (F64_TOL, 1e-11) if prec == "f64" else (F32_TOL, 1e-5)
    if cfg == 3 and prec == "f32":
        tol = 1e-6  # the C3 workload's stated tolerance (measured 3.5e-7)
    off = ~torch.eye(nb, dtype=torch.bool, device="cuda:0")
    H = qm.hamiltonian.Hamiltonian(s)
    g = torch.Generator(device="cuda:0").manual_seed(100 + cfg)
    x = torch.rand((nb * c, s.ng), device="cuda:0", dtype=dtype, generator=g) - 0.5
    y = torch.full_like(x, float("nan"))
    bands = [0, 1, nb // 2, nb - 1]
    rows = [b * c + cc for b in bands for cc in range(c)]
    src = x[rows].double().cpu().numpy()
    for ik in range(s.kpoints.shape[0]):
        y.fill_(float("nan"))
        H.apply(y, x, k_index=ik)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(y).all())
        want = orc.hpsi(s, src, k=s.kpoints[ik])
        err = _rel(y[rows].double().cpu().numpy(), want)
        assert err <= tol, (cfg, prec, ik, err)
        # Gram of the complex (or real) band block, FP64 sums over grid chunks
        G = torch.zeros((c, c, nb, nb), device="cuda:0", dtype=torch.float64)
        for a in range(0, s.ng, 1 << 20):
            xs = x[:, a:a + (1 << 20)].double().view(nb, c, -1)
            ys = y[:, a:a + (1 << 20)].double().view(nb, c, -1)
            for p in range(c):
                for q in range(c):
                    G[p, q] += xs[:, p] @ ys[:, q].T.contiguous()
        if c == 1:
            gre, gim = G[0, 0], torch.zeros_like(G[0, 0])
        else:  # x^H y = (xr yr + xi yi) + i (xr yi - xi yr)
            gre, gim = G[0, 0] + G[1, 1], G[0, 1] - G[1, 0]
        scale = max(float(gre[off].abs().max()), float(gim[off].abs().max()))
        defect = max(float((gre - gre.T).abs().max()), float((gim + gim.T).abs().max())) / scale
        print(f"  k {ik}: oracle rel err {err:.2e}, Hermiticity defect {defect:.2e}", flush=True)
        assert defect <= herm_tol, (cfg, prec, ik, defect)
    H.close()
    del x, y
    torch.cuda.empty_cache()


def test_hpsi_errors(qm, torch_cuda):
    torch = torch_cuda
    s = qm.system.config(0, nbands=2)
    H = qm.hamiltonian.Hamiltonian(s)
    x = torch.zeros((2, s.ng), device="cuda", dtype=torch.float64)
    with pytest.raises(qm._lib.QMError):
        H.apply(x, x)  # aliasing
    y = torch.zeros_like(x)
    with pytest.raises(qm._lib.QMError):
        H.apply(y, x, k_index=5)
    H.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_band_density_parity(qm, orc, torch_cuda, prec):
    torch = torch_cuda
    s = qm.system.config(2, nbands=3, grid=40)
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    occ = np.array([2.0, 1.25, 0.5])
    dtype = torch.float64 if prec == "f64" else torch.float32
    x = torch.from_numpy(psi).to(device="cuda", dtype=dtype)
    rho = H.band_density(x, occ)
    torch.cuda.synchronize()
    src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
    want = orc.band_density(s, src, occ)
    assert _rel(rho.cpu().numpy(), want) <= 1e-13
    H.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("cfg,kidx", [(4, 0), (2, 3), ("wlike4", 0), ("wlike2", 3)])
def test_nonlocal_forces_parity(qm, orc, torch_cuda, prec, cfg, kidx):
    """Forces and E_nl against the oracle; the W-like cases (32 projectors per
    atom) split each atom's projectors over two groups of 16."""
    torch = torch_cuda
    if isinstance(cfg, str):
        s = qm.system.wlike(int(cfg[-1]), nbands=4, grid=40)
    else:
        s = qm.system.config(cfg, nbands=4, grid=40)
    print(f"case forces {s.name} k {kidx} {prec}", flush=True)
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    dtype = torch.float64 if prec == "f64" else torch.float32
    x = torch.from_numpy(psi).to(device="cuda", dtype=dtype)
    e, f = H.nonlocal_forces(x, occ, k_index=kidx)
    src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
    e0, f0 = orc.nonlocal_forces(s, src, occ, k=s.kpoints[kidx])
    tol = 1e-12 if prec == "f64" else 1e-5
    assert abs(e - e0) <= tol * abs(e0)
    assert np.abs(f - f0).max() <= tol * np.abs(f0).max()
    H.close()


def _seg4(qm):
    s = qm.system.System("seg4", 16, 24, 12, (3.0, 4.5, 2.25), np.array([[1.0, 2.0, 1.0]]),
                         np.array([2], np.int32), np.array([1.2]), 2, 2, np.array([[0.5, 0.0, -0.25]]), 5)
    return s.build()


STENCIL_CASES = [
    ("C0 32^3 real (2 rows per wave)", lambda qm: qm.system.config(0, nbands=3)),
    ("64^3 real (1 row per wave)", lambda qm: qm.system.config(1, nbands=2)),
    ("C2 48^3 complex 4k (x windows)", lambda qm: qm.system.config(2, nbands=2, grid=48)),
    ("C3 80^3 (2 x windows)", lambda qm: qm.system.config(3, nbands=2, grid=80)),
    ("small 12^3 k", lambda qm: qm.system.small_system(n=12, L=4.0, natoms=3, rc=2.6, nbands=2, ncomp=2,
                                                       k_red=(0.5, 0.25, -0.125))),
    ("ragged 36x20x28 k", lambda qm: _ragged(qm)),
    ("16x24x12 k (4 rows per wave)", lambda qm: _seg4(qm)),
]


@pytest.mark.parametrize("label,build", STENCIL_CASES, ids=[c[0] for c in STENCIL_CASES])
def test_stencil_candidates_bit_identical(qm, orc, torch_cuda, label, build):
    """Every tile shape the autotuner may pick writes every output point and
    gives bit-identical results (same arithmetic per point), and the shape the
    autotuner settles on matches the oracle's kinetic + V_loc part."""
    torch = torch_cuda
    s = build(qm)
    H = qm.hamiltonian.Hamiltonian(s)
    names = H.stencil_candidates()
    assert len(names) >= 2
    psi = s.initial_bands()
    for prec, dtype, tol in (("f64", torch.float64, F64_TOL), ("f32", torch.float32, F32_TOL)):
        x = torch.from_numpy(psi).to(device="cuda", dtype=dtype)
        src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
        for ik in range(s.kpoints.shape[0]):
            print(f"case {label} prec {prec} k {ik}", flush=True)
            outs = []
            for cand in range(len(names)):
                H.stencil_force(cand)
                y = torch.full_like(x, float("nan"))
                H.apply(y, x, k_index=ik, part="local")
                outs.append(y)
            torch.cuda.synchronize()
            for cand in range(1, len(names)):
                assert torch.equal(outs[cand], outs[0]), (label, prec, ik, names[cand])
            H.stencil_force(-1)
            y = torch.full_like(x, float("nan"))
            H.apply(y, x, k_index=ik, part="local")
            torch.cuda.synchronize()
            kt = bool(np.any(np.abs(s.kpoints[ik]) >= 1e-8))
            assert 0 <= H.stencil_choice(32 if prec == "f32" else 64, s.nbands, kt=kt) < len(names)
            assert torch.equal(y, outs[0])
            want = orc.hpsi(s, src, k=s.kpoints[ik], nonlocal_=False)
            assert _rel(y.double().cpu().numpy(), want) <= tol, (label, prec, ik)
    H.close()


@pytest.mark.parametrize("label,build,dtype_name", [
    ("C0 32^3 real f32", lambda qm: qm.system.config(0, nbands=6), "float32"),
    ("C2 48^3 complex k f64", lambda qm: qm.system.config(2, nbands=3, grid=48), "float64"),
], ids=["C0-f32", "C2-k-f64"])
def test_hpsi_graph_capture_replays(qm, torch_cuda, label, build, dtype_name):
    """A full apply (stencil on the caller's stream, nonlocal projection forked
    onto the context's second stream and joined back) captured into a graph:
    every replay, with eager work and synchronisations between replays, equals
    the eager result bit for bit."""
    torch = torch_cuda
    s = build(qm)
    H = qm.hamiltonian.Hamiltonian(s)
    dtype = getattr(torch, dtype_name)
    ik = s.kpoints.shape[0] - 1
    x = torch.from_numpy(s.initial_bands()).to(device="cuda", dtype=dtype)
    want = torch.empty_like(x)
    H.apply(want, x, k_index=ik)  # eager; tunes the stencil shape before capture
    torch.cuda.synchronize()
    y = torch.full_like(x, float("nan"))
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g):
            H.apply(y, x, k_index=ik)
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(4):
        y.fill_(float("nan"))
        other = torch.empty_like(x)
        H.apply(other, x * 0.5, k_index=ik)  # eager work on the same context between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"case {label} replay {rep}", flush=True)
        assert torch.equal(y, want), (label, rep)


@pytest.mark.parametrize("label,build,ik", [
    ("C0 32^3 70 real bands", lambda qm: qm.system.config(0, nbands=70), 0),
    ("C2 48^3 24 complex bands k", lambda qm: qm.system.config(2, nbands=24, grid=48), 3),
    ("C0 32^3 W-like 40 bands (32 projectors)", lambda qm: qm.system.wlike(0, nbands=40), 0),
], ids=["C0-real", "C2-k", "W-like"])
def test_nonlocal_shapes_bit_identical(qm, torch_cuda, label, build, ik):
    """Every projection shape the nonlocal autotuner may pick (1, 2 or 4 column
    tiles per workgroup) gives the same bits, and the autotuner records its
    choice for the shape."""
    torch = torch_cuda
    s = build(qm)
    H = qm.hamiltonian.Hamiltonian(s)
    x = torch.from_numpy(s.initial_bands()).to("cuda", torch.float32)
    outs = []
    for tiles in (1, 2, 4):
        print(f"case {label}: {tiles} column tiles", flush=True)
        H.nonlocal_force(tiles)
        y = torch.full_like(x, float("nan"))
        H.apply(y, x, k_index=ik)
        outs.append(y)
    torch.cuda.synchronize()
    for y in outs[1:]:
        assert torch.equal(y, outs[0])
    H.nonlocal_force(0)
    y = torch.full_like(x, float("nan"))
    H.apply(y, x, k_index=ik)
    torch.cuda.synchronize()
    assert torch.equal(y, outs[0])
    kt = bool(np.any(np.abs(s.kpoints[ik]) >= 1e-8))
    assert H.nonlocal_choice(32, s.nbands, kt=kt) in (1, 2, 4)
    H.close()
