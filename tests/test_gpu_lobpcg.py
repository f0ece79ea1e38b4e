"""GPU LOBPCG (C ABI qm_lobpcg_*) against the FP64 LOBPCG oracle and the
ARPACK eigenvalues of the 2-atom bcc system (tests/golden/c0_eigenvalues.json).

Tolerances: FP64 path vs oracle after the same iterations from the same start
<= 1e-12 relative in every eigenvalue; converged FP64 eigenvalues vs ARPACK
<= 1e-10 Ha; mixed FP32/FP64 path vs ARPACK <= 1e-6 Ha on the converged bands.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.array(json.load(open(os.path.join(HERE, "golden", "c0_eigenvalues.json")))["eigenvalues_Ha"])


def _oracle_run(qm, orc, s, X, iters, calls, ik=0):
    import lobpcg_ref
    k = s.kpoints[ik]
    L = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a, k=k), s.ncomp, s.dV)
    ev = None
    for c in range(calls):
        X, ev, rn, _ = L.solve(X, iters, first_time=(c == 0))
    return X, ev


def _gpu_run(qm, torch, s, X, iters, calls, prec, ik=0):
    H = qm.hamiltonian.Hamiltonian(s)
    L = qm.lobpcg.LOBPCG(H, s.nbands, precision=prec)
    dtype = torch.float64 if prec == 64 else torch.float32
    x = torch.from_numpy(X).to(device="cuda", dtype=dtype)
    ev = rn = None
    for c in range(calls):
        ev, rn = L.solve(x, iters, k_index=ik, first_time=(c == 0))
    t = L.timings()
    out = x.double().cpu().numpy()
    L.close()
    H.close()
    return out, ev, rn, t


@pytest.fixture
def route(qm, request):
    qm.hamiltonian.subspace_force(request.param)
    yield request.param
    qm.hamiltonian.subspace_force(-1)


@pytest.mark.parametrize("route", [-1, 0], indirect=True, ids=["coop", "single"])
def test_lobpcg_fp64_matches_oracle_c0(qm, orc, torch_cuda, route):
    """8 iterations: <= 1e-12 relative.  20 iterations: within 10x the
    oracle's own spread under a 1e-15 relative perturbation of the start
    vectors (once converged bands make S ill-conditioned the iteration
    amplifies rounding; this bounds what any FP64 implementation can match)."""
    s = qm.system.config(0)
    X0 = s.initial_bands()
    print("case C0 FP64 8 iterations vs oracle", flush=True)
    _, ev_o = _oracle_run(qm, orc, s, X0, 8, 1)
    X, ev_g, rn, t = _gpu_run(qm, torch_cuda, s, X0, 8, 1, 64)
    err = np.abs(ev_g - ev_o).max() / np.abs(ev_o).max()
    print("8 iterations: max rel eigenvalue diff", err)
    assert err <= 1e-12
    nrm = (X * X).sum(axis=1) * s.dV  # returned vectors are dV-normalised
    assert np.abs(nrm - 1).max() < 1e-12
    print("case C0 FP64 20 iterations vs oracle", flush=True)
    _, ev_o = _oracle_run(qm, orc, s, X0, 20, 1)
    Xp = X0 * (1 + 1e-15 * np.random.default_rng(1).standard_normal(X0.shape))
    _, ev_p = _oracle_run(qm, orc, s, Xp, 20, 1)
    spread = np.abs(ev_p - ev_o).max()
    _, ev_g, rn, t = _gpu_run(qm, torch_cuda, s, X0, 20, 1, 64)
    print("timings ms", t)
    diff = np.abs(ev_g - ev_o).max()
    print(f"20 iterations: max |gpu - oracle| {diff:.3e} Ha, oracle self-spread {spread:.3e} Ha")
    assert diff <= 10 * spread + 1e-12


@pytest.mark.parametrize("prec,tol", [(64, 1e-10), (32, 1e-6)])
def test_lobpcg_converges_c0(qm, torch_cuda, prec, tol):
    """Converged eigenvalues vs ARPACK.  Bands closer than 1e-5 Ha (the lowest
    pair is split by 1.3e-6 Ha) are compared through their cluster mean."""
    s = qm.system.config(0)
    X0 = s.initial_bands()
    print(f"case C0 precision {prec}: 25 calls x 20 iterations", flush=True)
    _, ev, rn, t = _gpu_run(qm, torch_cuda, s, X0, 20, 25, prec)
    err = qm.lobpcg.cluster_mean_errors(ev, GOLD)
    print("per-band |error| (Ha):", np.array2string(np.abs(ev - GOLD[:24]), precision=2))
    print("cluster-mean |error| (Ha):", np.array2string(err, precision=2))
    print("timings ms (last call)", t)
    # bands 0..20 lie below the slowly converging top cluster
    assert err[:21].max() <= tol


@pytest.mark.parametrize("route", [-1, 0], indirect=True, ids=["coop", "heevd"])
def test_lobpcg_complex_kpoint_matches_oracle(qm, orc, torch_cuda, route):
    """Both routes of the complex subspace solve: the cooperative tridiagonal
    reduction + stedc + unmtr, and rocSOLVER heevd."""
    s = qm.system.config(2, nbands=8, grid=40)
    X0 = s.initial_bands()
    print(f"case C2 40^3 8 complex bands k=(1/4,1/4,1/4) FP64 2 x 6 iterations route {route}", flush=True)
    _, ev_o = _oracle_run(qm, orc, s, X0, 6, 2, ik=3)
    _, ev_g, rn, t = _gpu_run(qm, torch_cuda, s, X0, 6, 2, 64, ik=3)
    err = np.abs(ev_g - ev_o).max() / np.abs(ev_o).max()
    print("max rel eigenvalue diff", err)
    assert err <= 1e-12


def test_lobpcg_mixed_complex_close_to_fp64(qm, torch_cuda):
    """Complex k-point toy (2 atoms, 24^3): mixed FP32/FP64 vs FP64 on the bands the FP64 run has
    converged (|H x - e x| < 1e-5): cluster-mean eigenvalues within 1e-6 Ha."""
    s = qm.system.small_system(n=24, L=5.0, natoms=2, rc=1.6, nbands=8, ncomp=2, k_red=(0.25, 0.25, 0.25), seed=9)
    X0 = s.initial_bands()
    print("case 24^3 2-atom complex k=(1/4,1/4,1/4) mixed vs FP64, 8 x 20 iterations", flush=True)
    _, ev64, rn64, _ = _gpu_run(qm, torch_cuda, s, X0, 20, 8, 64)
    _, ev32, rn32, _ = _gpu_run(qm, torch_cuda, s, X0, 20, 8, 32)
    conv = np.sqrt(rn64) < 1e-5
    err = qm.lobpcg.cluster_mean_errors(ev32, ev64)
    print("mixed - fp64 (Ha):", ev32 - ev64, "converged:", conv)
    assert conv[:4].all()
    assert err[conv].max() <= 1e-6


@pytest.mark.parametrize("ik", [1, 3], ids=["k=(1/4,0,0)", "k=(1/4,1/4,1/4)"])
def test_lobpcg_mixed_complex_c2_cell(qm, torch_cuda, ik):
    """The C2 cell (54-atom bcc 3x3x3, a = 18) on a 40^3 grid with 64 complex
    bands at a k-point: mixed FP32/FP64 vs FP64 cluster means within 1e-6 Ha on
    the 54 states of the s-like band, the stated set.

    64 bands take the s band whole: ARPACK on the oracle operator (k =
    (1/4,0,0)) puts its lowest 24 states within 4 mHa, with an 8-fold degenerate
    group at bands 6-13, so a block of 8 bands cuts through a degenerate
    cluster and converges at the rate set by a zero gap; a block that holds the
    whole s band has the s-p gap above it.  Inside the band, clusters split by
    1e-8 .. 1e-3 Ha keep single Ritz vectors' residuals near 1e-4 for hundreds
    of iterations while their eigenvalues have settled, so the set is taken as
    converged when every FP64 cluster mean moved by < 1e-9 Ha over the last
    call of 25 iterations."""
    s = qm.system.config(2, nbands=64, grid=40)
    X0 = s.initial_bands()
    print(f"case C2 cell 40^3 64 complex bands k index {ik}: mixed vs FP64, 8 x 25 iterations", flush=True)
    res = {}
    for prec in (64, 32):
        H = qm.hamiltonian.Hamiltonian(s)
        L = qm.lobpcg.LOBPCG(H, s.nbands, precision=prec)
        x = torch_cuda.from_numpy(X0).to("cuda", torch_cuda.float64 if prec == 64 else torch_cuda.float32)
        evs = []
        for c in range(8):
            ev, rn = L.solve(x, 25, k_index=ik, first_time=(c == 0))
            evs.append(ev)
        res[prec] = (evs, rn, L.timings())
        L.close()
        H.close()
    (ev64s, rn64, t64), (ev32s, rn32, t32) = res[64], res[32]
    ev64, ev32 = ev64s[-1], ev32s[-1]
    nset = 54
    moved = qm.lobpcg.cluster_mean_errors(ev64s[-1][:nset], ev64s[-2][:nset])
    err = qm.lobpcg.cluster_mean_errors(ev32, ev64)
    print("FP64 sqrt residual norms of the s band:", np.array2string(np.sqrt(rn64[:nset]), precision=1))
    print("FP64 cluster-mean change over the last call (Ha):", float(moved.max()))
    print("per-band |mixed - fp64| (Ha):", np.array2string(np.abs(ev32 - ev64), precision=1))
    print("cluster-mean |error| (Ha):", np.array2string(err, precision=1))
    print("timings FP64 / mixed (ms, last call):", t64, t32)
    assert moved.max() < 1e-9
    assert err[:nset].max() <= 1e-6


@pytest.mark.parametrize("cond,label", [(0.5, "P dropped"), (2.0, "S eigenbasis (SVQB)")],
                         ids=["drop-P", "svqb"])
def test_lobpcg_fallbacks_match_oracle(qm, orc, torch_cuda, cond, label):
    """Subspace-solve fallbacks forced through qm_lobpcg_set_conditioning, FP64, C0:
    - 0.5: the [X R P] overlap fails the squared-pivot test every time P is present
      (P columns are unit but not orthogonal), so each step is redone on [X R];
    - 2.0: every pivot test fails, so every step takes the S eigenbasis.
    The oracle (LOBPCGRef with the same `cond`) takes the same branches;
    eigenvalues agree to <= 1e-12 relative after 2 calls of 4 iterations."""
    import lobpcg_ref
    s = qm.system.config(0)
    X0 = s.initial_bands()
    print(f"case C0 FP64 fallback {label} (cond {cond})", flush=True)
    Lo = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a), s.ncomp, s.dV, cond=cond)
    Xo = X0
    H = qm.hamiltonian.Hamiltonian(s)
    L = qm.lobpcg.LOBPCG(H, s.nbands, precision=64)
    L.set_conditioning(cond, 1e-4)
    x = torch_cuda.from_numpy(X0).to("cuda", torch_cuda.float64)
    for c in range(2):
        Xo, ev_o, rn_o, _ = Lo.solve(Xo, 4, first_time=(c == 0))
        ev_g, rn_g = L.solve(x, 4, first_time=(c == 0))
        err = np.abs(ev_g - ev_o).max() / np.abs(ev_o).max()
        print(f"call {c}: fallbacks {L.fallbacks}, max rel eigenvalue diff {err:.2e}")
        assert L.fallbacks > 0
        assert err <= 1e-12
    L.close()
    H.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_lobpcg_rollback_matches_oracle(qm, orc, torch_cuda, prec):
    """The normalisation rollback (lobpcg_z_multi_mpi.h:2611-2666) forced with a
    rollback limit of 0: the first update of the call is rejected, the previous
    vectors come back renormalised with their own Ritz values and residual
    norms, P = R is kept, and the next call (limit back at 1e-3) continues from
    there.  Against the oracle with the same limit: FP64 eigenvalues and
    residual norms to <= 1e-12 relative; mixed (FP64 Rayleigh quotients of
    the returned FP32 vectors) within 1e-5 relative, residual norms within 1e-3
    relative.  The returned vectors agree to the rounding of the two ordinary
    iterations before the rollback (1e-11 FP64, 2e-4 FP32)."""
    import lobpcg_ref
    s = qm.system.config(0)
    X0 = s.initial_bands()
    print(f"case C0 rollback precision {prec}", flush=True)
    Lo = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a), s.ncomp, s.dV)
    H = qm.hamiltonian.Hamiltonian(s)
    L = qm.lobpcg.LOBPCG(H, s.nbands, precision=prec)
    dtype = torch_cuda.float64 if prec == 64 else torch_cuda.float32
    x = torch_cuda.from_numpy(X0).to("cuda", dtype)
    # two ordinary iterations first, so that the rolled-back call starts with P
    Xo, _, _, _ = Lo.solve(X0, 2, first_time=True)
    L.solve(x, 2, first_time=True)
    Lo.limit_norm_x = 0.0
    L.set_rollback_limit(0.0)
    Xo, ev_o, rn_o, it_o = Lo.solve(Xo, 5, first_time=False)
    ev_g, rn_g = L.solve(x, 5, first_time=False)
    assert it_o == 1 and L.iterations == 1
    tol_e, tol_r = (1e-12, 1e-12) if prec == 64 else (1e-5, 1e-3)
    e_err = np.abs(ev_g - ev_o).max() / np.abs(ev_o).max()
    r_err = np.abs(rn_g - rn_o).max() / np.abs(rn_o).max()
    x_err = np.abs(x.double().cpu().numpy() - Xo).max() / np.abs(Xo).max()
    print(f"rollback call: eigenvalues {e_err:.2e}, residual norms {r_err:.2e}, vectors {x_err:.2e}")
    assert e_err <= tol_e and r_err <= tol_r
    # the vectors carry the two earlier iterations' rounding (FP64: ~2e-12;
    # FP32 trajectory: ~5e-5)
    assert x_err <= (1e-11 if prec == 64 else 2e-4)
    Lo.limit_norm_x = 1e-3
    L.set_rollback_limit(1e-3)
    Xo, ev_o, _, _ = Lo.solve(Xo, 4, first_time=False)
    ev_g, _ = L.solve(x, 4, first_time=False)
    err = np.abs(ev_g - ev_o).max() / np.abs(ev_o).max()
    print(f"next call (P = R kept): max rel eigenvalue diff {err:.2e}")
    assert err <= (1e-12 if prec == 64 else 1e-5)
    L.close()
    H.close()


def _locked_runs(qm, orc, torch, s, X0, schedule, prec=64, ik=0):
    """Runs the oracle and the GPU solver through the same schedule of
    (iterations, residual_threshold or None) calls; returns per-call
    (num_fix, eigenvalues) for both."""
    import lobpcg_ref
    k = s.kpoints[ik]
    Lo = lobpcg_ref.LOBPCGRef(lambda a: orc.hpsi(s, a, k=k), s.ncomp, s.dV)
    H = qm.hamiltonian.Hamiltonian(s)
    Lg = qm.lobpcg.LOBPCG(H, s.nbands, precision=prec)
    dtype = torch.float64 if prec == 64 else torch.float32
    x = torch.from_numpy(X0).to(device="cuda", dtype=dtype)
    Xo = X0
    ev_o = np.zeros(s.nbands)
    out = []
    for c, (iters, thr) in enumerate(schedule):
        Lo.residual_threshold = thr
        Lg.residual_threshold = thr
        Xo, ev_o, _, _ = Lo.solve(Xo, iters, first_time=(c == 0), eigen_values=ev_o)
        ev_g, _ = Lg.solve(x, iters, k_index=ik, first_time=(c == 0))
        out.append((Lo.num_fix if thr is not None else 0, ev_o.copy(), Lg.num_fix, ev_g))
    Xg = x.double().cpu().numpy()
    Lg.close()
    H.close()
    return out, Xo, Xg


@pytest.mark.parametrize("label,build,ik", [
    ("C0 32^3 24 real bands", lambda qm: qm.system.config(0), 0),
    ("C2 40^3 8 complex bands k=(1/4,1/4,1/4)", lambda qm: qm.system.config(2, nbands=8, grid=40), 3),
], ids=["C0-real", "C2-complex"])
def test_lobpcg_locking_matches_oracle(qm, orc, torch_cuda, label, build, ik):
    """Locking (qm_lobpcg_solve_locking) against the oracle, FP64.
    - Calls 1-2: a threshold every residual exceeds fixes N/2 bands on entry.
    - Call 3: the count stays there.
    - The last case fixes bands after an unlocked call with P kept, so P is
      projected too.
    Eigenvalues agree to <= 1e-12 relative, the fixed entries are left as
    passed, the fixed bands of x come back unchanged, and the active bands
    stay orthogonal to them.  Each schedule stops at 9 iterations: beyond
    about 10, converging bands make S ill-conditioned and rounding
    differences grow (3e-12 at 12 iterations here), the same limit as the
    unlocked 8-iteration parity test."""
    s = build(qm)
    X0 = s.initial_bands()
    N = s.nbands
    print(f"case {label}: locked from the start, 3 x 3 iterations", flush=True)
    out, Xo, Xg = _locked_runs(qm, orc, torch_cuda, s, X0, [(3, 1.0e6), (3, 1.0e6), (3, 0.0)], ik=ik)
    for c, (nf_o, ev_o, nf_g, ev_g) in enumerate(out):
        assert nf_o == nf_g == N // 2, (c, nf_o, nf_g)
        err = np.abs(ev_g[N // 2:] - ev_o[N // 2:]).max() / np.abs(ev_o[N // 2:]).max()
        print(f"call {c}: num_fix {nf_g}, max rel eigenvalue diff {err:.2e}")
        assert err <= 1e-12
        assert np.all(ev_g[:N // 2] == 0.0)  # never written: fixed before the first iteration
    nf = (N // 2) * s.ncomp
    assert np.abs(Xg[:nf] - X0[:nf]).max() <= 1e-14 * np.abs(X0).max()
    import lobpcg_ref
    z = lobpcg_ref.soa_to_cplx(Xg, s.ncomp)
    assert np.abs(np.conj(z[:N // 2]) @ z[N // 2:].T * s.dV).max() < 1e-10  # active stay orthogonal to fixed
    print(f"case {label}: unlocked 4 iterations, then locked with P kept, 3 + 2 iterations", flush=True)
    out, Xo, Xg = _locked_runs(qm, orc, torch_cuda, s, X0, [(4, None), (3, 1.0e6), (2, 1.0e6)], ik=ik)
    ev_first = out[0][3]
    for c, (nf_o, ev_o, nf_g, ev_g) in enumerate(out):
        lo = 0 if c == 0 else N // 2
        err = np.abs(ev_g[lo:] - ev_o[lo:]).max() / np.abs(ev_o[lo:]).max()
        print(f"call {c}: num_fix {nf_g}, max rel eigenvalue diff {err:.2e}")
        assert nf_o == nf_g
        assert err <= 1e-12
        if c > 0:
            assert np.array_equal(ev_g[:N // 2], ev_first[:N // 2])


@pytest.mark.parametrize("prec,thr,tol", [(64, 1.0e-6, 1e-10), (32, 1.0e-4, 1e-6)])
def test_lobpcg_locking_converges_c0(qm, torch_cuda, prec, thr, tol):
    """With locking on from the start (bands fixed once ||r|| < thr), the
    converged eigenvalues still match ARPACK: the locked ones keep the value
    from the call that last iterated them.  The bound is |error| <=
    ||r||^2 / gap, well under tol."""
    s = qm.system.config(0)
    X0 = s.initial_bands()
    print(f"case C0 precision {prec} locking threshold {thr}: 25 calls x 20 iterations", flush=True)
    H = qm.hamiltonian.Hamiltonian(s)
    L = qm.lobpcg.LOBPCG(H, s.nbands, precision=prec, residual_threshold=thr)
    x = torch_cuda.from_numpy(X0).to(device="cuda", dtype=torch_cuda.float64 if prec == 64 else torch_cuda.float32)
    counts = []
    for c in range(25):
        ev, rn = L.solve(x, 20, first_time=(c == 0))
        counts.append(L.num_fix)
    L.close()
    H.close()
    err = qm.lobpcg.cluster_mean_errors(ev, GOLD)
    print("num_fix per call:", counts)
    print("cluster-mean |error| (Ha):", np.array2string(err, precision=2))
    assert counts == sorted(counts) and counts[-1] > 0
    assert err[:21].max() <= tol


@pytest.mark.parametrize("label,build,prec,ik", [
    ("C0 32^3 24 real bands FP64", lambda qm: qm.system.config(0), 64, 0),
    ("C0 32^3 24 real bands mixed", lambda qm: qm.system.config(0), 32, 0),
    ("C2 40^3 8 complex bands k=(1/4,1/4,1/4) FP64", lambda qm: qm.system.config(2, nbands=8, grid=40), 64, 3),
], ids=["C0-f64", "C0-mixed", "C2-complex"])
def test_lobpcg_graph_replay_bit_identical(qm, torch_cuda, label, build, prec, ik):
    """The A.R and Gram segments replayed as HIP graphs give the same bits as
    eager execution: eigenvalues, residual norms and vectors, over two calls of
    14 iterations (the first call starts without P; iterations 0, 6 and 12
    recompute every Gram block)."""
    torch = torch_cuda
    s = build(qm)
    X0 = s.initial_bands()
    dtype = torch.float64 if prec == 64 else torch.float32
    out = {}
    for graphs in (False, True):
        print(f"case {label}: graphs {graphs}", flush=True)
        H = qm.hamiltonian.Hamiltonian(s)
        L = qm.lobpcg.LOBPCG(H, s.nbands, precision=prec)
        L.set_graphs(graphs)
        x = torch.from_numpy(X0).to(device="cuda", dtype=dtype)
        res, replays = [], []
        for c in range(2):
            ev, rn = L.solve(x, 14, k_index=ik, first_time=(c == 0))
            res.append((ev, rn))
            replays.append(L.graph_replays)
        out[graphs] = (res, x.cpu().numpy(), replays)
        L.close()
        H.close()
    (r0, x0, rep0), (r1, x1, rep1) = out[False], out[True]
    print("graph replays per call:", rep1)
    assert rep0 == [0, 0] and min(rep1) >= 10
    for (e0, n0), (e1, n1) in zip(r0, r1):
        assert np.array_equal(e0, e1) and np.array_equal(n0, n1)
    assert np.array_equal(x0, x1)
