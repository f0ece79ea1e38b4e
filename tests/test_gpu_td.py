"""Real-time propagators on the GPU (qm_td_taylor_*, qm_td_expv_*) against the
oracle (tests/test_oracle.py pins the oracle's Taylor series to the dense
exponential, O(dt^5) for order 4).

Tolerances: FP64 <= 1e-12 relative (max-norm of the difference over the
max-norm of the oracle result); FP32 <= 2e-6 relative, the FP32 H.psi bound,
since each of the order terms is an FP32 apply scaled by (dt |H|)^n / n! < 1."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = {"full": 3, "local": 1, "nonlocal": 2}


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _taylor(H, torch, x, nb, ik, dt, order, op):
    from qumasun_amd import _lib
    fn = H.lib.qm_td_taylor_f64 if x.dtype == torch.float64 else H.lib.qm_td_taylor_f32
    st = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(fn(H.ctx, ctypes.c_void_p(x.data_ptr()), nb, ik, float(dt), order, OPS[op], st))


CASES = [
    ("small 12^3 k=(1/2,1/4,-1/8) 3 atoms", lambda qm: qm.system.small_system(
        n=12, L=4.0, natoms=3, rc=2.6, nbands=3, ncomp=2, k_red=(0.5, 0.25, -0.125)), 0),
    ("C2 48^3 4 complex bands k=(1/4,1/4,1/4)", lambda qm: qm.system.config(2, nbands=4, grid=48), 3),
    ("C2 40^3 W-like 32 projectors k=(1/4,0,0)", lambda qm: qm.system.wlike(2, nbands=2, grid=40), 1),
]


@pytest.mark.parametrize("label,build,ik", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("op", ["full", "local", "nonlocal"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_td_taylor_parity(qm, orc, torch_cuda, label, build, ik, op, prec):
    torch = torch_cuda
    s = build(qm)
    assert s.ncomp == 2
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    dt = 0.02
    print(f"case Taylor-4 {label} op {op} {prec}", flush=True)
    dtype = torch.float64 if prec == "f64" else torch.float32
    x = torch.from_numpy(psi).to("cuda", dtype)
    _taylor(H, torch, x, s.nbands, ik, dt, 4, op)
    torch.cuda.synchronize()
    src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
    want = orc.td_taylor(s, src, dt, order=4, op=op, k=s.kpoints[ik])
    err = _rel(x.double().cpu().numpy(), want)
    print(f"rel err {err:.2e}")
    assert err <= (1e-12 if prec == "f64" else 2e-6)
    H.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_td_expv_parity(qm, orc, torch_cuda, prec):
    torch = torch_cuda
    s = qm.system.config(2, nbands=3, grid=48)
    H = qm.hamiltonian.Hamiltonian(s)
    psi = s.initial_bands()
    dtype = torch.float64 if prec == "f64" else torch.float32
    x = torch.from_numpy(psi).to("cuda", dtype)
    fn = H.lib.qm_td_expv_f64 if prec == "f64" else H.lib.qm_td_expv_f32
    st = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    qm._lib.check(fn(H.ctx, ctypes.c_void_p(x.data_ptr()), s.nbands, 0.37, st))
    torch.cuda.synchronize()
    src = psi if prec == "f64" else psi.astype(np.float32).astype(np.float64)
    want = orc.td_expv(s, src, 0.37)
    assert _rel(x.double().cpu().numpy(), want) <= (1e-14 if prec == "f64" else 1e-6)
    H.close()


def test_td_taylor_full_size_c2(qm, orc, torch_cuda):
    """C2 at full size (96^3, 256 complex bands): one order-4 step of the whole
    block at the last k-point in FP64; 3 bands against the oracle (bands
    propagate independently), and the block's norms, which the exact
    propagator conserves, kept to the truncation error of the step."""
    torch = torch_cuda
    s = qm.system.config(2)
    H = qm.hamiltonian.Hamiltonian(s)
    ik = 3
    g = torch.Generator(device="cuda:0").manual_seed(7)
    x = torch.rand((2 * s.nbands, s.ng), device="cuda:0", dtype=torch.float64, generator=g) - 0.5
    sel = [0, 101, 255]
    rows = [2 * b + c for b in sel for c in range(2)]
    src = x[rows].cpu().numpy()
    n0 = (x * x).view(s.nbands, 2, -1).sum(dim=(1, 2))
    dt = 2.0e-4
    print("case Taylor-4 C2 full size FP64", flush=True)
    _taylor(H, torch, x, s.nbands, ik, dt, 4, "full")
    torch.cuda.synchronize()
    want = orc.td_taylor(s, src, dt, order=4, op="full", k=s.kpoints[ik])
    err = _rel(x[rows].cpu().numpy(), want)
    n1 = (x * x).view(s.nbands, 2, -1).sum(dim=(1, 2))
    drift = float(((n1 - n0).abs() / n0).max())
    print(f"oracle rel err {err:.2e}, max relative norm change {drift:.2e}")
    assert err <= 1e-12
    assert drift <= 1e-9
    H.close()
    del x
    torch.cuda.empty_cache()
