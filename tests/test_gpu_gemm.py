"""GPU Gram (A^T B, FP64 result) and update (Z C) products through the C ABI
against float64 numpy products of the same (rounded) inputs.

Tolerances: FP64 <= 1e-13 relative to sum |a||b|; FP32 Gram <= 1e-6 of
sum |a||b| (FP32 products, FP64 sums of 256-term FP32 partials); FP32 update
<= 1e-5 relative (FP32 MFMA accumulation over m terms)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 4096), (24, 48, 32768), (7, 5, 1000), (96, 64, 262144), (130, 70, 20160), (64, 192, 3),
          # both sides >= 128 columns and K a multiple of the panel: the LDS-tiled kernel
          # (the 64^3 x 128 LOBPCG block, partial tiles, a K split over few tiles)
          (384, 256, 262144), (130, 200, 20160), (512, 384, 110592),
          # the same sizes with K off the panel size: the untiled kernel
          (130, 200, 20161)]


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gram(qm, torch_cuda, m, n, k, prec):
    torch = torch_cuda
    print(f"case gram m={m} n={n} k={k} {prec}", flush=True)
    H = qm.hamiltonian.Hamiltonian(qm.system.small_system(n=8, natoms=0))
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    dt = torch.float64 if prec == "f64" else torch.float32
    a = torch.from_numpy(A).to("cuda", dt)
    b = torch.from_numpy(B).to("cuda", dt)
    C = H.gram(a, b).cpu().numpy()
    Ar, Br = a.double().cpu().numpy(), b.double().cpu().numpy()
    want = Ar @ Br.T
    scale = np.abs(Ar) @ np.abs(Br).T
    tol = 1e-13 if prec == "f64" else 1e-6
    assert (np.abs(C - want) <= tol * scale + 1e-300).all(), np.max(np.abs(C - want) / scale)
    H.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gram_unaligned_columns(qm, torch_cuda, prec):
    """Column blocks that do not start on a 16-byte boundary take the scalar-load
    path (A as a view starting one element into its buffer)."""
    torch = torch_cuda
    H = qm.hamiltonian.Hamiltonian(qm.system.small_system(n=8, natoms=0))
    m, n, k = 160, 144, 4096
    rng = np.random.default_rng(7)
    dt = torch.float64 if prec == "f64" else torch.float32
    buf = torch.from_numpy(rng.standard_normal(m * k + 1)).to("cuda", dt)
    a = buf[1:].view(m, k)
    b = torch.from_numpy(rng.standard_normal((n, k))).to("cuda", dt)
    print(f"case gram unaligned {prec}", flush=True)
    C = H.gram(a, b).cpu().numpy()
    Ar, Br = a.double().cpu().numpy(), b.double().cpu().numpy()
    scale = np.abs(Ar) @ np.abs(Br).T
    tol = 1e-13 if prec == "f64" else 1e-6
    assert (np.abs(C - Ar @ Br.T) <= tol * scale).all()
    H.close()


@pytest.mark.parametrize("m,n,k", [(48, 16, 32768), (72, 48, 20160), (9, 130, 1000), (384, 256, 262144),
                                   # LDS-tiled kernel: n >= 128, K (= m) a multiple of the panel, whole point tiles
                                   (96, 200, 20480), (512, 384, 110592),
                                   # untiled: m off the panel size
                                   (100, 200, 20480)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_update(qm, torch_cuda, m, n, k, prec):
    torch = torch_cuda
    print(f"case update m={m} n={n} k={k} {prec}", flush=True)
    H = qm.hamiltonian.Hamiltonian(qm.system.small_system(n=8, natoms=0))
    rng = np.random.default_rng(m + n)
    Z = rng.standard_normal((m, k))
    C = rng.standard_normal((m, n))
    dt = torch.float64 if prec == "f64" else torch.float32
    z = torch.from_numpy(Z).to("cuda", dt)
    c = torch.from_numpy(C).to("cuda", dt)
    Y = H.update(z, c).double().cpu().numpy()
    want = c.double().cpu().numpy().T @ z.double().cpu().numpy()
    scale = np.abs(c.double().cpu().numpy()).T @ np.abs(z.double().cpu().numpy())
    tol = 1e-13 if prec == "f64" else 1e-5
    assert (np.abs(Y - want) <= tol * scale).all(), np.max(np.abs(Y - want) / scale)
    H.close()
