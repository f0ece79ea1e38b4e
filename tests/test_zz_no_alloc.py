"""qm_hpsi_apply_f32_noalloc (the H.psi entry point that allocates nothing)
against qm_hpsi_apply_f32, through the C ABI.

Both run the same kernels with the same arithmetic (every stencil tile shape
computes the same values), so the outputs must be bit-identical.  Each case
makes two calls in a row, the second with fewer bands, so the second reads
work memory that still holds the first call's values.  A band count whose
work memory exceeds the fixed 64 MiB is refused with QM_ERR_ARG before any
launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QM_ERR_ARG = -1


def _apply(H, fn, out, x, nb, ik, torch):
    st = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    return fn(H.ctx, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(x.data_ptr()), nb, H.sys.ncomp, ik, st)


@pytest.mark.parametrize("label,build,nbs,ik", [
    ("C1 64^3 128 real bands, then 40", lambda qm: qm.system.config(1), (128, 40), 0),
    ("C2 48^3 24 complex bands k=(1/4,1/4,0), then 8", lambda qm: qm.system.config(2, nbands=24, grid=48), (24, 8), 2),
    ("C0 32^3 W-like 32 projectors, 24 bands then 6", lambda qm: qm.system.wlike(0, nbands=24), (24, 6), 0),
], ids=["C1-real", "C2-kpoint", "C0-wlike"])
def test_noalloc_matches_apply(qm, torch_cuda, label, build, nbs, ik):
    torch = torch_cuda
    s = build(qm)
    H = qm.hamiltonian.Hamiltonian(s)
    lib = H.lib
    X = torch.from_numpy(s.initial_bands()).to("cuda", torch.float32)
    for nb in nbs:
        print(f"case noalloc {label}: nb {nb}", flush=True)
        x = X[:nb * s.ncomp].contiguous()
        want = torch.empty_like(x)
        got = torch.full_like(x, float("nan"))
        qm._lib.check(_apply(H, lib.qm_hpsi_apply_f32, want, x, nb, ik, torch))
        qm._lib.check(_apply(H, lib.qm_hpsi_apply_f32_noalloc, got, x, nb, ik, torch))
        torch.cuda.synchronize()
        diff = (got - want).abs().max().item()
        print(f"max |noalloc - apply| = {diff}")
        assert torch.equal(got, want)
    H.close()


def test_noalloc_refuses_oversize(qm, torch_cuda):
    torch = torch_cuda
    s = qm.system.config(2, nbands=256)  # 96^3, 54 atoms of 9929 voxels: 1080 chunks
    H = qm.hamiltonian.Hamiltonian(s)
    nchunk = sum(-(-int(s.vox_off[a + 1] - s.vox_off[a]) // 512) for a in range(s.natoms))
    need = 8 * nchunk * 16 * 2 * 256
    print(f"case noalloc oversize: {need} bytes of partial sums for 256 complex bands", flush=True)
    assert need > 64 << 20
    x = torch.zeros((2 * 256, s.ng), device="cuda", dtype=torch.float32)
    out = torch.zeros_like(x)
    st = _apply(H, H.lib.qm_hpsi_apply_f32_noalloc, out, x, 256, 0, torch)
    assert st == QM_ERR_ARG
    assert b"64 MiB" in H.lib.qm_last_error()
    H.close()
