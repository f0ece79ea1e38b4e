"""Host-side mirror of LOBPCG::Eigen_z_multi_mpi (lobpcg_z_multi_mpi.h:1732-2852)
on the GPU.

`LOBPCG(hamiltonian, nbands, precision)` owns the solver's work arena on the
device (the reference's X, R, P, AX, AR, AP and tmpM blocks, 1779-1786, plus
the subspace matrices, 2016-2022) and keeps the P block between calls like the
reference's `keep` argument.  `solve(x, iter_max, k_index, first_time)` runs
the iterations in place on the torch tensor `x` (SoA band layout, FP64 for
precision 64, FP32 for precision 32) and returns the eigenvalues (Ha) and the
squared residual norms of the last residual block.

With `residual_threshold` set, the solver locks converged low bands as the
reference does (p_num_fix / residual_threshold, 1962-1986): `num_fix` is kept
on the object between calls (reset it with `reset_locking()` at each SCF step,
as qumasun_kpoint.scfloop.h:76-80 does), and the eigenvalue / residual entries
of fixed bands keep their values from the call that last iterated them.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def cluster_mean_errors(ev, ref, split=1.0e-5):
    """Per-band |ev - ref| with bands of `ref` closer than `split` grouped into
    clusters and compared through their means (rotations inside a near-degenerate
    cluster move individual Ritz values by up to the split, not the cluster trace)."""
    ev = np.asarray(ev)
    ref = np.asarray(ref)[:ev.size]
    err = np.empty(ev.size)
    i = 0
    while i < ev.size:
        j = i + 1
        while j < ev.size and ref[j] - ref[j - 1] < split:
            j += 1
        err[i:j] = abs(ev[i:j].mean() - ref[i:j].mean())
        i = j
    return err


class LOBPCG:
    PHASES = ("hpsi", "gram", "eigensolve", "update", "other", "total")

    def __init__(self, hamiltonian, nbands, precision=64, ncomp=None, residual_threshold=None):
        self.H = hamiltonian
        self.residual_threshold = residual_threshold  # None: no locking
        self.num_fix = 0
        self._ev = np.zeros(nbands)
        self._rn = np.zeros(nbands)
        self.lib = hamiltonian.lib
        self.nbands = nbands
        self.ncomp = hamiltonian.sys.ncomp if ncomp is None else ncomp
        self.precision = precision
        h = ctypes.c_void_p()
        _lib.check(self.lib.qm_lobpcg_create(ctypes.byref(h), hamiltonian.ctx, nbands, self.ncomp, precision))
        self.h = h
        self.iterations = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.qm_lobpcg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, x, iter_max, k_index=0, first_time=True):
        import torch
        want = torch.float64 if self.precision == 64 else torch.float32
        assert x.is_cuda and x.is_contiguous() and x.dtype == want
        assert x.numel() == self.nbands * self.ncomp * self.H.sys.ng
        ev, rn = self._ev, self._rn  # in/out: entries of fixed bands are kept
        it = ctypes.c_int(0)
        st = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
        if self.residual_threshold is None:
            _lib.check(self.lib.qm_lobpcg_solve(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(ev.ctypes.data),
                                                ctypes.c_void_p(rn.ctypes.data), iter_max, k_index, int(first_time),
                                                ctypes.byref(it), st))
        else:
            nf = ctypes.c_int(self.num_fix)
            _lib.check(self.lib.qm_lobpcg_solve_locking(
                self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(ev.ctypes.data), ctypes.c_void_p(rn.ctypes.data),
                iter_max, k_index, int(first_time), ctypes.byref(nf), float(self.residual_threshold), ctypes.byref(it),
                st))
            self.num_fix = nf.value
        self.iterations = it.value
        return ev.copy(), rn.copy()

    def reset_locking(self):
        """num_fix = 0 (the reference resets it at each SCF step)."""
        self.num_fix = 0

    def set_conditioning(self, min_pivot_ratio_fp64=1e-14, min_pivot_ratio_fp32=1e-4):
        _lib.check(self.lib.qm_lobpcg_set_conditioning(self.h, min_pivot_ratio_fp64, min_pivot_ratio_fp32))

    def set_rollback_limit(self, limit=1.0e-3):
        """|1 - |x'|^2| above which an update is rolled back (0: the first update of a call)."""
        _lib.check(self.lib.qm_lobpcg_set_rollback_limit(self.h, float(limit)))

    @property
    def fallbacks(self):
        return self.lib.qm_lobpcg_fallbacks(self.h)

    def eigen_retries(self):
        """(subspace eigensolves redone in the last solve, note on the last one)"""
        buf = ctypes.create_string_buffer(512)
        n = self.lib.qm_lobpcg_eigen_retries(self.h, buf, 512)
        return n, buf.value.decode()

    def set_graphs(self, enable=True):
        """HIP-graph replay of the A.R and Gram segments of an iteration (default on)."""
        _lib.check(self.lib.qm_lobpcg_set_graphs(self.h, int(bool(enable))))

    @property
    def graph_replays(self):
        return self.lib.qm_lobpcg_graph_replays(self.h)

    def timings(self):
        t = np.zeros(6)
        _lib.check(self.lib.qm_lobpcg_timings(self.h, ctypes.c_void_p(t.ctypes.data)))
        return dict(zip(self.PHASES, t.tolist()))
