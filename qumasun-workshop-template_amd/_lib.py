"""ctypes binding of libqumasun_hip.so (the C ABI in include/qumasun_hip.h).

The library is built in-tree by build.py.  There is no fallback: if the
library is missing, loading raises, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libqumasun_hip.so")

# every entry point declared in include/qumasun_hip.h
EXPORTS = (
    "qm_last_error", "qm_create", "qm_destroy", "qm_set_vloc", "qm_set_projectors", "qm_set_kpoints",
    "qm_hpsi_apply_f64", "qm_hpsi_apply_f32", "qm_hpsi_apply_f32_noalloc", "qm_local_apply_f64", "qm_local_apply_f32",
    "qm_nonlocal_add_f64", "qm_nonlocal_add_f32", "qm_band_density_f64", "qm_band_density_f32",
    "qm_nonlocal_forces_f64", "qm_nonlocal_forces_f32", "qm_gram_f64", "qm_gram_f32", "qm_update_f64",
    "qm_update_f32", "qm_lobpcg_create", "qm_lobpcg_destroy",
    "qm_lobpcg_solve", "qm_lobpcg_solve_locking", "qm_lobpcg_set_conditioning", "qm_lobpcg_fallbacks", "qm_lobpcg_timings",
    "qm_lobpcg_set_graphs", "qm_lobpcg_graph_replays",
    "qm_stencil_num_candidates", "qm_stencil_candidate_name", "qm_stencil_force", "qm_stencil_choice",
    "qm_sygvd_f64", "qm_release_captures", "qm_lobpcg_set_rollback_limit",
    "qm_td_taylor_f64", "qm_td_taylor_f32", "qm_td_expv_f64", "qm_td_expv_f32",
    "qm_set_kinetic", "qm_td_kinetic_fft_f64", "qm_td_kinetic_fft_f32",
    "qm_nonlocal_force", "qm_nonlocal_choice",
    "qm_subspace_force", "qm_coop_set_spin_limit", "qm_coop_spin_limit_default", "qm_hegvd_z64",
    "qm_lobpcg_eigen_retries",
    "qm_hartree_f64", "qm_xc_lda_f64", "qm_grid_dot_f64", "qm_mixer_create", "qm_mixer_destroy", "qm_mixer_init",
    "qm_mixer_predict", "qm_occupancy",
    "qm_upf_load", "qm_upf_free", "qm_upf_info", "qm_upf_radial", "qm_upf_eval", "qm_upf_tables",
)

QM_OK = 0
_STATUS = {-1: "QM_ERR_ARG", -2: "QM_ERR_HIP", -3: "QM_ERR_STATE", -4: "QM_ERR_NUMERIC", -5: "QM_ERR_NOMEM"}


class QMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load():
    """Loads the HIP library (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: run `python qumasun-workshop-template_amd/build.py` "
                      "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(LIB_PATH)
    I, D, P, L64 = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64
    L.qm_last_error.restype = ctypes.c_char_p
    L.qm_last_error.argtypes = []
    L.qm_create.argtypes = [ctypes.POINTER(P), I, I, I, D, D, D, I]
    L.qm_destroy.argtypes = [P]
    L.qm_destroy.restype = None
    L.qm_set_vloc.argtypes = [P, P]
    L.qm_set_projectors.argtypes = [P, I, P, P, P, P, P, P, P, P]
    L.qm_set_kpoints.argtypes = [P, I, P]
    for n in ("qm_hpsi_apply_f64", "qm_hpsi_apply_f32", "qm_hpsi_apply_f32_noalloc", "qm_local_apply_f64",
              "qm_local_apply_f32", "qm_nonlocal_add_f64", "qm_nonlocal_add_f32"):
        getattr(L, n).argtypes = [P, P, P, I, I, I, P]
    for n in ("qm_band_density_f64", "qm_band_density_f32"):
        getattr(L, n).argtypes = [P, P, P, P, I, I, P]
    for n in ("qm_nonlocal_forces_f64", "qm_nonlocal_forces_f32"):
        getattr(L, n).argtypes = [P, P, P, I, I, I, P, P, P]
    for n in ("qm_gram_f64", "qm_gram_f32", "qm_update_f64", "qm_update_f32"):
        getattr(L, n).argtypes = [P, P, I, P, I, L64, P, P]
    L.qm_lobpcg_create.argtypes = [ctypes.POINTER(P), P, I, I, I]
    L.qm_lobpcg_destroy.argtypes = [P]
    L.qm_lobpcg_destroy.restype = None
    L.qm_lobpcg_solve.argtypes = [P, P, P, P, I, I, I, ctypes.POINTER(I), P]
    L.qm_lobpcg_solve_locking.argtypes = [P, P, P, P, I, I, I, ctypes.POINTER(I), D, ctypes.POINTER(I), P]
    L.qm_lobpcg_timings.argtypes = [P, P]
    L.qm_lobpcg_set_conditioning.argtypes = [P, D, D]
    L.qm_lobpcg_fallbacks.argtypes = [P]
    L.qm_lobpcg_set_rollback_limit.argtypes = [P, D]
    for n in ("qm_td_taylor_f64", "qm_td_taylor_f32"):
        getattr(L, n).argtypes = [P, P, I, I, D, I, I, P]
    for n in ("qm_td_expv_f64", "qm_td_expv_f32"):
        getattr(L, n).argtypes = [P, P, I, D, P]
    L.qm_set_kinetic.argtypes = [P, I]
    L.qm_nonlocal_force.argtypes = [P, I]
    L.qm_nonlocal_choice.argtypes = [P, I, I, I, I, ctypes.POINTER(I)]
    for n in ("qm_td_kinetic_fft_f64", "qm_td_kinetic_fft_f32"):
        getattr(L, n).argtypes = [P, P, I, I, D, P]
    L.qm_lobpcg_set_graphs.argtypes = [P, I]
    L.qm_lobpcg_graph_replays.argtypes = [P]
    L.qm_stencil_num_candidates.argtypes = []
    L.qm_stencil_candidate_name.argtypes = [I, ctypes.c_char_p, I]
    L.qm_stencil_force.argtypes = [P, I]
    L.qm_stencil_choice.argtypes = [P, I, I, I, I, ctypes.POINTER(I)]
    L.qm_sygvd_f64.argtypes = [P, I, P, P, P, P, P, I, P]
    L.qm_release_captures.argtypes = [P]
    L.qm_subspace_force.argtypes = [I]
    L.qm_coop_set_spin_limit.argtypes = [ctypes.c_long]
    L.qm_coop_spin_limit_default.argtypes = []
    L.qm_coop_spin_limit_default.restype = ctypes.c_long
    L.qm_hegvd_z64.argtypes = [P, I, P, P, P, P, P, I, P]
    L.qm_lobpcg_eigen_retries.argtypes = [P, ctypes.c_char_p, I]
    L.qm_hartree_f64.argtypes = [P, P, P, P]
    L.qm_xc_lda_f64.argtypes = [P, P, P, P, P]
    L.qm_grid_dot_f64.argtypes = [P, P, P, ctypes.POINTER(D), P]
    L.qm_mixer_create.argtypes = [ctypes.POINTER(P), P, I, I, D]
    L.qm_mixer_destroy.argtypes = [P]
    L.qm_mixer_destroy.restype = None
    L.qm_mixer_init.argtypes = [P, P, P]
    L.qm_mixer_predict.argtypes = [P, P, ctypes.POINTER(D), P]
    L.qm_occupancy.argtypes = [I, I, P, P, D, D, P, ctypes.POINTER(D)]
    L.qm_upf_load.argtypes = [ctypes.POINTER(P), ctypes.c_char_p]
    L.qm_upf_free.argtypes = [P]
    L.qm_upf_free.restype = None
    L.qm_upf_info.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(D), ctypes.POINTER(D)]
    L.qm_upf_radial.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(D), ctypes.POINTER(D)]
    L.qm_upf_eval.argtypes = [P, I, I, P, P, P]
    L.qm_upf_tables.argtypes = [P, I, P, I, I, I, D, D, D, P, P, P, P, P, P, P, P, P]
    for n in EXPORTS:
        if n not in ("qm_last_error", "qm_destroy", "qm_lobpcg_destroy", "qm_upf_free", "qm_mixer_destroy"):
            getattr(L, n).restype = I
    _lib = L
    return L


def check(status):
    if status != QM_OK:
        msg = load().qm_last_error()
        raise QMError(status, msg.decode() if msg else "")
    return status
