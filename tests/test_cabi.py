"""The C-ABI library loads without a GPU and exports every entry point that
include/qumasun_hip.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(qm_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol(qm):
    lib = ctypes.CDLL(qm._lib.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header(qm):
    assert set(qm._lib.EXPORTS) == _declared()


def test_load_sets_signatures(qm):
    L = qm._lib.load()
    assert L.qm_hpsi_apply_f64.restype == ctypes.c_int
    assert L.qm_last_error.restype == ctypes.c_char_p
    assert L.qm_last_error() in (b"", None) or isinstance(L.qm_last_error(), bytes)


def test_errors_without_device(qm):
    """Argument checks run before any device call and report through qm_last_error."""
    L = qm._lib.load()
    ctx = ctypes.c_void_p()
    st = L.qm_create(ctypes.byref(ctx), 2, 8, 8, 0.1, 0.1, 0.1, 0)
    assert st == -1  # QM_ERR_ARG: grid dimensions must be >= 4
    assert b"grid" in L.qm_last_error()


def test_stencil_candidate_queries(qm):
    """The tile-shape autotuner's candidate list is queryable without a GPU;
    out-of-range queries are argument errors, not crashes."""
    lib = qm._lib.load()
    n = lib.qm_stencil_num_candidates()
    assert n >= 2
    buf = ctypes.create_string_buffer(128)
    names = set()
    for i in range(n):
        assert lib.qm_stencil_candidate_name(i, buf, len(buf)) == qm._lib.QM_OK
        assert buf.value.startswith((b"fd8_wave_kernel", b"fd8_slab_kernel"))
        names.add(buf.value)
    assert len(names) == n  # every candidate is a distinct shape
    assert lib.qm_stencil_candidate_name(n, buf, len(buf)) == -1
    assert lib.qm_stencil_candidate_name(-1, buf, len(buf)) == -1
    small = ctypes.create_string_buffer(8)
    assert lib.qm_stencil_candidate_name(0, small, len(small)) == qm._lib.QM_OK
    assert len(small.value) == 7  # truncated, NUL-terminated
    assert lib.qm_stencil_force(None, 0) == -1


def test_sygvd_argument_errors(qm):
    lib = qm._lib.load()
    assert lib.qm_sygvd_f64(None, 4, None, None, None, None, None, 2, None) == -1


def test_lobpcg_locking_argument_errors(qm):
    import ctypes as C
    lib = qm._lib.load()
    nf = C.c_int(0)
    assert lib.qm_lobpcg_solve_locking(None, None, None, None, 1, 0, 1, C.byref(nf), 0.0, None, None) == -1
    assert b"NULL" in lib.qm_last_error()


def test_lobpcg_graph_switch_argument_errors(qm):
    lib = qm._lib.load()
    assert lib.qm_lobpcg_set_graphs(None, 1) == -1
    assert b"NULL" in lib.qm_last_error()
    assert lib.qm_lobpcg_graph_replays(None) == -1
