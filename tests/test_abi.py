"""The C-ABI library loads on a CPU-only host and exports every entry point
declared in include/stfem.h (no compute calls that need a GPU)."""
import ctypes
import os
import re

import stfem_amd as st

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "stfem.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stfem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(st.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(st._SIGS)


def test_errors_are_reported():
    L = st.load()
    # bad time type: argument error with a message, no crash
    rc = L.stfem_fe_time_weights(7, 1, 1.0, 1, None, None, None, None)
    assert rc == -1
    assert b"time step type" in L.stfem_last_error()
    # bad spec rejected before any device work
    h = ctypes.c_void_p()
    rc = L.stfem_problem_create(b"dim=4", None, None, 0, ctypes.byref(h))
    assert rc == -1 and not h.value
    rc = L.stfem_problem_create(b"dim=2;nts=3", None, None, 0, ctypes.byref(h))
    assert rc == -1 and b"power of two" in L.stfem_last_error()


def test_spec_values_are_checked():
    """Out-of-range solver and kernel keys are rejected with a message before
    any device work (these run on a host without a GPU)."""
    L = st.load()
    h = ctypes.c_void_p()
    for spec, msg in [(b"dim=2;vanka_device_setup=9", b"vanka_device_setup"),
                      (b"dim=2;orthogonalization=2", b"orthogonalization"),
                      (b"dim=2;op_jb=5", b"op_jb"),
                      (b"dim=2;vanka_kernel=1", b"vanka_kernel"),
                      (b"dim=2;coarse_coef=4", b"coarse_coef"),
                      (b"dim=2;basis=61", b"basis")]:
        rc = L.stfem_problem_create(spec, None, None, 0, ctypes.byref(h))
        assert rc == -1 and not h.value, spec
        assert msg in L.stfem_last_error(), (spec, L.stfem_last_error())
