"""CPU checks of the C-ABI boundary: libppmg.so loads, exports every entry
point include/ppmg.h declares, and rejects bad arguments before touching the
device (argument checks need no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ppmg.h")
LIB = os.path.join(ROOT, "ppm_legacy_amd", "libppmg.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(ppmg_\w+)\s*\(", text)))


def test_header_declares_core_and_registry():
    names = declared_functions()
    for n in ["ppmg_p2m", "ppmg_m2p", "ppmg_p2m_adjoint", "ppmg_util_rank", "ppmg_rmsh_remesh",
              "ppmg_rmsh_comp_weights", "ppmg_rmsh_create_part", "ppmg_impose_part_bc",
              "ppmg_interp_p2m", "ppmg_interp_m2p", "ppmg_init", "ppmg_mktopo_field"]:
        assert n in names


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        from ppm_legacy_amd import build
        build.build(verbose=False)
    return LIB


def test_library_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT\s+(ppmg_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    L = ctypes.CDLL(built)
    for n in declared_functions():
        assert getattr(L, n) is not None


def test_python_binding_lists_every_declared_symbol():
    from ppm_legacy_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_functions()


def test_argument_errors_without_device(built):
    from ppm_legacy_amd import _lib
    L = _lib.lib()
    plan = ctypes.c_void_p()
    assert L.ppmg_plan_create(ctypes.byref(plan)) == 0
    m = _lib.PpmgMesh()
    m.dim = 3
    for d in range(3):
        m.nm[d] = 9
        m.max_phys[d] = 1.0
    for f in range(6):
        m.bcdef[f] = 1
    fld = _lib.PpmgField()
    fld.data = 1
    fld.lda = 1
    # kernel 3 (M'3/6) is not an interpolation kernel: ppm_interp_p2m.f:3568-3573
    info = L.ppmg_p2m(ctypes.byref(m), 3, 0, ctypes.c_int64(1), ctypes.c_void_p(8), 1, ctypes.c_void_p(8),
                      ctypes.byref(fld), plan, None)
    assert info == -2 and _lib.last_error()[1] == _lib.ERR_ARGUMENT
    info = L.ppmg_p2m(ctypes.byref(m), 2, 0, ctypes.c_int64(-1), None, 1, None, ctypes.byref(fld), plan, None)
    assert info == -2 and _lib.last_error()[1] == _lib.ERR_ARGUMENT
    info = L.ppmg_m2p(ctypes.byref(m), 2, 0, ctypes.c_int64(4), None, 0, ctypes.byref(fld), None, plan, None)
    assert info == -2
    bad = _lib.PpmgMesh()
    bad.dim = 4
    info = L.ppmg_impose_part_bc(ctypes.byref(bad), 0, ctypes.c_int64(0), None, None)
    assert info == -2 and "dim" in _lib.last_error()[2]
    # adjoint without a forward on the plan
    info = L.ppmg_p2m_adjoint(ctypes.byref(m), 2, 0, ctypes.c_int64(4), None, 1, None, ctypes.byref(fld), None,
                              None, plan, None)
    assert info == -2
    # registry calls before ppm_init
    tid, mid = ctypes.c_int(), ctypes.c_int()
    mn = (ctypes.c_double * 3)(0, 0, 0)
    mx = (ctypes.c_double * 3)(1, 1, 1)
    bc = (ctypes.c_int * 6)(1, 1, 1, 1, 1, 1)
    nm = (ctypes.c_int * 3)(9, 9, 9)
    g = (ctypes.c_int * 3)(1, 1, 1)
    info = L.ppmg_mktopo_field(3, mn, mx, bc, nm, g, ctypes.byref(tid), ctypes.byref(mid))
    assert info == -2 and _lib.last_error()[1] == _lib.ERR_NOINIT
    assert L.ppmg_plan_destroy(plan) == 0


def test_product_path_has_no_oracle_dependency():
    """The package never imports or links the CPU oracle."""
    pkg = os.path.join(ROOT, "ppm_legacy_amd")
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f"{f} mentions the oracle"
