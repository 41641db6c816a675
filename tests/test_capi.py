"""CPU checks of the drop-in boundary: libhgdt.so loads, exports exactly the entry points that
include/hgdt.h declares, the Python binding covers them, and without a GPU every compute entry
point refuses loudly (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "hgdt.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hgdt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("hgdt_operator_assemble", "hgdt_operator_apply", "hgdt_operator_apply_inverse",
                     "hgdt_operator_append_laplace", "hgdt_operator_append_inner_coupling",
                     "hgdt_operator_adjoint_log_diffusion_gradient", "hgdt_grid_create_union"):
        assert required in names


def test_library_exports_every_declared_symbol(hgdt_mod):
    lib_path = hgdt_mod.library_path()
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hgdt_[a-z0-9_]+)", out))
    declared = set(declared_functions())
    assert declared <= exported, sorted(declared - exported)
    assert {s for s in exported if s.startswith("hgdt_")} == declared  # nothing undeclared leaks out


def test_python_binding_covers_header(hgdt_mod):
    from hgdt import _lib

    assert set(_lib.SIGNATURES) == set(declared_functions())
    lib = hgdt_mod.load_library()
    for name in declared_functions():
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU refusal path")
def test_refuses_without_gpu(hgdt_mod):
    lib = hgdt_mod.load_library()
    count = ctypes.c_int32(-1)
    assert lib.hgdt_device_count(ctypes.byref(count)) == 8  # HGDT_ERR_DEVICE
    with pytest.raises(hgdt_mod.device_error):
        hgdt_mod.make_cube_grid([0, 0], [1, 1], [4, 4])
    # invalid arguments are reported with their own code, never a crash
    assert lib.hgdt_grid_destroy(None) == 0
    h = ctypes.c_void_p()
    assert lib.hgdt_grid_create(0, 0, None, 0, None, ctypes.byref(h)) == 1


def test_status_codes_map_reference_exceptions():
    from hgdt import _lib

    assert _lib._EXC[7].__name__ == "linear_solver_failed"
    assert _lib._EXC[2].__name__ == "shapes_do_not_match"
    assert _lib._EXC[3].__name__ == "operator_error"
