"""The C-ABI library: it loads on a machine without a GPU, exports every
function that include/rgs.h declares, the Python binding declares the same
set, and calls fail with a status code (never abort) when no device exists."""
import ctypes
import os
import re
import subprocess

import pytest

from rustscan_amd import _lib, build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rgs.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = re.findall(r"\b(rgs_[a-z0-9_]+)\s*\(", text)
    # typedef'd structs are not followed by "(", so this is the function list
    return sorted(set(names))


@pytest.fixture(scope="module")
def libpath():
    if not os.path.exists(build.LIB_PATH):
        build.build_hip()
    return build.LIB_PATH


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["rgs_ctx_create", "rgs_ctx_destroy", "rgs_last_error", "rgs_splats_create",
                     "rgs_splats_import_aos", "rgs_splats_export_aos", "rgs_render_forward", "rgs_render_backward",
                     "rgs_l1_loss_grad", "rgs_adam_step", "rgs_train_step", "rgs_autotune", "rgs_get_counts",
                     "rgs_sync", "rgs_sort_pairs_u32", "rgs_inclusive_scan_u32"]:
        assert required in names, required


def test_library_exports_every_declared_symbol(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (rgs_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared in rgs.h but not exported: {missing}"
    # and through the dynamic loader
    L = ctypes.CDLL(libpath)
    for n in declared_functions():
        assert hasattr(L, n), n


def test_integration_index_lists_the_header():
    """The entry point index of INTEGRATION.md names every declared function,
    and nothing else."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    section = text.split("## 5. Entry point index", 1)[1]
    listed = set(re.findall(r"`(rgs_[a-z0-9_]+)`", section))
    declared = set(declared_functions())
    assert listed == declared, (sorted(declared - listed), sorted(listed - declared))


def test_python_binding_covers_the_header():
    bound = set(_lib._SIGNATURES)
    declared = set(declared_functions())
    assert bound == declared, (sorted(declared - bound), sorted(bound - declared))


def test_exports_are_plain_c_symbols(libpath):
    """No C++ mangling on the boundary: every rgs_ symbol is extern "C"."""
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True, text=True, check=True).stdout
    mangled = [l for l in out.splitlines() if " T _Z" in l and "rgs_" in l and "rgs_ctx" not in l]
    assert not [l for l in mangled if re.search(r"_Z\d+rgs_", l)], mangled


def test_version_string(libpath):
    L = _lib.lib()
    v = L.rgs_version().decode()
    assert re.search(r"\d+\.\d+", v) and "gfx950" in v, v


def test_calls_fail_with_status_without_device(libpath):
    """In a process without a HIP device, context creation reports an error
    code instead of aborting, and null handles are rejected."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.rgs_ctx_create(0, ctypes.byref(h))
    assert rc != _lib.RGS_OK
    assert not h.value
    assert L.rgs_ctx_create(0, None) == _lib.RGS_ERR_INVALID
    assert L.rgs_sync(None) != _lib.RGS_OK


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """The product path raises when the HIP library is absent; nothing falls back."""
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "librgs_hip.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.lib()
