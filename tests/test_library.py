"""CPU checks of the C-ABI library: it loads, and exports every symbol that
include/split_operator_hip.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "split_operator_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(so_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_crate_surface():
    names = declared_functions()
    for required in ["so_wave_function_norm", "so_fft_transformation_create", "so_matrix_transformation_set_matrix",
                     "so_border_dumping_create", "so_leak_control_create", "so_state_saver_create",
                     "so_propagation_propagate", "so_propagation_mean_energy", "so_operation_stack_add_saver"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from split_operator_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(declared_functions()) - {"so_last_error"} <= set(_lib.SIGNATURES)


def test_product_package_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "split_operator_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                with open(os.path.join(dirpath, fn)) as f:
                    src = f.read()
                assert "import oracle" not in src and "from oracle" not in src, fn


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from split_operator_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()
