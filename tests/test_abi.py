"""The C-ABI library loads without a GPU and exports every function that
include/tonto_gpu.h and include/tonto_mlp.h declare (no compute calls here)."""
import ctypes
import os
import re

from tonto_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="tonto_gpu.h", prefix="tgpu_"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("tgpu_create", "tgpu_destroy", "tgpu_set_cutoffs", "tgpu_schwarz",
                 "tgpu_make_r_JK", "tgpu_make_r_JK_device", "tgpu_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    names = declared_functions() + declared_functions("tonto_mlp.h", "tmlp_")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_only_exported_symbols():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    assert all(hasattr(lib, n) for n in gpu.EXPORTS)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "tonto_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".inc")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "liboracle" not in src, f
