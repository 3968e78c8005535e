"""The C-ABI library loads and exports every entry point include/hdog.h declares.

Runs without a GPU: only symbol lookup, no compute call.
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hdog.h")
LIB = os.path.join(ROOT, "cubic-informatics_amd", "hdog", "libhdog.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hdog_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("hdog_create", "hdog_normalize3d", "hdog_init_gaussian3d", "hdog_dog3d",
              "hdog_hessian_pd", "hdog_ccl3d", "hdog_regional_features",
              "hdog_process_substack"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (same HIP runtime resolution as the product binding)
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_match_header():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "cubic-informatics_amd"))
    from hdog import _capi
    assert sorted(_capi.EXPORTED) == declared_functions()
