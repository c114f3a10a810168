"""The C-ABI library loads and exports every symbol include/dl_gpu.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from dynlattice_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "dl_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dl_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_bound_symbols():
    assert set(declared_symbols()) == set(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_last_error_callable_without_device():
    L = gpu.load()
    assert isinstance(L.dl_last_error(), bytes)


def test_null_context_is_an_input_error():
    L = gpu.load()
    assert L.dl_sync(None) == gpu.DL_ERR_INPUT
    assert L.dl_destroy(None) == gpu.DL_OK
