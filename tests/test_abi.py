"""The C-ABI library loads and exports every symbol include/dl_gpu.h declares
(no compute call: this runs without a GPU)."""
import ctypes

import numpy as np
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


def _rod(n=3):
    from dynlattice_amd import lattice as L
    return L.straight_rod(n, 0.3, 1e-3, 2e9, 0.3, 1200.0)


def _rejected(p, words):
    """dl_create refuses the mesh before any device call: DL_ERR_INPUT with
    a message naming the problem"""
    import pytest
    with pytest.raises(gpu.DLError) as e:
        gpu.GpuModel(p)
    assert e.value.code == gpu.DL_ERR_INPUT, e.value
    assert words in str(e.value), e.value


def test_invalid_meshes_are_input_errors():
    from dynlattice_amd import lattice as L
    p = _rod()
    p.elem_nodes = p.elem_nodes[:0]
    p.elem_strut = p.elem_strut[:0]
    _rejected(p, "empty mesh")
    p = _rod()
    p.elem_nodes = p.elem_nodes.copy()
    p.elem_nodes[1, 1] = 99
    _rejected(p, "out of range")
    p = _rod()
    p.elem_nodes = p.elem_nodes.copy()
    p.elem_nodes[2] = [3, 3]
    _rejected(p, "degenerate element")
    p = _rod()
    p.elem_strut = np.array([0, 1, 0], np.int32)
    p.n_struts = 2
    p.sections = L.Sections.circle(1e-3, 2e9, 0.3, 1200.0, n=2)
    _rejected(p, "contiguous")
    p = _rod()
    p.elem_nodes = np.array([[0, 1], [2, 3], [1, 2]], np.int32)
    _rejected(p, "chain")
    p = _rod()
    p.elem_strut = np.array([0, 0, 2], np.int32)
    p.n_struts = 3
    p.sections = L.Sections.circle(1e-3, 2e9, 0.3, 1200.0, n=3)
    _rejected(p, "without elements")


def test_invalid_plastic_data_is_an_input_error():
    from dynlattice_amd import lattice as L
    p = _rod()
    p.plastic = L.Plastic.resultant_capacities(np.full(1, 1e-3), 20e6)
    p.plastic.exponent = (0.5,) * 6
    _rejected(p, "exponents")
