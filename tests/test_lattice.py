"""Mesh generators (the data format on the input side of the path)."""
import math

import numpy as np
import pytest

from dynlattice_amd import lattice as L


@pytest.mark.parametrize("n,m", [(1, 1), (2, 3), (3, 4), (10, 4)])
def test_bcc_counts(n, m):
    p = L.bcc_lattice(n, m, 1e-2, np.full(8 * n ** 3, 1e-4), 1e9, 0.3, 1000.0)
    assert p.n_elems == 8 * m * n ** 3
    assert p.n_nodes == (n + 1) ** 3 + n ** 3 + 8 * (m - 1) * n ** 3
    # chains: consecutive elements of a strut share nodes
    same = p.elem_strut[1:] == p.elem_strut[:-1]
    assert np.all(p.elem_nodes[1:, 0][same] == p.elem_nodes[:-1, 1][same])
    # every strut has length sqrt(3)/2 cell
    x = p.coords[p.elem_nodes[:, 1]] - p.coords[p.elem_nodes[:, 0]]
    Ls = np.bincount(p.elem_strut, weights=np.sqrt((x * x).sum(1)))
    assert np.allclose(Ls, math.sqrt(3) / 2 * 1e-2, rtol=1e-12)


def test_config_sizes():
    p = L.bcc_c2()
    assert (p.n_elems, p.n_nodes, p.n_struts) == (32000, 26331, 8000)
    assert len(p.groups["bottom"]) == 121 and len(p.groups["top"]) == 121


def test_rng_is_reproducible_and_normal():
    a = L.normal(2, 20000)
    assert np.array_equal(a, L.normal(2, 20000))
    assert abs(a.mean()) < 0.03 and abs(a.std() - 1) < 0.02
    u = L.uniform01(5, 100000)
    assert u.min() > 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.005
    # splitmix64 reference value (seed 0, first output)
    assert int(L.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF


def test_circle_section():
    s = L.Sections.circle(np.array([0.01]), 210e9, 0.3, 7850.0)
    assert s.area[0] == math.pi * 1e-4
    assert s.J[0] == 2 * s.I1[0]
    assert abs(s.shear_modulus[0] - 210e9 / 2.6) < 1e-3
    assert s.shear_factor[0] == 0.9


def test_stable_dt_cantilever():
    p = L.cantilever_c1()
    c = math.sqrt(210e9 / 7850.0)
    assert abs(p.dt - 0.5 * (1 / 32) / c) < 1e-15
