"""The 2M-element elastoplastic crush (bcc-40x40x40, 4 elements per strut,
contact on) at its full size on the GPU, through properties that do not
depend on the size:

* the internal force (elements plus contact) is self-equilibrated: the
  translational forces and the moments about the origin sum to zero;
* at interior strut nodes (two elements, no contact) the device's nodal
  force equals the sum of the oracle's element forces at the same state,
  within the 1e-11 bar of the force parity tests, on 3000 seeded nodes;
* evaluations and steps are bitwise reproducible, and a run replayed from
  captured graphs equals direct launches bit for bit after 24 steps.
The state is the rest state perturbed by seeded displacements (1 um) and
rotations (1e-4 rad), which keeps contact and plastic flow off, so that the
sampled nodes see element forces only."""
import numpy as np
import pytest

from dynlattice_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    return L.bcc_c3(40, contact=True)


def _state(p, seed=21):
    rng = np.random.default_rng(seed)
    d = np.zeros((p.n_nodes, 6))
    d[:, :3] = rng.normal(size=(p.n_nodes, 3)) * 1e-6
    d[:, 3:] = rng.normal(size=(p.n_nodes, 3)) * 1e-4
    return d


def test_fullsize_force_balance_and_element_parity(c3):
    from dynlattice_amd.gpu import GpuModel
    from oracle.ref_cpu import RefModel
    p = c3
    assert p.n_elems == 2_048_000 and p.n_nodes == 1_668_921
    d = _state(p)
    g = GpuModel(p)
    g.set_state(d, np.zeros_like(d))
    f = g.int_vector()
    f2 = g.int_vector()
    g.close()
    assert np.array_equal(f, f2)
    scale = np.abs(f).sum(axis=0)
    # translational balance
    assert np.all(np.abs(f[:, :3].sum(axis=0)) <= 1e-12 * scale[:3]), f[:, :3].sum(axis=0)
    # moment balance about the origin: sum of x_n x f_n + m_n over all nodes
    x = p.coords + d[:, :3]
    mom = np.cross(x, f[:, :3]).sum(axis=0) + f[:, 3:].sum(axis=0)
    mscale = np.abs(np.cross(x, f[:, :3])).sum(axis=0) + scale[3:]
    assert np.all(np.abs(mom) <= 1e-12 * mscale), mom
    # interior strut nodes: element k's end b plus element k+1's end a
    en = p.elem_nodes
    same = p.elem_strut[:-1] == p.elem_strut[1:]
    inner = np.nonzero(same & (en[:-1, 1] == en[1:, 0]))[0]
    pick = np.random.default_rng(5).choice(inner, 3000, replace=False)
    import copy
    q = copy.copy(p)
    q.contact = None  # element forces only; the oracle's contact set-up is not needed
    r = RefModel(q)
    dv = np.ascontiguousarray(d.reshape(-1))
    ref = np.array([r.elem_fint(e, dv)[1] + r.elem_fint(e + 1, dv)[0] for e in pick])
    got = f[en[pick, 1]]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("sampled interior nodes: max rel diff", err)
    assert err <= 1e-11


def test_fullsize_steps_are_reproducible_and_graphs_equal_direct(c3):
    from dynlattice_amd.gpu import GpuModel
    p = c3
    out = []
    for graph in (0, 0, 8):
        g = GpuModel(p)
        g.set_reaction_group(p.groups["bottom"], 2)
        g.set_graph(graph)
        g.step(24)
        out.append(g.get_state() + (g.history(24)["reaction"], g.plastic_active()))
        g.close()
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(a, b)
