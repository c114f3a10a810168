"""The allocation-free step (dl_explicit_step_noalloc) against dl_explicit_step:
two calls in a row on the GPU's shared scratch array, the second with a
smaller lattice and fewer steps, each equal bit for bit to the ordinary
entry point on a twin context (final state, reaction, dissipation and
contact-pair histories); sizes beyond the scratch are refused."""
import numpy as np
import pytest

from tests.helpers import small_bcc

pytestmark = pytest.mark.gpu


def _case(n, contact):
    p = small_bcc(n=n, m=3, plastic=True, exponent=2.0, contact=contact)
    p.velo0[p.groups["top"], 2] = -40.0
    return p


def _pair(p):
    from dynlattice_amd.gpu import DL_TRACK_DISSIPATION, GpuModel
    out = []
    for _ in range(2):
        g = GpuModel(p, flags=DL_TRACK_DISSIPATION)
        g.set_reaction_groups(np.r_[p.groups["top"], p.groups["bottom"]],
                              [0, len(p.groups["top"]), len(p.groups["top"]) + len(p.groups["bottom"])], 2)
        out.append(g)
    return out


def test_noalloc_step_equals_explicit_step():
    cases = [(_case(3, True), 120), (_case(2, False), 40)]  # second call: smaller input
    for p, K in cases:
        a, b = _pair(p)
        a.step(K)
        ha = a.history(K)
        ra = a.group_history(K)
        hb = b.step_noalloc(K)
        da, va = a.get_state()
        db, vb = b.get_state()
        print(p.n_elems, "elements", K, "steps; contacts", ha["n_contacts"].max(),
              "dissipation", ha["e_diss"][-1])
        assert ha["e_diss"][-1] > 0
        assert np.array_equal(da, db) and np.array_equal(va, vb)
        assert np.array_equal(hb["reaction"], ra)
        assert np.array_equal(hb["e_diss"], ha["e_diss"])
        assert np.array_equal(hb["n_contacts"], ha["n_contacts"])
        assert np.array_equal(a.plastic_active(), b.plastic_active())
        a.close()
        b.close()


def test_noalloc_step_refuses_more_rows_than_the_scratch():
    from dynlattice_amd.gpu import DL_ERR_INPUT, NOALLOC_SCRATCH_DOUBLES, DLError, GpuModel
    p = _case(2, False)
    g = GpuModel(p)
    too_many = NOALLOC_SCRATCH_DOUBLES // 4 + 1  # one group: 4 doubles per step
    with pytest.raises(DLError) as e:
        g.step_noalloc(too_many)
    assert e.value.code == DL_ERR_INPUT
    g.step_noalloc(8)  # still usable afterwards
    g.close()
