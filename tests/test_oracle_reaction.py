"""The oracle's reaction history is GroupOutput's (GroupOutputModule.cpp:
100-116,162): after each step (or MilneDevice attempt), with LOAD_CASE
"output", the model's internal force at the new state -- elastic predictor
from the committed plastic state (SpecialCosseratRodModel.cpp:762) plus the
contact forces of a fresh search (RodContactModel.cpp:153-166) -- summed over
the group's dof.

Pinned against the oracle's GET_INT_VECTOR (itself pinned by closed forms
and invariants in test_oracle_element.py) evaluated step by step at the
state each step ends in, and against the two mistakes a restatement can make:
the start state instead of the end state (row 0 of a crush from rest would
be exactly zero) and the trial instead of the output load case (differs as
soon as the material yields)."""
import numpy as np

from oracle.ref_cpu import RefModel, milne_opts
from tests.helpers import small_bcc


def _group_sum(f, group, dof):
    return float(f[group, dof].sum())


def _stepwise(p, K, inelastic, group, dof=2):
    """K single steps of the oracle, after each the group sum of its
    GET_INT_VECTOR at the new state"""
    r = RefModel(p)
    out = []
    for _ in range(K):
        r.leapfrog(1)
        out.append(_group_sum(r.int_vector(inelastic=inelastic), group, dof))
    return np.array(out), r


def test_reaction_rows_are_the_output_state_sums_elastic():
    p = small_bcc(n=2, m=3, contact=True)
    top = p.groups["top"]
    K = 60
    o = RefModel(p).leapfrog(K, group=top, group_dof=2)["reaction"]
    ref, _ = _stepwise(p, K, inelastic=True, group=top)
    assert np.abs(o).max() > 0
    assert np.allclose(o, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    # row 0 is the state after the first step: the lattice starts at rest and
    # unstressed, so a start-state sum would be exactly zero there
    assert o[0] != 0.0
    r0 = RefModel(p)
    assert _group_sum(r0.int_vector(), top, 2) == 0.0


def test_reaction_rows_use_the_output_load_case_plastic():
    p = small_bcc(n=2, m=3, plastic=True, exponent=2.0)
    p.velo0 *= 8.0  # a faster crush, so that struts yield within the run
    top = p.groups["top"]
    K = 150
    rr = RefModel(p)
    o = rr.leapfrog(K, group=top, group_dof=2)["reaction"]
    out_case, r1 = _stepwise(p, K, inelastic=False, group=top)
    assert np.allclose(o, out_case, rtol=1e-13, atol=1e-13 * np.abs(out_case).max())
    # the material yielded, and the trial load case would give other values
    ep = rr.plastic_state()[0]
    assert np.abs(ep).max() > 0
    trial, _ = _stepwise(p, K, inelastic=True, group=top)
    assert np.abs(trial - out_case).max() > 1e-8 * np.abs(out_case).max()


def test_milne_rows_are_the_sums_after_each_attempt():
    p = small_bcc(n=2, m=3, contact=True)
    top = p.groups["top"]
    n = 40
    opts = milne_opts(p.dt, 1e-4)
    o = RefModel(p).milne(n, opts, group=top, group_dof=2)
    # attempt by attempt: the sum at the state each attempt leaves
    q = RefModel(p)
    ref = []
    dt = p.dt
    for _ in range(n):
        a = q.milne(1, milne_opts(dt, 1e-4, dt_min=p.dt / 100, dt_max=p.dt * 100))
        dt = float(a["dt_next"][0])
        ref.append(q.output_reaction(top, 2))
    ref = np.array(ref)
    assert o["accepted"].sum() > 0
    assert np.allclose(o["reaction"], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    # and it is the GET_INT_VECTOR of the output load case at that state
    f = q.int_vector(inelastic=False)
    assert abs(_group_sum(f, top, 2) - ref[-1]) <= 1e-13 * abs(ref[-1])
