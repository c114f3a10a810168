"""The all-cores oracle (libref_cpu_omp.so, the CPU baseline on every host
core) computes what the serial oracle computes: per-strut accumulation then
per-node sums in strut order changes only the association of the sums at
shared nodes, so trajectories agree to rounding and discrete sets exactly."""
import os

import numpy as np

from oracle.ref_cpu import RefModel
from tests.conftest import rel_inf
from tests.helpers import small_bcc


def test_omp_oracle_matches_serial_oracle():
    os.environ.setdefault("OMP_NUM_THREADS", "4")
    p = small_bcc(n=2, m=3, plastic=True, contact=True, exponent=2.0)
    p.velo0[p.groups["top"], 2] = -40.0
    outs = []
    for omp in (False, True):
        r = RefModel(p, omp=omp, grid=True)
        o = r.leapfrog(200, group=p.groups["bottom"])
        outs.append((r.disp.copy(), r.velo.copy(), o["reaction"], o["n_contacts"],
                     r.plastic_state()[0].copy()))
    (d0, v0, r0, c0, e0), (d1, v1, r1, c1, e1) = outs
    assert rel_inf(d1, d0) <= 1e-11
    assert rel_inf(v1, v0) <= 1e-9
    assert rel_inf(r1, r0) <= 1e-10
    assert np.array_equal(c1, c0)
    assert rel_inf(e1, e0) <= 1e-9
