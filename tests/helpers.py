"""Shared test cases: small problems with seeded, non-trivial states."""
import numpy as np

from dynlattice_amd import lattice as L


def small_bcc(n=2, m=3, plastic=False, contact=False, seed=7, exponent=1.0):
    R = 8 * n ** 3
    r = 0.5e-3 * (1.0 + 0.05 * L.normal(seed, R))
    p = L.bcc_lattice(n, m, 10e-3, r, 2e9, 0.3, 1200.0)
    if plastic:
        p.plastic = L.Plastic.resultant_capacities(r, 20e6, kin_ratio=0.01)
        p.plastic.exponent = (float(exponent),) * 6
        # the cutting plane crosses sum-abs corners slowly: raise maxIter
        # (a configuration property of the reference, ElastoPlasticRodMaterial.cpp:145)
        p.plastic.max_iter = 100
    if contact:
        p.contact = L.Contact(radius=r.copy(), penalty_sts=1e6, penalty_nts=1e6)
    L.crush_bcs(p, -5.0)
    p.dt = L.stable_dt(p, 0.5)
    return p


def random_state(p, seed=0, du=1e-4, dth=0.05, dv=0.0):
    rng = np.random.default_rng(seed)
    N = p.n_nodes
    d = np.zeros((N, 6))
    d[:, :3] = rng.normal(size=(N, 3)) * du
    d[:, 3:] = rng.normal(size=(N, 3)) * dth
    v = np.zeros((N, 6))
    if dv:
        v[:] = rng.normal(size=(N, 6)) * dv
    return d, v
