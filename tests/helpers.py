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


def random_segments(n_struts=60, m=2, box=0.05, length=0.02, radius=1.5e-3, seed=11,
                    kS=1e6, kN=2e6, spread=3.0):
    """Randomly placed straight struts (m elements each) in a cube: many
    segment pairs within contact distance, covering the STS, NTS and
    speculative NTS branches of the narrowphase.  The initial state (used by
    INIT for the blacklist) spreads the struts apart by ``spread``, so the
    packed rest configuration (zero displacement) is full of new contacts."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, box, size=(n_struts, 3))
    d = rng.normal(size=(n_struts, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    b = a + length * d
    t = np.linspace(0, 1, m + 1)
    coords = (a[:, None, :] + t[None, :, None] * (b - a)[:, None, :]).reshape(-1, 3)
    base = np.arange(n_struts) * (m + 1)
    en = np.stack([(base[:, None] + np.arange(m)[None, :]).ravel(),
                   (base[:, None] + np.arange(1, m + 1)[None, :]).ravel()], 1).astype(np.int32)
    es = np.repeat(np.arange(n_struts, dtype=np.int32), m)
    r = np.full(n_struts, radius)
    p = L.Problem(coords=coords, elem_nodes=en, elem_strut=es, n_struts=n_struts,
                  sections=L.Sections.circle(r * 0.5, 2e9, 0.3, 1200.0))
    p.contact = L.Contact(radius=r, penalty_sts=kS, penalty_nts=kN)
    p.finalize()
    c = coords.mean(0)
    strut_mid = np.repeat(0.5 * (a + b), m + 1, axis=0)
    p.disp0[:, :3] = (spread - 1.0) * (strut_mid - c)
    p.dt = L.stable_dt(p, 0.5)
    return p
