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


# Comparison metrics of the reference's transient validators, restated from
# tests/metrics.py:8-26 (interp_on_reference, relative_L2).
def interp_on_reference(x_sim, y_sim, x_ref):
    """linear interpolation of (x_sim, y_sim), sorted by x, at x_ref"""
    order = np.argsort(x_sim)
    return np.interp(x_ref, x_sim[order], y_sim[order])


def relative_L2(y_sim, y_ref):
    """||y_sim - y_ref|| / ||y_ref|| (the largest |y_ref| when the norm is
    essentially zero)"""
    denom = np.linalg.norm(y_ref)
    if denom < 1e-12:
        denom = max(np.max(np.abs(y_ref)), 1e-12)
    return float(np.linalg.norm(y_sim - y_ref) / denom)


def transient_test1_errors(sim, golden):
    """tests/transient/test1.py's comparison with Simo & Vu-Quoc 1988 (Ex.
    5.1, the spin-up of a flexible beam): sim rows are (free node dx .. rz,
    hub rz, t); returns the relative L2 errors of the hub angle psi, the
    rotational deviation psi - rz(tip) in degrees, and the tip's axial and
    tangential deviations from the rigid shadow beam, each at the
    reference's abscissae."""
    import pandas as pd

    def ref(name, csv=False):
        path = f"{golden}/transient_test1_ref_{name}.csv"
        if csv:
            return np.loadtxt(path, delimiter=",")
        return pd.read_csv(path, delimiter=";", decimal=",").values
    bc, ang, u1, u2 = ref("BC", True), ref("angle"), ref("u1"), ref("u2")
    t = sim[:, -1]
    psi = np.unwrap(sim[:, -2])
    px, py = 10 + sim[:, 0], sim[:, 1]
    rz = np.unwrap(sim[:, 5])
    delta = np.stack([px - 10 * np.cos(psi), py - 10 * np.sin(psi)], -1)
    out = np.stack([np.cos(psi), np.sin(psi)], -1)
    tan = np.stack([-np.sin(psi), np.cos(psi)], -1)
    dev_rz = np.rad2deg(psi - rz)
    dev_u1 = (delta * out).sum(axis=1)
    dev_u2 = (delta * tan).sum(axis=1)
    return {"psi": relative_L2(interp_on_reference(t, psi, bc[:, 0]), bc[:, 1]),
            "rz": relative_L2(interp_on_reference(t, dev_rz, ang[:, 0]), ang[:, 1]),
            "u1": relative_L2(interp_on_reference(t, dev_u1, u1[:, 0]), u1[:, 1]),
            "u2": relative_L2(interp_on_reference(t, dev_u2, u2[:, 0]), u2[:, 1])}


def energy_rise(time, load, e_tot):
    """tests/transient/test3.py's energy criterion: the largest relative rise
    of the total energy between samples while the load is zero (-inf when
    it never rises)"""
    d = np.diff(e_tot) / e_tot[1:]
    rise = (d > 0) & (load[1:] == 0)
    return float(d[rise].max()) if rise.any() else float("-inf")


def hub_struts(seed=41, n_outer=12, m=3, radius=2.5e-3, spine_m=12, n_branch=6, kS=1e6,
               kN=2e6):
    """Struts with crowded joints: a hub shared by n_outer struts (more than
    eight struts at one node), a ring of struts between the outer ends, and
    a spine of spine_m elements with n_branch struts leaving from its
    interior nodes (a strut with more than four joint nodes).  Exercises the
    contact filter "no node of X in Y's strut" at joints of every size."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_outer, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    centre = np.array([0.02, 0.02, 0.02])
    outer = centre + 0.015 * d
    coords = [centre] + list(outer)
    en, es = [], []

    def strut(i, j, k, n_el):
        a, b = coords[i], coords[j]
        prev = i
        for t in range(1, n_el):
            coords.append(a + (b - a) * t / n_el)
            en.append((prev, len(coords) - 1))
            es.append(k)
            prev = len(coords) - 1
        en.append((prev, j))
        es.append(k)

    k = 0
    for i in range(n_outer):
        strut(0, 1 + i, k, m)
        k += 1
    for i in range(0, n_outer, 2):  # ring pieces between outer ends
        strut(1 + i, 1 + (i + 1) % n_outer, k, m)
        k += 1
    # spine through the hub region with branches from its interior nodes
    a = len(coords)
    coords.append(centre + np.array([-0.018, 0.004, -0.003]))
    b = len(coords)
    coords.append(centre + np.array([0.018, -0.003, 0.004]))
    first = len(coords)
    strut(a, b, k, spine_m)
    k += 1
    inner = list(range(first, first + spine_m - 1))
    for q, n in enumerate(inner[1::2][:n_branch]):
        coords.append(coords[n] + 0.01 * rng.normal(size=3))
        strut(n, len(coords) - 1, k, m)
        k += 1
    coords = np.array(coords)
    en = np.array(en, dtype=np.int32)
    es = np.array(es, dtype=np.int32)
    r = np.full(k, radius)
    p = L.Problem(coords=coords, elem_nodes=en, elem_strut=es, n_struts=k,
                  sections=L.Sections.circle(r * 0.5, 2e9, 0.3, 1200.0))
    p.contact = L.Contact(radius=r, penalty_sts=kS, penalty_nts=kN)
    p.finalize()
    p.dt = L.stable_dt(p, 0.5)
    return p
