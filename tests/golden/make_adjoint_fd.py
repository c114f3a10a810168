"""Central finite differences of the oracle for the adjoint gradient check.

    python tests/golden/make_adjoint_fd.py [--workers 8]

Case: lattice.bcc_adjoint_check() (64 struts, elastic, penalty contact, top
plate at -20 m/s), 1600 LeapFrog steps of the oracle (oracle/ref_cpu.c, grid
broadphase).  Objective: energy absorbed through the plate plus the peak
plate force, J = dt |v| sum_n F_n + max_n F_n with F_n = -reaction_n of the
top nodes along z.  Design variables: every strut's radius r_s (section
A = pi r^2, I1 = I2 = pi r^4 / 4, J = I1 + I2 and the contact radius all
follow r_s).  For each strut, dJ/dr_s by central differences with
h = 1e-6 r_s.  Writes adjoint_fd_bcc2.json next to this script.
"""
import argparse
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dynlattice_amd import lattice as L  # noqa: E402

NSTEPS = 1600
H_REL = 1e-6
V_PLATE = -20.0


def objective(reaction, dt):
    F = reaction * np.sign(V_PLATE)
    return dt * abs(V_PLATE) * F.sum() + F.max()


def run(radius):
    from oracle.ref_cpu import RefModel
    base = L.bcc_adjoint_check()
    p = L.bcc_adjoint_check(radius=radius, dt=base.dt)
    ref = RefModel(p, grid=True)
    o = ref.leapfrog(NSTEPS, group=p.groups["top"], group_dof=2)
    return objective(o["reaction"], p.dt)


def fd(s):
    r0 = L.bcc_adjoint_check().contact.radius
    h = H_REL * r0[s]
    rp = r0.copy(); rp[s] += h
    rm = r0.copy(); rm[s] -= h
    return (run(rp) - run(rm)) / (2 * h)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=8)
    a = ap.parse_args()
    p = L.bcc_adjoint_check()
    struts = list(range(p.n_struts))
    with Pool(a.workers) as pool:
        J0 = pool.apply_async(run, (p.contact.radius,))
        g = pool.map(fd, struts)
        J0 = J0.get()
    out = {"case": "lattice.bcc_adjoint_check()", "nsteps": NSTEPS, "h_rel": H_REL,
           "v_plate": V_PLATE, "dt": p.dt, "objective": J0, "struts": struts, "dJ_dr": g,
           "oracle": "oracle/ref_cpu.c, grid broadphase, LeapFrog"}
    with open(os.path.join(HERE, "adjoint_fd_bcc2.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"J = {J0:.9e}; |dJ/dr| range {min(map(abs, g)):.3e} .. {max(map(abs, g)):.3e}")


if __name__ == "__main__":
    main()
