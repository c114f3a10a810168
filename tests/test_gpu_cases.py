"""The reference's transient test 2 end to end on the GPU: the case files
(tests/golden/transient/test2.pro + includes + test2.geo) are read, the
swinging rod runs with MilneDevice to t = 1 s as the case prescribes
(additive rotation update, sampling "t % 1e-4 < deltaTime"), disp.gz is
written in the CSVOutput layout, and the reference's validator
(tests/transient/test2.py, restated below on the same pandas calls) accepts
it: every snapshot of Lang et al. 2011 within curve distance 0.05.

The case leaves the precision at JIVE's library default, which is not in
the reference's sources; the run here uses 1e-4 (the oracle shows the
comparison unchanged from 1e-4 to 1e-6)."""
import os

import numpy as np
import pytest

from tests.test_oracle_milne import curve_distance_2d

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _validate(disp_path, ref_path):
    """tests/transient/test2.py's reading and comparison"""
    import pandas as pd
    data = pd.read_csv(disp_path, index_col=["time", "label"]).xs("disp", level="label")
    data.columns = pd.MultiIndex.from_tuples(
        [(n[:n.find("[")], int(n[n.find("[") + 1:n.find("]")])) for n in data.columns],
        names=["dof", "node"])
    dx = data.xs("dx", axis=1, level="dof").copy()
    dz = data.xs("dz", axis=1, level="dof").copy()
    end_dx, end_dz = dx[1].copy(), dz[1].copy()
    nmax = data.columns.unique("node").max()
    for i in range(1, nmax):
        dx[i] = dx[i + 1]
        dz[i] = dz[i + 1]
    dx[nmax] = end_dx
    dz[nmax] = end_dz
    lines = open(ref_path).read().splitlines()
    times = lines[0].split(",")
    rows = [ln.split(",") for ln in lines[2:]]
    dists = []
    n_nodes = len(dx.columns)
    for j in range(0, len(times), 2):
        t_val = float(times[j][1:])
        pts = np.array([(float(r[j]), float(r[j + 1])) for r in rows
                        if j + 1 < len(r) and r[j] != "" and r[j + 1] != ""])
        if t_val == 0.0:
            dx.loc[0.0] = [0.0] * len(dx.columns)
            dz.loc[0.0] = [0.0] * len(dz.columns)
        if t_val not in dx.index:
            dx.loc[t_val] = np.nan
            dx = dx.sort_index().interpolate()
        if t_val not in dz.index:
            dz.loc[t_val] = np.nan
            dz = dz.sort_index().interpolate()
        sim_x = (dx.loc[t_val] + 1 / (n_nodes - 1) * np.array(dx.columns.unique())).values
        sim_z = dz.loc[t_val].values
        dists.append(curve_distance_2d(np.column_stack([sim_x, sim_z]), pts))
    return dists, float(max(data.index))


def test_transient_test2_case_on_gpu(tmp_path):
    from dynlattice_amd import case as CS
    c = CS.load(os.path.join(HERE, "golden", "transient", "test2.pro"),
                case_name="tests/transient/test2")
    paths, t_end, steps = CS.run(c, out_dir=str(tmp_path), precision=1e-4)
    print("t_end", t_end, "steps", steps)
    assert t_end > 1.0
    dists, last = _validate(paths["disp"], os.path.join(HERE, "golden", "transient_test2_ref.csv"))
    print("snapshot distances", np.round(dists, 4))
    assert last > 0.9 and len(dists) == 11
    assert max(dists) <= 0.05
