"""Reader of the GMSH .geo subset the reference's test cases use, and the
line mesher GMSHInputModule gets from the GMSH SDK (GMSHInputModule.cpp:
211-236 opens and meshes the file, :270-300 numbers the nodes, :305-420
builds one element group per curve).

Supported: ``DefineConstant[ s = {0.1, Name "size"} ];`` (overridable by
name through Input.input.onelab), assignments ``a = <expr>;``,
``Point(i) = {x, y, z, lc};``, ``Line(i) = {p, q};``, // and /* */ comments;
expressions in JEM/GMSH arithmetic syntax.

Meshing: every line is split into n equal elements with
n = floor(L / lc + 0.99) (at least 1), lc the points' characteristic length
(equal at both ends), GMSH's rule for uniform sizes (GMSH 4.9
meshGEdge.cpp; a third-party library not in the reference's sources:
exact for the integer L / lc of the reference's tests, parity unpinned
otherwise).  Node numbering follows GMSH's getNodes order as
GMSHInputModule stores it: the points by tag (0-based JIVE index =
position), then every line's interior nodes in line-tag order, from its
first point to its second.  Each line is one strut (element group
beam_<k>, k = 1, 2, ... in line order).

Second order (Input.input.order = 2) gives GMSH's 3-node line elements:
the n elements' end and midside nodes, 2n + 1 equally spaced nodes per
line.  The element here is the 2-node one, so a second-order mesh becomes
2n two-node elements on that same node set.  The interior nodes are
numbered along the line (GMSH lists the midside nodes after the end
nodes; no reference test reads an interior node of a second-order mesh).
"""
from __future__ import annotations

import re

import numpy as np


class GeoError(ValueError):
    pass


def _strip(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _num(expr, env):
    from .jemprops import number
    e = expr
    for k in sorted(env, key=len, reverse=True):
        e = re.sub(rf"\b{re.escape(k)}\b", f"({env[k]!r})", e)
    return number(e)


def read(path: str, onelab: dict | None = None):
    """points {tag: (x, y, z, lc)}, lines [(tag, p, q)]"""
    onelab = onelab or {}
    text = _strip(open(path).read())
    env = {}
    points, lines = {}, []
    for stmt in text.split(";"):
        s = stmt.strip()
        if not s:
            continue
        m = re.match(r"DefineConstant\s*\[(.*)\]\s*$", s, re.S)
        if m:
            for dm in re.finditer(r"([A-Za-z_]\w*)\s*=\s*\{\s*([^,}]+)\s*(?:,\s*Name\s*\"([^\"]*)\")?[^}]*\}",
                                  m.group(1)):
                name, val, label = dm.group(1), dm.group(2), dm.group(3)
                key = (label or name).split("/")[-1]
                env[name] = float(onelab[key]) if key in onelab else _num(val, env)
            continue
        m = re.match(r"Point\s*\(\s*(\d+)\s*\)\s*=\s*\{(.*)\}\s*$", s, re.S)
        if m:
            vals = [_num(v, env) for v in m.group(2).split(",")]
            if len(vals) not in (3, 4):
                raise GeoError(f"{path}: Point needs x, y, z[, lc]")
            points[int(m.group(1))] = tuple(vals + [0.0] * (4 - len(vals)))
            continue
        m = re.match(r"Line\s*\(\s*(\d+)\s*\)\s*=\s*\{\s*(\d+)\s*,\s*(\d+)\s*\}\s*$", s)
        if m:
            lines.append((int(m.group(1)), int(m.group(2)), int(m.group(3))))
            continue
        m = re.match(r"([A-Za-z_]\w*)\s*=\s*(.+)$", s, re.S)
        if m:
            env[m.group(1)] = _num(m.group(2), env)
            continue
        raise GeoError(f"{path}: unsupported statement: {s[:60]!r}")
    return points, sorted(lines)


def mesh(points: dict, lines: list, order: int = 1):
    """coords [N, 3], elem_nodes [E, 2], elem_strut [E], strut_names"""
    if order not in (1, 2):
        raise GeoError("line elements of order 1 or 2 only")
    tags = sorted(points)
    index = {t: i for i, t in enumerate(tags)}
    coords = [points[t][:3] for t in tags]
    en, es, names = [], [], []
    for k, (tag, p, q) in enumerate(lines):
        if p not in points or q not in points:
            raise GeoError(f"Line({tag}) uses an undefined point")
        a, b = np.array(points[p][:3]), np.array(points[q][:3])
        lp, lq = points[p][3], points[q][3]
        if lp <= 0 or abs(lp - lq) > 1e-12 * max(lp, lq):
            raise GeoError(f"Line({tag}): equal positive characteristic lengths required")
        n = max(1, int(np.linalg.norm(b - a) / lp + 0.99)) * order
        chain = [index[p]]
        for j in range(1, n):
            coords.append(tuple(a + (b - a) * (j / n)))
            chain.append(len(coords) - 1)
        chain.append(index[q])
        for j in range(n):
            en.append((chain[j], chain[j + 1]))
            es.append(k)
        names.append(f"beam_{k + 1}")
    return (np.array(coords, np.float64), np.array(en, np.int32), np.array(es, np.int32), names)
