"""ISM-V2 SpecMesh files for the reader's tests (test infrastructure).

quarter_annulus() writes the file format ReadSpecMeshFile_QuadMesh reads
(src/geom/QuadMeshClass.f90:2100-2465): a version line, the counts, corner
positions, one line of six integers per edge, and per element its corners,
its four side flags, a curve of gPolyDeg+1 points on Chebyshev-Gauss-Lobatto
parameters -cos(j pi / gPolyDeg) for each flagged side, and the side names.

The mesh is the quarter annulus 1 <= r <= 2, 0 <= theta <= pi/2 in nr x nt
elements whose inner and outer arcs are curved sides.  Elements may list
their corners from any starting corner (rotation) or clockwise (mirror,
negative Jacobian), which changes which local side holds which curve and
gives edges whose second element runs the other way (negative side).
"""
import math

import numpy as np

# local side k runs from corner EDGE[k][0] to EDGE[k][1] (QuadMesh edgeMap)
EDGE = ((0, 1), (1, 2), (3, 2), (0, 3))


def quarter_annulus(nr, nt, g, seed=0, mirror=(), names=None):
    """Returns the text of the mesh file.  mirror: elements (0-based) listed
    clockwise.  names: side names for (inner arc, outer arc, theta = 0,
    theta = pi/2); interior sides are '---'."""
    rng = np.random.default_rng(seed)
    names = names or ("DIRICHLET", "NEUMANN_WALL", "ROBIN", "RADIATION")
    r = np.linspace(1.0, 2.0, nr + 1)
    th = np.linspace(0.0, 0.5 * math.pi, nt + 1)

    def node(i, j):
        return 1 + i + (nr + 1) * j

    s = -np.cos(np.arange(g + 1) * math.pi / g)
    lines = [" ISM-V2", ""]
    nodes = [(r[i] * math.cos(th[j]), r[i] * math.sin(th[j])) for j in range(nt + 1) for i in range(nr + 1)]
    for x, y in nodes:
        lines.append(f"{float(x)!r} {float(y)!r} 0.0")
    elems, edge_set = [], {}
    for j in range(nt):
        for i in range(nr):
            ccw = [node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)]
            # the counter-clockwise side from ccw[m] to ccw[m+1]: (curve or None, name)
            side = {}
            for m in range(4):
                a, b = ccw[m], ccw[(m + 1) % 4]
                curve, name = None, "---"
                if m == 0 and j == 0:
                    name = names[2]
                if m == 2 and j == nt - 1:
                    name = names[3]
                if m == 1 and i == nr - 1:
                    curve, name = (r[i + 1], th[j], th[j + 1]), names[1]
                if m == 3 and i == 0:
                    curve, name = (r[i], th[j + 1], th[j]), names[0]
                side[(a, b)] = (curve, name)
            e = len(elems)
            rot = int(rng.integers(4))
            order = [ccw[(k + rot) % 4] for k in range(4)]
            if e in mirror:
                order = [order[0], order[3], order[2], order[1]]
            flags, curves, sname = [], [], []
            for k in range(4):
                a, b = order[EDGE[k][0]], order[EDGE[k][1]]
                if (a, b) in side:
                    (curve, name), fwd = side[(a, b)], True
                else:
                    (curve, name), fwd = side[(b, a)], False
                sname.append(name)
                key = (min(a, b), max(a, b))
                edge_set.setdefault(key, []).append((e + 1, k + 1))
                if curve is None:
                    flags.append(0)
                    continue
                flags.append(1)
                R, t0, t1 = curve
                t = 0.5 * (s + 1.0)
                if not fwd:
                    t = 1.0 - t
                pts = [(R * math.cos(t0 + (t1 - t0) * u), R * math.sin(t0 + (t1 - t0) * u)) for u in t]
                curves.append(pts)
            elems.append((order, flags, curves, sname))
    for (n1, n2), users in edge_set.items():
        (e1, s1), (e2, s2) = users[0], users[1] if len(users) > 1 else (0, 0)
        lines.append(f"{n1:6d} {n2:6d} {e1:6d} {e2:6d} {s1:6d} {s2:6d}")
    for order, flags, curves, sname in elems:
        lines.append(" ".join(f"{v:11d}" for v in order))
        lines.append(" ".join(f"{v:11d}" for v in flags))
        for pts in curves:
            for x, y in pts:
                lines.append(f"{float(x)!r} {float(y)!r} 0.0")
        lines.append(" " + " ".join(sname) + " ")
    lines[1] = f"{len(nodes):12d} {len(edge_set):11d} {len(elems):11d} {g:11d}"
    return "\n".join(lines) + "\n"
