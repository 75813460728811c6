"""ISM-V2 SpecMesh files: the library's reader against the reference's own.

ReadSpecMeshFile_QuadMesh (src/geom/QuadMeshClass.f90:2100-2465, with
ConstructEdges :1489-1627 and the transfinite-interpolation geometry of
MappedGeometryClass_2D.f90:536-710) is compiled unmodified
(oracle/ref.mk -> oracle/_ref/ref_specmesh).  The library's reader
(self_specmesh_read, host code) must build the same edge table, node types
and bitwise the same element geometry:

* tests/golden/ref_specmesh_annulus.npz holds a curved quarter annulus
  (tests/specmesh_gen.py: rotated corner orders, one clockwise element)
  and what the reference built from it; these tests run everywhere;
* the live tests run the reference on further generated meshes and on the
  ISM-V2 mesh files shipped with the reference, when both are present;
* the closed-form tests pin what the mathematics fixes without the
  reference: a straight-sided mesh is the bilinear map of its corners, the
  Jacobian integrates to the area, curved sides lie on their arcs, normals
  are unit and outward, and the edge table's orientation follows the
  corners.  Each is shown to reject a planted mistake.
These are host-only file operations; no GPU is used.
"""
import glob
import math
import os

import numpy as np
import pytest

import ref_run as R
import specmesh_gen as G
from self_amd import _lib as L
from self_amd.mesh import SpecMesh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_specmesh_annulus.npz")
REF_MESHES = [
    "src/highend/shallowwater/Testing/box-step.mesh",
    "src/geom/Testing/square.mesh",
    "src/geom/Testing/MeshFiles/CircleFlow.mesh",
    "examples/hydrostaticprimitive/kelvinwaves/run/circle.mesh",
    "examples/cgsem-elliptic2d/paraboloid/run/square.mesh",
    "examples/shallowwater/shelfsteepening/run/box.mesh",
]
REF_ROOT = "/root/reference"
FLAGS = {"DIRICHLET": -200, "NEUMANN_WALL": -206, "ROBIN": -202, "RADIATION": -101}


def nodes(N, quad):
    x, w = np.zeros(N + 1), np.zeros(N + 1)
    L.check(L.lib().self_legendre_quadrature(N, quad, L.ptr(x), L.ptr(w)))
    return x, w


def read(path, N, quad=1):
    return SpecMesh(path, N, nodes(N, quad)[0])


def mismatches(m, ref):
    """Names of the arrays that differ from the reference's (element-last)."""
    bad = []
    for k, ours in (("edges", m.edges.T), ("elemNodes", m.elemNodes.T), ("nodeType", m.nodeType),
                    ("nodeX", m.nodeX), ("nodeY", m.nodeY)):
        if not np.array_equal(ours, ref[k]):
            bad.append(k)
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        if not np.array_equal(getattr(m, k).transpose(2, 1, 0), ref[k]):
            bad.append(k)
    if not np.array_equal(m.nHat.transpose(3, 2, 1, 0), ref["nHat"]):
        bad.append("nHat")
    if not np.array_equal(m.nLen.transpose(2, 1, 0), ref["nLen"]):
        bad.append("nLen")
    return bad


def write(tmp_path, text, name="m.mesh"):
    p = tmp_path / name
    p.write_text(text)
    return p


# ---------------------------------------------------------------- golden
@pytest.mark.parametrize("N,quad,scale", [(3, 1, None), (5, -1, None), (4, 1, (2.0, 3.0))])
def test_reader_matches_the_reference_fixture(tmp_path, N, quad, scale):
    g = np.load(GOLDEN)
    path = tmp_path / "annulus.mesh"
    path.write_bytes(g["mesh"].tobytes())
    tag = f"{N}_{quad}_" + ("" if scale is None else "s_")
    ref = {k[len(tag):]: g[k] for k in g.files if k.startswith(tag) and "_" not in k[len(tag):]}
    m = read(path, N, quad)
    if scale is not None:
        m.ScaleTheMesh(*scale)
    assert (m.nElems, m.nEdges, m.nNodes) == (int(ref["nElems"]), int(ref["nEdges"]), int(ref["nNodes"]))
    assert mismatches(m, ref) == []
    # the fixture exercises reversed neighbours and a negative Jacobian
    assert (m.edges[:, 3] < 0).any() and (m.J < 0).any() and (m.J > 0).any()


def test_fixture_present():
    assert glob.glob(GOLDEN)


# ----------------------------------------------------------- live reference
live = pytest.mark.skipif(not R.available("specmesh"),
                          reason="oracle/_ref/ref_specmesh not built (needs the reference sources)")


@live
@pytest.mark.parametrize("nr,nt,g,seed,mirror,N,quad", [
    (2, 3, 1, 3, (), 4, 1),
    (4, 2, 7, 4, (0, 5), 6, -1),
    (1, 1, 3, 5, (), 2, 1),
    (3, 3, 5, 6, (4,), 9, -1),
])
def test_live_reference_generated(tmp_path, nr, nt, g, seed, mirror, N, quad):
    path = write(tmp_path, G.quarter_annulus(nr, nt, g, seed=seed, mirror=mirror))
    assert mismatches(read(path, N, quad), R.run_specmesh(path, N, quad)) == []


@live
@pytest.mark.skipif(not os.path.isdir(REF_ROOT), reason="reference mesh files absent")
@pytest.mark.parametrize("rel", REF_MESHES)
def test_live_reference_mesh_files(rel):
    path = os.path.join(REF_ROOT, rel)
    assert mismatches(read(path, 4), R.run_specmesh(path, 4, 1)) == []


@live
@pytest.mark.skipif(not os.path.isdir(REF_ROOT), reason="reference mesh files absent")
@pytest.mark.parametrize("rel,N,quad,scale", [
    ("examples/shallowwater/dipole/run/box.mesh", 10, 1, (4.0e5, 4.0e5)),  # the dipole run's own
    ("src/geom/Testing/MeshFiles/CircleFlow.mesh", 5, -1, (1.5, 0.5)),
])
def test_live_reference_scale_the_mesh(rel, N, quad, scale):
    path = os.path.join(REF_ROOT, rel)
    m = read(path, N, quad).ScaleTheMesh(*scale)
    assert mismatches(m, R.run_specmesh(path, N, quad, scale=scale)) == []


# --------------------------------------------------------------- closed form
def bilinear_error(m, s):
    """Largest deviation of x, y and the metric terms from the bilinear map
    of each element's corners (exact for straight sides)."""
    err = 0.0
    a = s[None, :]  # i -> s
    b = s[:, None]  # j -> p
    for e in range(m.nElems):
        c = m.elemNodes[e] - 1
        X, Y = m.nodeX[c], m.nodeY[c]
        for V, ours, dS, dP in ((X, m.x[e], m.dxds[e], m.dxdp[e]), (Y, m.y[e], m.dyds[e], m.dydp[e])):
            val = 0.25 * (V[0] * (1 - a) * (1 - b) + V[1] * (1 + a) * (1 - b) + V[2] * (1 + a) * (1 + b)
                          + V[3] * (1 - a) * (1 + b))
            ds = 0.25 * ((V[1] - V[0]) * (1 - b) + (V[2] - V[3]) * (1 + b)) + 0 * a
            dp = 0.25 * ((V[3] - V[0]) * (1 - a) + (V[2] - V[1]) * (1 + a)) + 0 * b
            err = max(err, np.abs(ours - val).max(), np.abs(dS - ds).max(), np.abs(dP - dp).max())
        det = m.dxds[e] * m.dydp[e] - m.dxdp[e] * m.dyds[e]
        err = max(err, np.abs(m.J[e] - det).max())
    return err


def straight_mesh(tmp_path, seed=2):
    """A distorted 3x2 quad mesh with straight sides (gPolyDeg 1), corners
    listed from rotated starting corners."""
    rng = np.random.default_rng(seed)
    nx, ny = 3, 2
    xs = np.linspace(0, 3, nx + 1)
    ys = np.linspace(0, 2, ny + 1)
    pts = [(x + 0.15 * rng.uniform(-1, 1), y + 0.15 * rng.uniform(-1, 1)) for y in ys for x in xs]
    idx = lambda i, j: 1 + i + (nx + 1) * j  # noqa: E731
    elems, edges = [], set()
    for j in range(ny):
        for i in range(nx):
            ccw = [idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)]
            r = int(rng.integers(4))
            order = ccw[r:] + ccw[:r]
            elems.append(order)
            for k in range(4):
                a, b = order[G.EDGE[k][0]], order[G.EDGE[k][1]]
                edges.add((min(a, b), max(a, b)))
    lines = [" ISM-V2", f"{len(pts)} {len(edges)} {len(elems)} 1"]
    lines += [f"{float(x)!r} {float(y)!r} 0.0" for x, y in pts]
    lines += [f"{a} {b} 0 0 0 0" for a, b in sorted(edges)]
    for o in elems:
        lines += [" ".join(map(str, o)), "0 0 0 0", " --- --- --- --- "]
    return write(tmp_path, "\n".join(lines) + "\n", "straight.mesh")


def test_straight_sides_are_the_bilinear_map(tmp_path):
    path = straight_mesh(tmp_path)
    s = nodes(5, -1)[0]
    m = read(path, 5, -1)
    assert bilinear_error(m, s) < 1e-14
    # planted mistakes: swapped metric terms, transposed element arrays
    m.dxdp[:], m.dyds[:] = m.dyds.copy(), m.dxdp.copy()
    assert bilinear_error(m, s) > 1e-3
    m = read(path, 5, -1)
    m.x[:] = m.x.transpose(0, 2, 1)
    assert bilinear_error(m, s) > 1e-3


def tangent_errors(m):
    """Largest |nHat . t| over the sides of a straight-sided mesh, t the
    side's direction from its (scaled) corner positions, and the largest
    |nLen - |t| / 2| (the side's half length: d/ds of the bilinear map)."""
    err = 0.0
    for e in range(m.nElems):
        c = m.elemNodes[e] - 1
        for k in range(4):
            a, b = c[G.EDGE[k][0]], c[G.EDGE[k][1]]
            t = np.array([m.nodeX[b] - m.nodeX[a], m.nodeY[b] - m.nodeY[a]])
            err = max(err, np.abs(m.nHat[e, k] @ t).max() / np.hypot(*t),
                      np.abs(m.nLen[e, k] - 0.5 * np.hypot(*t)).max() / np.hypot(*t))
    return err


def test_scaling_recomputes_normals(tmp_path):
    path = straight_mesh(tmp_path)
    s = nodes(4, 1)[0]
    m = read(path, 4)
    keep = m.nHat.copy()
    m.ScaleTheMesh(2.0, 3.0)
    assert tangent_errors(m) < 1e-13 and bilinear_error(m, s) < 1e-13
    # planted mistake: normals left as they were before scaling
    m.nHat[:] = keep
    assert tangent_errors(m) > 1e-2
    # uniform scaling: same directions, lengths scaled
    u = read(path, 4)
    u0 = read(path, 4)
    u.ScaleTheMesh(4.0e5, 4.0e5)
    assert np.allclose(u.nHat, u0.nHat, atol=1e-14)
    assert np.allclose(u.nLen, 4.0e5 * u0.nLen, rtol=1e-14)
    assert np.array_equal(u.J, (4.0e5 * 4.0e5) * u0.J)


def area_and_arc_errors(m, s, w):
    """|integral of |J| - 3 pi / 4| and the largest |r - R| of the nodes on
    the curved sides (inner arc R = 1, outer R = 2; Gauss-Lobatto nodes)."""
    area = float(np.sum(np.abs(m.J) * w[None, :, None] * w[None, None, :]))
    r = np.hypot(m.x, m.y)
    arc = 0.0
    for e in range(m.nElems):
        edge = [r[e, 0, :], r[e, :, -1], r[e, -1, :], r[e, :, 0]]  # south, east, north, west
        for k in range(4):
            rr = edge[k]
            for R in (1.0, 2.0):
                if np.all(np.abs(rr - R) < 1e-3):
                    arc = max(arc, np.abs(rr - R).max())
    return abs(area - 0.75 * math.pi), arc


def inward_normals(m):
    """(arc sides seen, arc sides whose normal is not the outward radial
    direction: -r on the inner arc, +r on the outer; Gauss-Lobatto nodes)."""
    seen = bad = 0
    for ed in m.edges:
        if ed[2] >= 0:
            continue
        e, k = ed[0] - 1, ed[1] - 1
        xs = [m.x[e, 0, :], m.x[e, :, -1], m.x[e, -1, :], m.x[e, :, 0]][k]
        ys = [m.y[e, 0, :], m.y[e, :, -1], m.y[e, -1, :], m.y[e, :, 0]][k]
        rr = np.hypot(xs, ys)
        radial = (m.nHat[e, k, :, 0] * xs + m.nHat[e, k, :, 1] * ys) / rr
        for R, sign in ((1.0, -1.0), (2.0, 1.0)):
            if np.all(np.abs(rr - R) < 1e-6):
                seen += 1
                bad += not np.all(sign * radial > 1 - 1e-6)
    return seen, bad


def test_annulus_area_arcs_and_normals(tmp_path):
    path = write(tmp_path, G.quarter_annulus(4, 6, 8, seed=9, mirror=(3, 10)))
    s, w = nodes(8, -1)
    m = read(path, 8, -1)
    darea, arc = area_and_arc_errors(m, s, w)
    assert darea < 1e-8 and arc < 1e-8  # degree-8 interpolation of the arcs
    assert np.allclose(np.hypot(m.nHat[..., 0], m.nHat[..., 1]), 1.0, atol=1e-15)
    assert inward_normals(m) == (12, 0)
    # planted mistake: normals without the Jacobian's sign turn inward on a
    # clockwise element (element 3 has a side on the outer arc)
    m.nHat[3] *= -1
    assert inward_normals(m)[1] > 0
    # planted mistake: a dropped 1/2 in a metric term changes the area
    m.J *= 0.5
    assert area_and_arc_errors(m, s, w)[0] > 1.0


def orientation_errors(m):
    """Edge table checked against the corners: every element side is on
    exactly one edge; the two sides of an interior edge join the same corner
    pair, and s2 < 0 exactly when the second element's side starts at the
    other corner than the first's (QuadMesh edgeMap); start/inc follow the
    sign."""
    seen = np.zeros((m.nElems, 4), dtype=int)
    errs = 0
    for ed in m.edges:
        e1, s1, e2, s2, st, inc = (int(v) for v in ed)
        seen[e1 - 1, s1 - 1] += 1
        a1, b1 = (m.elemNodes[e1 - 1, G.EDGE[s1 - 1][t]] for t in (0, 1))
        if e2 > 0:
            seen[e2 - 1, abs(s2) - 1] += 1
            a2, b2 = (m.elemNodes[e2 - 1, G.EDGE[abs(s2) - 1][t]] for t in (0, 1))
            errs += {a1, b1} != {a2, b2}
            errs += (s2 < 0) != (a2 != a1)
        errs += (st, inc) != ((m.N - 1, -1) if s2 < 0 else (1, 1))
    errs += int(np.abs(seen - 1).sum())
    return errs


def test_edge_table_follows_the_corners(tmp_path):
    path = write(tmp_path, G.quarter_annulus(3, 5, 3, seed=12, mirror=(2,)))
    m = read(path, 4)
    assert orientation_errors(m) == 0
    assert (m.edges[:, 3] < 0).sum() > 0
    # named sides: e2 is their flag, their corners take it as node type
    for name, flag in FLAGS.items():
        assert (m.edges[:, 2] == flag).any(), name
    assert set(np.unique(m.nodeType)) <= {1, *FLAGS.values()}
    assert (m.nodeType == 1).sum() == (3 - 1) * (5 - 1)  # interior corners
    # planted mistake: one flipped side sign
    m.edges[np.argmax(m.edges[:, 2] > 0), 3] *= -1
    assert orientation_errors(m) > 0


def test_fortran_exponents_and_crlf_read_the_same(tmp_path):
    """Node and curve lines written with Fortran's D exponent, and CRLF line
    ends, read as the plain file (list-directed READ accepts both)."""
    text = G.quarter_annulus(2, 3, 4, seed=8)
    lines = text.split("\n")
    nNodes = int(lines[1].split()[0])
    for i in range(2, 2 + nNodes):
        x, y, z = (float(v) for v in lines[i].split())
        lines[i] = f"{x:.17E} {y:.17E} {z:.1E}".replace("E", "D")
    a = read(write(tmp_path, text, "a.mesh"), 4)
    b = read(write(tmp_path, "\r\n".join(lines), "b.mesh"), 4)
    assert mismatches(b, {"edges": a.edges.T, "elemNodes": a.elemNodes.T, "nodeType": a.nodeType,
                          "nodeX": a.nodeX, "nodeY": a.nodeY, **{k: getattr(a, k).transpose(2, 1, 0) for k in
                          ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J", "nLen")},
                          "nHat": a.nHat.transpose(3, 2, 1, 0)}) == []


# --------------------------------------------------------------- bad input
def test_rejects_files_it_cannot_read(tmp_path):
    text = G.quarter_annulus(2, 2, 2, seed=1)
    with pytest.raises(L.SelfError, match="ISM-V2"):
        read(write(tmp_path, "\n".join(text.split("\n")[1:])), 3)  # no version line
    lines = text.split("\n")
    nNodes, nEdges, nElems, g = (int(v) for v in lines[1].split())
    first_elem = 2 + nNodes + nEdges
    lines[1] = f"{nNodes} {nEdges + 1} {nElems} {g}"
    lines.insert(first_elem, "1 2 1 0 1 0")  # one edge line too many
    with pytest.raises(L.SelfError, match="lists 13 edges, its elements have 12"):
        read(write(tmp_path, "\n".join(lines), "e.mesh"), 3)
    lines = text.split("\n")
    lines[first_elem + 1] = "2 0 0 0"
    with pytest.raises(L.SelfError, match="flag 2"):
        read(write(tmp_path, "\n".join(lines), "f.mesh"), 3)
    with pytest.raises(L.SelfError, match="ends before"):
        read(write(tmp_path, "\n".join(text.split("\n")[: first_elem + 3]), "t.mesh"), 3)
    with pytest.raises(L.SelfError, match="cannot open"):
        read(tmp_path / "missing.mesh", 3)
