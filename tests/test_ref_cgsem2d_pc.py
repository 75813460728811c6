"""The oracle's restatement of the reference's FEM preconditioner
(CGsemPreconditioner_2D_Class.f90: the linear mesh through the Gauss-Lobatto
nodes of ConstructLinearMeshFromOther, QuadMeshClass.f90:1863-2096, and an
inner CG of pcIterates iterations to pcTolerance, :664-746) against the
reference itself, compiled unmodified by oracle/ref.mk, through the
preconditioned CG of CGsemElliptic_2D (:1114-1207, the call at :1169).

With the reference's as-built mask (quirk Q6) the oracle follows the
reference's solve: the same ioerr and iteration count, the residual history
and the solution to round-off (the outer operator runs on the oracle's own
mesh, equal to the reference's to about 1e-14; the preconditioner is built
from the reference's node positions).  Fixtures tests/golden/ref_cgsempc_*.npz
come from tests/golden/make_ref_cgsem2d.py; the live test runs the reference
binary when it is built.  A closed-form pin of the preconditioner itself: run
to convergence it inverts the N = 1 operator, assembled independently here
from the bilinear element formula and solved densely.
"""
import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po
import ref_run as R

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_cgsempc_*.npz")))
F = lambda a: np.ascontiguousarray(np.asfortranarray(a).ravel(order="F"))  # noqa: E731


def _oracle_solve(g, as_built=True):
    N, nX, nY = int(g["N"]), int(g["nX"]), int(g["nY"])
    o = po.CGSEM2D(N, nX, nY, tolerance=float(g["tolerance"]), maxIters=int(g["maxIters"]), q6_as_built=as_built)
    o.s.dirichletValue = 0.0
    o.fluxCoeff[:] = F(g["fluxCoeff"])
    o.source[:] = F(g["source"])
    o.solution[:] = 0.0
    o.build_preconditioner(F(g["x"]), F(g["y"]), pcTolerance=float(g["pcTolerance"]),
                           pcIterates=int(g["pcIterates"]))
    io, it, pit = o.solve_fem()
    return o, io, it, pit


def _check(g):
    o, io, it, pit = _oracle_solve(g)
    n_ref = int(np.count_nonzero(g["resi"])) - 1
    assert io == int(g["ioerr"]) and it == n_ref, (io, it, int(g["ioerr"]), n_ref)
    assert np.abs(o.resi[:it + 1] - g["resi"][:it + 1]).max() <= 1e-13 * g["resi"][0]
    ref = F(g["sol_after_solve"])
    assert np.abs(o.solution - ref).max() <= 1e-13 * np.abs(ref).max()
    assert pit > it  # the inner CG ran


def test_fixtures_present():
    assert len(GOLDEN) >= 2


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_preconditioned_solve_follows_the_reference(path):
    _check(dict(np.load(path)))


@pytest.mark.skipif(not R.available("cgsem2d"), reason="oracle/_ref/ref_cgsem2d not built (needs the reference tree)")
def test_live_reference_run():
    inp = R.random_inputs_cgsem2d(3, 3, 4, 31)
    inp["solution"][:] = 0.0
    _check(R.run_cgsem2d(3, 3, 4, 300, 1e-10, 0.0, inp, pcIterates=7, pcTolerance=1e-3))


def test_converged_preconditioner_inverts_the_linear_operator():
    """pcIterates large and pcTolerance tiny: the preconditioner output w of
    b solves A1 w = b on the interior nodes, A1 the assembled N = 1 operator
    built independently here from the element formula (bilinear elements,
    two-point Gauss-Lobatto rule)."""
    N, nX, nY = 3, 3, 2
    o = po.CGSEM2D(N, nX, nY, tolerance=1e-12, maxIters=10)
    n1, E = N + 1, o.nElems
    x = o.mesh.x.reshape(E, n1, n1)
    y = o.mesh.y.reshape(E, n1, n1)
    o.build_preconditioner(o.mesh.x, o.mesh.y, pcTolerance=1e-15, pcIterates=500)
    # global node numbering by position
    ids, pos = {}, []
    gid = np.zeros((E, n1, n1), dtype=int)
    for e in range(E):
        for j in range(n1):
            for i in range(n1):
                k = (round(x[e, j, i], 12), round(y[e, j, i], 12))
                if k not in ids:
                    ids[k] = len(pos)
                    pos.append(k)
                gid[e, j, i] = ids[k]
    nG = len(pos)
    A = np.zeros((nG, nG))
    # bilinear element with corners c (SW, SE, NE, NW); two-point GLL rule at the corners, weights 1
    for e in range(E):
        for j in range(N):
            for i in range(N):
                cn = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
                cx = np.array([x[e, b, a] for a, b in cn])
                cy = np.array([y[e, b, a] for a, b in cn])
                g = [gid[e, b, a] for a, b in cn]
                for q, (s, p) in enumerate([(-1, -1), (1, -1), (1, 1), (-1, 1)]):
                    dNs = 0.25 * np.array([-(1 - p), (1 - p), (1 + p), -(1 + p)])
                    dNp = 0.25 * np.array([-(1 - s), -(1 + s), (1 + s), (1 - s)])
                    xs, xp, ys, yp = dNs @ cx, dNp @ cx, dNs @ cy, dNp @ cy
                    Jq = xs * yp - xp * ys
                    gx = (yp * dNs - ys * dNp) / Jq
                    gy = (xs * dNp - xp * dNs) / Jq
                    K = (np.outer(gx, gx) + np.outer(gy, gy)) * Jq
                    for a_ in range(4):
                        for b_ in range(4):
                            A[g[a_], g[b_]] -= K[a_, b_]  # the reference's operator is -div(grad)
    xs_all = np.array([p[0] for p in pos])
    ys_all = np.array([p[1] for p in pos])
    interior = ~((np.isclose(xs_all, 0) | np.isclose(xs_all, 1)) | (np.isclose(ys_all, 0) | np.isclose(ys_all, 1)))
    Ai = A[np.ix_(interior, interior)]
    rng = np.random.default_rng(5)
    bg = np.zeros(nG)
    bg[interior] = rng.standard_normal(interior.sum())
    b = bg[gid].ravel()
    w, io, it = o.precondition(b)
    wg = np.zeros(nG)
    wg[gid.ravel()] = w
    ref = np.linalg.solve(Ai, bg[interior])
    assert np.abs(wg[interior] - ref).max() <= 1e-10 * np.abs(ref).max(), it
    assert np.all(wg[~interior] == 0.0)
