"""The HIP CGSEM-2D solve on an ISM-V2 mesh file against the compiled
reference's solution on the same file (tests/golden/
ref_cgsem2dspec_annulus_n4.npz: curved quarter annulus, N = 4 GLL, zero
start, zero Dirichlet values, random k and source).  The mesh is the
library's SpecMesh reader's (bitwise the reference's,
tests/test_ref_cgsem2d_specmesh.py); from a zero start the reference's
as-built mask and the intended one reach the same solution, so the solved
field must agree to 1e-10 (relative max), as on the default mesh."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_cgsem2dspec_annulus_n4.npz")


def test_solve_on_spec_mesh_reaches_the_reference_solution(gpu_lib, tmp_path):
    from test_ref_cgsem2d_specmesh import read_on_gll
    from self_amd.krylov import CGsemElliptic2D, CGsemParams
    g = dict(np.load(GOLDEN))
    N = int(g["N"])
    print(f"cgsem2d on the annulus spec mesh, N={N}", flush=True)
    mesh = read_on_gll(g, tmp_path)
    p = CGsemParams(polyDeg=N, tolerance=float(g["tolerance"]), maximumIterates=int(g["maxIters"]))
    m = CGsemElliptic2D(p, mesh=mesh)
    shape = (m.nElems, N + 1, N + 1)
    T = lambda a: np.ascontiguousarray(a.transpose(2, 1, 0))  # (i,j,e) -> (e,j,i)  # noqa: E731
    m.SetFluxCoefficient(T(g["fluxCoeff"]))
    m.SetSource(T(g["source"]))
    m.SetDirichletBoundaryCondition(np.zeros(shape))
    m.SetSolution(np.zeros(shape))
    io, it, _ = m.Solve("cgsem_pcg", precond="none", tolerance=float(g["tolerance"]),
                        maxIters=int(g["maxIters"]))
    ref = T(g["sol_after_solve"])
    got = m.GetSolution()
    e = np.abs(got - ref).max() / np.abs(ref).max()
    assert e <= 1e-10, (e, io, it)
    m.Trash()
