// fempc.hpp -- the CG spectral-element solver's preconditioner on the device.
//
// CGsemPreconditioner_2D (src/cgsem/CGsemPreconditioner_2D_Class.f90): the
// same operator, -div(k grad u), discretised at N = 1 on the linear mesh
// whose corners are the Gauss-Lobatto nodes of the original mesh
// (ConstructLinearMeshFromOther_QuadMesh, src/geom/QuadMeshClass.f90:
// 1863-2096), inverted approximately by pcIterates iterations of plain CG to
// the relative tolerance pcTolerance (Solve, :664-746).
//
// Here the linear operator is assembled once, over the interior global
// degrees of freedom of the original solver (the same numbering), into an
// ELL matrix stored slot-major (slot k of every row contiguous, so a
// wavefront's row reads coalesce); the element matrices are the reference's
// FluxDivergence at N = 1 (two-point Gauss-Lobatto rule, straight-sided
// bilinear sub-elements).  The inner CG keeps its scalars on the device and
// runs as two launches per iteration (each workgroup sums the previous
// launch's partials itself, in one fixed order, so all hold the same bits);
// every kernel returns at once when the inner CG has converged or the outer
// solver is done, so the host issues pcIterates iterations without reading
// anything back.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace self {

struct SolverState;

struct PcState {
    double num, den, alpha, beta, r0, rNorm, tol;
    double numP[2];  // r.r at the start of iteration it, in slot it & 1
    int iter, done, maxIters, iterations;
};

constexpr int kPcParts = 512;  // workgroups of the reduction kernels

struct FemPC {
    size_t n = 0;  // interior global degrees of freedom
    int width = 0; // ELL slots per row
    int maxIters = 0;
    double tol = 0;
    int *col = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
    double *r = nullptr, *z = nullptr, *v[2] = {nullptr, nullptr};
    float *r32 = nullptr, *z32 = nullptr, *v32[2] = {nullptr, nullptr};
    double *partials = nullptr;  // [2][kPcParts]: product step (v.z), update (r.r)
    PcState *st = nullptr;
    size_t bytes = 0;
    long long applications = 0, innerTotal = 0;  // host counters (set by readback only)

    ~FemPC();
    // Assemble from the element-local node positions x, y and coefficient k
    // (nElems x (N+1)^2, element-local layout) and the local -> global map
    // gid (-1 on Dirichlet nodes) with nG interior degrees of freedom.
    int build(int N, size_t nElems, const double *x, const double *y, const double *k, const int *gid, size_t nG,
              int pcIterates, double pcTolerance);
    // x = M^-1 b (FP64 or FP32 arithmetic); outer: the calling solver's
    // state (its done / innerExit flags end the work early), may be null.
    int apply(const double *b, double *x, const SolverState *outer, hipStream_t s);
    int apply32(const float *b, float *x, const SolverState *outer, hipStream_t s);
    // inner iterations of the last application (synchronizes the stream)
    int last_iterations(hipStream_t s, int &its);
};

}  // namespace self
