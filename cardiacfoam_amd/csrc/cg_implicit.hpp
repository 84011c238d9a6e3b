// Implicit monodomain step on the hex block (solveDiffusionImplicit,
// monodomainSolver.C:121-150): orthogonal part of fvm::laplacian as a symmetric
// 7-point operator, tensor correction explicit in the source, Jacobi-PCG.
#pragma once
#include <string>
#include "cg_internal.hpp"

namespace cg {

struct ImplicitSolver;

int implicit_create(ImplicitSolver** out, const HexGeom& g, const double* h_sigma9,
                    double chiCm, double dt, int prec, hipStream_t s, std::string* err);
void implicit_destroy(ImplicitSolver* S);
// ionic launch of the following step, queued behind the current solve (gated on its
// convergence): the step constants of that step
struct NextIonic {
    ExplicitStep st;
};
// one step: ionic (V^n) + right-hand side (unless ionic_queued: the previous call already
// queued it), solve for Vnew, activation update; with `next`, the next step's ionic kernel
// (V^{n+1}, V^n) is queued behind the read-back of this step's solver state.  extrapolate
// (mixed path): initial guess 2 V^n - V^{n-1} instead of V^n.
int implicit_step(ImplicitSolver* S, int model, int prec, int tissue, void* states,
                  const ModelParams& P, const double* Vn, const double* Vm1, double* Vnew,
                  double* Iion, uint8_t* flag, double* tact, const ExplicitStep& st, double tol,
                  int max_iter, Profiler* prof, bool ionic_queued, const NextIonic* next,
                  bool extrapolate, std::string* err);
void implicit_stats(const ImplicitSolver* S, int64_t* iters, double* res);
// fvSolution controls beyond tolerance and maxIter: norm 0 = relative L2, 1 = OpenFOAM's
// normalised residual (with rel_tol); PIMPLE nOuterCorrectors (n_outer <= 1: one solve)
// and the initial-residual tolerance residual_control of the outer loop
struct SolveControls {
    int norm = 0;
    double rel_tol = 0.0;
    int n_outer = 1;
    double residual_control = 0.0;
};
int implicit_set_controls(ImplicitSolver* S, const SolveControls& c);
// last step: linear solves (outer correctors), initial residual of the last corrector and
// final residual of the last solve, both in the controls' norm
void implicit_report(const ImplicitSolver* S, int* solves, double* init_res, double* final_res);
// y = A x, A = coefft chiCm/dt I + orthogonal Laplacian matrix; prec 0: double, 1: float
// (device pointers).  Returns 0, 9 when the fp32 operator was not built, 4 on HIP error.
int implicit_apply(ImplicitSolver* S, int prec, double coefft, const void* x, void* y,
                   Profiler* prof);

// standalone pieces used by tests through the C-ABI
hipError_t launch_hex_laplacian(const HexGeom& g, const HexCoefs& k, const double* V,
                                const double* od, unsigned od_mask, double* out, hipStream_t s);

} // namespace cg
