// lduMatrix (diag + symmetric face coefficients in lduAddressing order) on an arbitrary
// mesh: ELL rows on the device, y = A x, and PCG with OpenFOAM's residual normalisation.
#pragma once
#include <string>
#include "cg_common.hpp"

namespace cg {

struct LduSolver;

// lower[f] / upper[f]: the two cells of face f (lduAddressing lowerAddr / upperAddr)
int ldu_create(LduSolver** out, int64_t n_cells, int64_t n_faces, const int32_t* lower,
               const int32_t* upper, std::string* err);
void ldu_destroy(LduSolver* L);
int64_t ldu_rows(const LduSolver* L);
int ldu_width(const LduSolver* L);
// diag[n], upper[n_faces] (lower == upper); device or host pointers
int ldu_set_coeffs(LduSolver* L, const double* diag, const double* upper, bool device);
// y = A x; with events, the kernel alone is bracketed (device pointers)
int ldu_amul(LduSolver* L, const double* x, double* y, bool device, hipEvent_t ev0,
             hipEvent_t ev1);
int ldu_solve(LduSolver* L, double* psi, const double* b, bool device, double tol, double rel,
              int max_iter, int* iters, double* init_res, double* final_res);

} // namespace cg
