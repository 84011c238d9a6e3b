// Host interface of the adjoint of the explicit monodomain step (adjoint.hip).
#pragma once
#include <string>
#include <vector>
#include "cg_internal.hpp"

namespace cg {

struct Adjoint;

// stimulus boxes of a tissue (all of them, with their time windows)
struct ActiveBoxesAll {
    int n;
    double lo[MAX_BOXES][3], hi[MAX_BOXES][3], t0[MAX_BOXES], dur[MAX_BOXES], I[MAX_BOXES];
    const uint8_t* mask;  // device box-membership mask from the mesh's centres, or null
};

// Snapshot of the initial state (V0 [n], ionic states [4][n] of BuenoOrovio, FP64) and the
// kappa = 1 conductivity diagonal; probes: cell ids; implicit: backward-Euler diffusion
// (Euler ddt) instead of the explicit update.  Returns 0 or an error code.
int adjoint_create(Adjoint** out, const HexGeom& g, const double* sigma0_diag, double dt,
                   double chiCm, const ModelParams& P, const double* V0, const void* y0_states,
                   const ActiveBoxesAll& boxes, int checkpoint_every, int n_probes,
                   const int64_t* probes, bool implicit, hipStream_t s, std::string* err);
void adjoint_destroy(Adjoint* A);
// n_steps from the initial state with conductivity kappa * sigma0 (device [n]) and the
// constants theta[k] at indices theta_index[k]; Vm after each step at the probes into
// probes [n_steps][n_probes] (device); t0: start time of every step.  When the checkpoint
// interval covers the run and every step's state fits in store_budget bytes, all states
// are kept and the backward pass re-runs nothing.
int adjoint_forward(Adjoint* A, const double* kappa, const double* theta, const int* theta_index,
                    int n_theta, int n_steps, const std::vector<double>& t0, double* probes,
                    size_t store_budget, std::string* err);
// dLoss/dprobes (device [n_steps][n_probes]) -> dLoss/dkappa (device [n]) and
// dLoss/d(tau_fi, tau_si) (host [2]).
int adjoint_backward(Adjoint* A, const double* dprobes, double* dkappa, double* dtheta,
                     std::string* err);

} // namespace cg
