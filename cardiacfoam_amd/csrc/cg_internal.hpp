// Internal launch interface between the C-ABI (capi.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "models.hpp"

namespace cg {

constexpr int MAX_BOXES = 8;

// Box stimulus active at one step (myocardiumDomain.C:307-342): cells whose centre lies
// inside [lo, hi] (inclusive, boundBox::contains) receive the summed intensity (A/m^3).
struct ActiveBoxes {
    int n;
    double lo[MAX_BOXES][3];
    double hi[MAX_BOXES][3];
    double intensity[MAX_BOXES];
};

struct HexGeom {
    int nx, ny, nz;
    double hx, hy, hz;
};

// Coefficients of the orthogonal FV Laplacian per unit volume on the hex block:
// face between c and c+e_d carries a_d = (n.sigma_f.n) A_d / (h_d V).  Either uniform
// scalars (ax, ay, az) or per-face arrays indexed by the lower cell (0 on the wall).
struct HexCoefs {
    double ax, ay, az;
    const double* fx;
    const double* fy;
    const double* fz;
};

struct ExplicitStep {
    double dt;         // s
    double dt_model;   // model time units
    double chiCm;      // F/m^3
    double c, c0, c00; // ddt coefficients (Euler: 1,1,0; backward: 1.5,2,0.5)
    double t_old;      // time at the step start (activation interpolation)
    ActiveBoxes boxes;
    const double* corr;  // explicit tensor correction div (per unit volume) or null
};

// ionic step (tissue mode) on SoA states; Vm in V, Im in V/s; prec 0 = fp64, 1 = fp32
hipError_t launch_ionic_tissue(int model, int prec, int tissue, void* states, int64_t n,
                               const double* Vm, double* Im, double dt_model,
                               const ModelParams& P, const void* gscale, hipStream_t s);

// single-cell steps (V inside the ODE) for n independent cells, n_steps steps of
// dt_model starting at model time t0; V trace of the last step not stored.
hipError_t launch_ionic_cells(int model, int prec, int tissue, void* states, int64_t n,
                              double t0_model, double dt_model, int n_steps,
                              const ModelParams& P, const void* gscale, const Stimulus& stim,
                              hipStream_t s);

// fused explicit monodomain step on the hex block: ionic + diffusion + activation
hipError_t launch_explicit_hex(int model, int prec, int tissue, const HexGeom& g,
                               const HexCoefs& k, void* states, const double* V,
                               const double* Vold, double* Vnew, double* Iion, uint8_t* flag,
                               double* tact, const ExplicitStep& st, const ModelParams& P,
                               hipStream_t s);

// tensor correction div(SfGammaCorr . interp(grad V)) on the hex block (sigma9 per cell)
hipError_t launch_hex_correction(const HexGeom& g, const double* sigma9, const double* V,
                                 double* grad3, double* out, hipStream_t s);

} // namespace cg

#include <vector>
#include <string>

namespace cg {

// Per-kernel HIP-event timing on the handle's stream (the trace equivalent of the
// clockTime phases of staggeredElectrophysicsAdvanceScheme.C:54-74).
enum KernelId { K_IONIC = 0, K_SPMV, K_UPDATE, K_PUPDATE, K_RESIDUAL, K_CORR, K_ACTIVATION,
                K_INVDIAG, K_ADDCORR, K_COUNT };
const char* kernel_name(int id);

struct Profiler {
    bool enabled = false;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Rec { int id; hipEvent_t a, b; };
    std::vector<Rec> recs;
    double total_ms[K_COUNT] = {0};
    long long launches[K_COUNT] = {0};

    hipEvent_t next()
    {
        if (used == pool.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    int begin(int id)
    {
        if (!enabled) return -1;
        if (used + 2 > 8192) resolve(true);
        Rec r{id, next(), next()};
        hipEventRecord(r.a, stream);
        recs.push_back(r);
        return (int)recs.size() - 1;
    }
    void end(int tok)
    {
        if (tok >= 0) hipEventRecord(recs[tok].b, stream);
    }
    size_t mark() const { return recs.size(); }
    // drop records from index i on (launches that exited early on a converged flag)
    void invalidate_from(size_t i)
    {
        for (size_t k = i; k < recs.size(); ++k) recs[k].id = -1;
    }
    // accumulate recorded intervals; sync first when asked
    void resolve(bool sync)
    {
        if (recs.empty()) return;
        if (sync) hipStreamSynchronize(stream);
        for (auto& r : recs) {
            if (r.id < 0) continue;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                total_ms[r.id] += ms;
                launches[r.id] += 1;
            }
        }
        recs.clear();
        used = 0;
    }
    void reset()
    {
        resolve(true);
        for (int i = 0; i < K_COUNT; ++i) { total_ms[i] = 0; launches[i] = 0; }
    }
    ~Profiler()
    {
        for (auto e : pool) hipEventDestroy(e);
    }
};

} // namespace cg
