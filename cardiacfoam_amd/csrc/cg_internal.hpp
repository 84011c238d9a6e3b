// Internal launch interface between the C-ABI (capi.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "models.hpp"

namespace cg {

constexpr int MAX_BOXES = 8;

// Box stimulus active at one step (myocardiumDomain.C:307-342): cells whose centre lies
// inside [lo, hi] (inclusive, boundBox::contains) receive the summed intensity (A/m^3).
// On a hex block the centres are the (i+1/2)h lattice unless the caller gave the mesh's
// own centres: then mask[c] holds bit b for every box b that contains cell c's centre
// (computed once on the host) and bit[q] names the box of active slot q.
struct ActiveBoxes {
    int n;
    double lo[MAX_BOXES][3];
    double hi[MAX_BOXES][3];
    double intensity[MAX_BOXES];
    const uint8_t* mask;
    uint8_t bit[MAX_BOXES];
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

// Per-cell ODE integrator of the ionic step: Rush-Larsen (method 0), or OpenFOAM's
// adaptive RKF45 (1) or Euler (2) with the reference's tolerances and per-cell trial step.
struct OdeArgs {
    int method;
    double t0, t1;     // model time at the step start and end (tStart, tEnd)
    double dt_clamp;   // deltaT in model units; a cell's trial step is min(step, dt_clamp)
    double* step;      // [n] per-cell trial step in model time units (RKF45)
    int* fail;         // set to 1 by a cell whose integration underflowed or ran out of steps
    double abs_tol, rel_tol;
    long long max_steps;
    // single cells over several steps: PDE time t_s (s) advanced by dt_s per step, the
    // step start taken as (t + dt) - dt like Time's accumulated value; ap: time / 12.9
    double t_s, dt_s;
    int ap;
};

struct ExplicitStep {
    double dt;         // s
    double dt_model;   // model time units
    double chiCm;      // F/m^3
    double c, c0, c00; // ddt coefficients (Euler: 1,1,0; backward: 1.5,2,0.5)
    double t_old;      // time at the step start (activation interpolation)
    ActiveBoxes boxes;
    const double* od;  // off-diagonal conductivity (xy, xz, yz) [3][n], or null
    unsigned od_mask;  // which of xy, xz, yz are non-zero somewhere (bits 0..2)
    // launch gate: when gate != null the kernel runs only if *gate == gate_val (a solver
    // generation written by an earlier kernel of the stream)
    const long long* gate;
    long long gate_val;
    OdeArgs ode;
    // implicit path: the previous step's activation-time update (myocardiumDomain.C:
    // 345-374), done by this step's ionic kernel, which reads V^n and V^{n-1} anyway;
    // act_flag null: none
    uint8_t* act_flag;
    double* act_tact;
    double act_t_old, act_dt;
    // activation threshold (V): the first crossing of V > act_v + 1e-15, interpolated
    // linearly to act_v; 0 is the reference's (myocardiumDomain.C:362-371)
    double act_v = 0.0;
};

// Model kernels, one set per arithmetic type: _fp64 (IEEE double, model_fp64.hip) and
// _fp32 (single precision with hardware reciprocal/exp2/log2, model_fp32.hip).
#define CG_DECLARE_MODEL_ENTRIES(SFX)                                                      \
    /* ionic step (tissue mode) on SoA states; Vm in V, Im in V/s */                       \
    hipError_t launch_ionic_tissue_##SFX(int model, int tissue, void* states, int64_t n,   \
                                         const double* Vm, double* Im, double dt_model,    \
                                         const ModelParams& P, const void* gscale,         \
                                         const OdeArgs& ode, hipStream_t s);               \
    /* n independent single cells (V inside the ODE), n_steps steps from model time t0;  \
     * with trace, cell tcell stores its states after every step, [n_steps][NS] */         \
    hipError_t launch_ionic_cells_##SFX(int model, int tissue, void* states, int64_t n,    \
                                        double t0, double dt, int n_steps,                 \
                                        const ModelParams& P, const void* gscale,          \
                                        const Stimulus& stim, const OdeArgs& ode,          \
                                        double* trace, int64_t tcell, hipStream_t s);      \
    /* fused explicit monodomain step on the hex block: ionic + diffusion + activation */  \
    hipError_t launch_explicit_hex_##SFX(int model, int tissue, const HexGeom& g,          \
                                         const HexCoefs& a, void* states, const double* V, \
                                         const double* Vold, double* Vnew, double* Iion,   \
                                         uint8_t* flag, double* tact,                      \
                                         const ExplicitStep& st, const ModelParams& P,     \
                                         hipStream_t s);                                   \
    /* ionic step + right-hand side of the implicit monodomain system */                   \
    hipError_t launch_ionic_rhs_##SFX(int model, int tissue, const HexGeom& g,             \
                                      void* states, const double* Vn, const double* Vm1,   \
                                      double* x0, double* Iion, double* b,                 \
                                      const ExplicitStep& st, const ModelParams& P,        \
                                      hipStream_t s);

CG_DECLARE_MODEL_ENTRIES(fp64)
CG_DECLARE_MODEL_ENTRIES(fp32)

// detector state, FP64 SoA [field][lane]
enum {
    DS_V1, DS_V2, DS_V3, DS_DVP,  // V_{i-1}, V_{i-2}, V_{i-3}, dv_{i-2}
    DS_CUR,                       // current beat: 8 fields from here
    DS_OLD = DS_CUR + 8,          // closing beat
    DS_N = DS_OLD + 8
};
enum { BF_TU, BF_TEND, BF_BASE, BF_PEAK, BF_JP, BF_V90, BF_TREP, BF_FLAGS };
enum { BEAT_OPEN = 1, BEAT_HAS = 2, BEAT_FOUND = 4 };
constexpr int BEAT_REC = 5;  // t_up, v_base, v_peak, v90, t_repol90 (NaN when not found)
// RATES (kind 0) / ALGEBRAIC (kind 1) of a model at n host-given rows, FP64 (model_fp64.hip);
// model_nalgebraic: the length of the model's algebraic list on the device (0: none)
int model_nalgebraic(int model);
hipError_t launch_ionic_eval(int model, int tissue, int64_t n, const double* st, const double* t,
                             int kind, int solve_v, const Stimulus& stim, const ModelParams& P,
                             const double* gsc, double* out, hipStream_t s);
// per-beat flags of the restitution classification (cg_ionic_get_restitution)
constexpr unsigned RB_EARLY = 1, RB_VALID_APD = 2, RB_VALID_DI = 4, RB_TRUNCATED = 8;
// beats [n][max_beats][BEAT_REC] (model units), nbeats [n] -> out [n][2] (APD90, DI in s,
// NaN when rejected), flags [n][max_beats] (or null); asynchronous on s
hipError_t launch_restitution(const double* beats, const int* nbeats, int64_t n, int max_beats,
                              double inv_per_s, double* out, uint8_t* flags, hipStream_t s);

struct BatchArgs {
    int64_t n;             // lanes (= cells)
    int64_t s0;            // steps already done
    int chunk;             // steps of this launch
    double dt;             // model time units
    const int64_t* order;  // lane -> cell
    const int64_t* len;    // [cell] run length in steps
    const Stimulus* prot;  // [cell] protocol, or null: `stim` for every cell
    Stimulus stim;
    // beat detection (det == null: off)
    double* det;           // [DS_N][n]
    void* ring;            // [ring][n] samples of T
    int ring_len;
    double* beats;         // [cell][max_beats][BEAT_REC]
    int* nbeats;           // [cell]
    int max_beats;
    double thr, baseline, peak_window;  // mV per model time unit, model time units
    double vscale, voff;                // V (mV) = y[0] * vscale - voff
    // integrator: 0 Rush-Larsen; 1 RKF45, 2 OpenFOAM's Euler, each step [s dt, (s+1) dt]
    // from the cell's kept trial step (clamped to dt), as cg_ionic_run_cells does
    int method;
    double abs_tol, rel_tol;
    long long max_steps;
    double* step;          // [cell] kept trial steps (model time units)
    int* fail;             // set to 1 by a cell whose integration failed
};

// batched single-cell protocol launch (batch_kernels.hpp); variant = kernel shape
hipError_t launch_cells_protocol_fp64(int model, int tissue, int variant, void* states,
                                      const ModelParams& P, const void* gscale,
                                      const BatchArgs& a, hipStream_t s);
hipError_t launch_cells_protocol_fp32(int model, int tissue, int variant, void* states,
                                      const ModelParams& P, const void* gscale,
                                      const BatchArgs& a, hipStream_t s);
constexpr int BATCH_VARIANTS = 6;
const char* batch_variant_name(int v);

// precision dispatch: prec 0 = fp64, 1 = fp32
inline hipError_t launch_ionic_tissue(int model, int prec, int tissue, void* states, int64_t n,
                                      const double* Vm, double* Im, double dt_model,
                                      const ModelParams& P, const void* gscale,
                                      const OdeArgs& ode, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    return (prec ? launch_ionic_tissue_fp32 : launch_ionic_tissue_fp64)(
        model, tissue, states, n, Vm, Im, dt_model, P, gscale, ode, s);
}

inline hipError_t launch_ionic_cells(int model, int prec, int tissue, void* states, int64_t n,
                                     double t0_model, double dt_model, int n_steps,
                                     const ModelParams& P, const void* gscale,
                                     const Stimulus& stim, const OdeArgs& ode, hipStream_t s,
                                     double* trace = nullptr, int64_t tcell = -1)
{
    if (n <= 0 || n_steps <= 0) return hipSuccess;
    return (prec ? launch_ionic_cells_fp32 : launch_ionic_cells_fp64)(
        model, tissue, states, n, t0_model, dt_model, n_steps, P, gscale, stim, ode, trace,
        tcell, s);
}

inline hipError_t launch_explicit_hex(int model, int prec, int tissue, const HexGeom& g,
                                      const HexCoefs& a, void* states, const double* V,
                                      const double* Vold, double* Vnew, double* Iion,
                                      uint8_t* flag, double* tact, const ExplicitStep& st,
                                      const ModelParams& P, hipStream_t s)
{
    return (prec ? launch_explicit_hex_fp32 : launch_explicit_hex_fp64)(
        model, tissue, g, a, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
}

inline hipError_t launch_ionic_rhs(int model, int prec, int tissue, const HexGeom& g,
                                   void* states, const double* Vn, const double* Vm1,
                                   double* x0, double* Iion, double* b,
                                   const ExplicitStep& st, const ModelParams& P, hipStream_t s)
{
    return (prec ? launch_ionic_rhs_fp32 : launch_ionic_rhs_fp64)(
        model, tissue, g, states, Vn, Vm1, x0, Iion, b, st, P, s);
}

} // namespace cg

#include <vector>
#include <string>

namespace cg {

// Per-kernel HIP-event timing on the handle's stream (the trace equivalent of the
// clockTime phases of staggeredElectrophysicsAdvanceScheme.C:54-74).
enum KernelId { K_IONIC = 0, K_SPMV, K_UPDATE, K_RESIDUAL0, K_RESIDUAL, K_ACTIVATION,
                K_INVDIAG, K_APPLY, K_ECG, K_COUNT };
const char* kernel_name(int id);

struct Profiler {
    bool enabled = false;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> all, free_;
    struct Rec { int id; hipEvent_t a, b; };
    std::vector<Rec> recs;
    double total_ms[K_COUNT] = {0};
    long long launches[K_COUNT] = {0};

    hipEvent_t next()
    {
        if (free_.empty()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            all.push_back(e);
            return e;
        }
        hipEvent_t e = free_.back();
        free_.pop_back();
        return e;
    }
    int begin(int id)
    {
        if (!enabled) return -1;
        if (recs.size() >= 4096) resolve(true);
        Rec r{id, next(), next()};
        if (!r.a || !r.b || hipEventRecord(r.a, stream) != hipSuccess) r.id = -1;
        recs.push_back(r);
        return (int)recs.size() - 1;
    }
    void end(int tok)
    {
        if (tok >= 0 && hipEventRecord(recs[tok].b, stream) != hipSuccess) recs[tok].id = -1;
    }
    size_t mark() const { return recs.size(); }
    // drop records [i, j) (launches that exited early on a finished solve or a closed gate)
    void invalidate_range(size_t i, size_t j)
    {
        for (size_t k = i; k < j && k < recs.size(); ++k) recs[k].id = -1;
    }
    // accumulate the records whose launches have completed (all of them after a sync when
    // asked), keep the others; record indices are valid until the next resolve
    void resolve(bool sync)
    {
        if (recs.empty()) return;
        if (sync && hipStreamSynchronize(stream) != hipSuccess) return;
        std::vector<Rec> keep;
        for (auto& r : recs) {
            if (hipEventQuery(r.b) == hipErrorNotReady) {
                keep.push_back(r);
                continue;
            }
            float ms = 0.f;
            if (r.id >= 0 && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                total_ms[r.id] += ms;
                launches[r.id] += 1;
            }
            if (r.a) free_.push_back(r.a);
            if (r.b) free_.push_back(r.b);
        }
        recs.swap(keep);
    }
    void reset()
    {
        resolve(true);
        for (int i = 0; i < K_COUNT; ++i) { total_ms[i] = 0; launches[i] = 0; }
    }
    ~Profiler()
    {
        for (auto e : all) hipEventDestroy(e);
    }
};

} // namespace cg
