// C-ABI of include/cardiacgpu.h: handles, argument checks, host-side time stepping,
// and launches of the gfx950 kernels.  No torch types, no global state except the
// thread-local error message.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <cmath>

#include "../../include/cardiacgpu.h"
#include "cg_internal.hpp"
#include "cg_implicit.hpp"
#include "cg_adjoint.hpp"
#include "cg_ldu.hpp"
#include "foam_format.hpp"
#include "cg_ecg.hpp"
#include "cg_general.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(CG_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));           \
    } while (0)

constexpr int kNumModels = 11;
const char* const kModelKeys[kNumModels] = {"BuenoOrovio", "TNNP", "ORd", "AlievPanfilov",
                                            "monodomainFDAManufactured", "Courtemanche",
                                            "PerisYague_2022", "Stewart", "Grandi",
                                            "ToRORd_dynCl", "Gaur"};

// state names, order of the *_STATES_NAMES tables / Names.H enums
const char* const kBONames[4] = {"u", "v", "w", "s"};
const char* const kAPNames[2] = {"u", "recovery_r"};
const char* const kFDANames[4] = {"V", "u1", "u2", "u3"};
const char* const kTNNPNames[17] = {"V", "K_i", "Na_i", "Ca_i", "Xr1", "Xr2", "Xs", "m", "h",
                                    "j", "d", "f", "fCa", "s", "r", "Ca_SR", "g"};
const char* const kORdNames[41] = {
    "membrane_V", "CaMK_CaMKt", "Na_i", "Na_ss", "K_i", "K_ss", "Ca_ss", "Ca_nsr", "Ca_jsr",
    "Ca_i", "INa_m", "INa_hf", "INa_hs", "INa_j", "INa_hsp", "INa_jp", "INaL_mL", "INaL_hL",
    "INaL_hLp", "Ito_a", "Ito_iF", "Ito_iS", "Ito_ap", "Ito_iFp", "Ito_iSp", "ICaL_d",
    "ICaL_ff", "ICaL_fs", "ICaL_fcaf", "ICaL_fcas", "ICaL_jca", "ICaL_ffp", "ICaL_fcafp",
    "ICaL_nca", "IKr_xrf", "IKr_xrs", "IKs_xs1", "IKs_xs2", "IK1_xk1", "ryr_Jrelnp",
    "ryr_Jrelp"};
const char* const kCRNNames[21] = {
    "membrane_V", "sodium_Nai", "potassium_Ki", "calcium_Cai", "calcium_CaUp", "calcium_CaRel",
    "ina_m", "ina_h", "ina_j", "ito_oa", "ito_oi", "ikur_ua", "ikur_ui", "ikr_xr", "iks_xs",
    "ical_d", "ical_f", "ical_fCa", "cajsr_u", "cajsr_v", "cajsr_w"};
const char* const kPYSNames[22] = {
    "membrane_V", "sodium_Nai", "potassium_Ki", "chloride_Cli", "calcium_Cai", "calcium_CaUp",
    "calcium_CaRel", "ina_m", "ina_h", "ina_j", "ikur_ua", "ikur_uif", "ikur_uis", "ikr_xr",
    "iks_xs", "ical_d", "ical_f", "ical_fCa", "iclca_qCa", "ryr_u", "ryr_v", "ryr_w"};
const char* const kSTWNames[20] = {
    "membrane_V", "Ihyperpolarization_activated_current_y_gate_y",
    "Irapid_time_dependent_potassium_current_Xr1_gate_Xr1",
    "Irapid_time_dependent_potassium_current_Xr2_gate_Xr2",
    "Islow_time_dependent_potassium_current_Xs_gate_Xs", "Ifast_sodium_current_m_gate_m",
    "Ifast_sodium_current_h_gate_h", "Ifast_sodium_current_j_gate_j",
    "IL_type_Ca_current_d_gate_d", "IL_type_Ca_current_f_gate_f",
    "IL_type_Ca_current_f2_gate_f2", "IL_type_Ca_current_fCass_gate_fCass",
    "Itransient_outward_current_s_gate_s", "Itransient_outward_current_r_gate_r", "Ca_i",
    "Ca_SR", "Ca_ss", "R_prime", "Na_i", "K_i"};
const char* const kGRDNames[41] = {
    "membrane_V", "Ca_i", "Ca_jn", "Ca_sl", "Ca_sr", "Csqn", "Ical_d", "Ical_f",
    "Ical_fCaB_jn", "Ical_fCaB_sl", "Ikr_xr", "Iks_xs", "Ikur_ikur_r", "Ikur_s", "Ina_h",
    "Ina_j", "Ina_m", "Inal_hl", "Inal_ml", "Ito_x", "Ito_y", "K_i", "ryr_i", "ryr_o",
    "ryr_ryr_r", "Na_i", "Na_jn", "Na_sl", "buffCa_CaM", "buffCa_Myoc", "buffCa_Myom",
    "buffCa_SLH_jn", "buffCa_SLH_sl", "buffCa_SLL_jn", "buffCa_SLL_sl", "buffCa_SRB",
    "buffCa_TnCHc", "buffCa_TnCHm", "buffCa_TnCL", "buffNa_NaB_jn", "buffNa_NaB_sl"};
const char* const kTORNames[45] = {
    "V", "CaMKt", "Nai", "Nass", "Ki", "Kss", "Cass", "Cansr", "Cajsr", "Cai", "Cli", "Clss",
    "INa_m", "INa_h", "INa_j", "INa_hp", "INa_jp", "INaL_mL", "INaL_hL", "INaL_hLp", "Ito_a",
    "Ito_iF", "Ito_iS", "Ito_ap", "Ito_iFp", "Ito_iSp", "ICaL_d", "ICaL_ff", "ICaL_fs",
    "ICaL_fcaf", "ICaL_fcas", "ICaL_jca", "ICaL_ffp", "ICaL_fcafp", "ICaL_nca_ss",
    "ICaL_nca_i", "IKr_C1", "IKr_C2", "IKr_C3", "IKr_I", "IKr_O", "IKs_xs1", "IKs_xs2",
    "Jrel_np", "Jrel_p"};
const char* const kGAURNames[29] = {
    "cell_v", "nai", "nass", "ki", "kss", "cai", "cai2", "cass", "cansr", "cajsr", "cacsr",
    "I_Na_m", "I_Na_h", "I_Na_j", "INaL_ml", "INaL_hl", "ICaL_d", "ICaL_fca", "IKr_xr",
    "IKs_xs1", "IKs_xs2", "ITo_aa", "CICR_Jrel2", "CICR_Jrel1", "CaMK_CaMKt", "CICR_tjsrol",
    "CICR_A", "ICaL_fs", "ICaL_ff"};

struct ModelInfo {
    int ns, np;
    double vscale, voff, iscale, tscale;
};

bool model_info(int model, ModelInfo* mi)
{
    switch (model) {
    case cg::MODEL_BO:
        *mi = {cg::BuenoOrovio::NS, cg::BuenoOrovio::NP, cg::BuenoOrovio::VSCALE,
               cg::BuenoOrovio::VOFF, cg::BuenoOrovio::ISCALE, cg::BuenoOrovio::TSCALE};
        return true;
    case cg::MODEL_TNNP:
        *mi = {cg::TNNP::NS, cg::TNNP::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_ORD:
        *mi = {cg::ORd::NS, cg::ORd::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_AP:
        *mi = {cg::AlievPanfilov::NS, cg::AlievPanfilov::NP, cg::AlievPanfilov::VSCALE,
               cg::AlievPanfilov::VOFF, cg::AlievPanfilov::ISCALE, cg::AlievPanfilov::TSCALE};
        return true;
    case cg::MODEL_FDA:
        *mi = {cg::FdaManufactured::NS, cg::FdaManufactured::NP, 1.0, 0.0, 1.0, 1.0};
        return true;
    case cg::MODEL_CRN:
        *mi = {cg::Courtemanche::NS, cg::Courtemanche::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_PYS:
        *mi = {cg::PerisYague::NS, cg::PerisYague::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_STW:
        *mi = {cg::Stewart::NS, cg::Stewart::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_GRD:
        *mi = {cg::Grandi::NS, cg::Grandi::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_TOR:
        *mi = {cg::ToRORd::NS, cg::ToRORd::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    case cg::MODEL_GAUR:
        *mi = {cg::Gaur::NS, cg::Gaur::NP, 1.0, 0.0, 1.0, 1000.0};
        return true;
    default: return false;
    }
}

bool model_host_init(int model, int tissue, double* p, double* y)
{
    switch (model) {
    case cg::MODEL_BO: return cg::BuenoOrovio::host_init(tissue, p, y);
    case cg::MODEL_TNNP: return cg::TNNP::host_init(tissue, p, y);
    case cg::MODEL_ORD: return cg::ORd::host_init(tissue, p, y);
    case cg::MODEL_AP: return cg::AlievPanfilov::host_init(tissue, p, y);
    case cg::MODEL_FDA: return cg::FdaManufactured::host_init(tissue, p, y);
    case cg::MODEL_CRN: return cg::Courtemanche::host_init(tissue, p, y);
    case cg::MODEL_PYS: return cg::PerisYague::host_init(tissue, p, y);
    case cg::MODEL_STW: return cg::Stewart::host_init(tissue, p, y);
    case cg::MODEL_GRD: return cg::Grandi::host_init(tissue, p, y);
    case cg::MODEL_TOR: return cg::ToRORd::host_init(tissue, p, y);
    case cg::MODEL_GAUR: return cg::Gaur::host_init(tissue, p, y);
    default: return false;
    }
}

// reference of state q in the FP32 offset storage (StateRef in models.hpp; 0 = plain),
// rounded to FP32 as the kernels use it
double state_ref(int model, int q)
{
    double r = 0.0;
    switch (model) {
    case cg::MODEL_TNNP: r = cg::StateRef<cg::TNNP>::at(q); break;
    case cg::MODEL_ORD: r = cg::StateRef<cg::ORd>::at(q); break;
    case cg::MODEL_CRN: r = cg::StateRef<cg::Courtemanche>::at(q); break;
    case cg::MODEL_PYS: r = cg::StateRef<cg::PerisYague>::at(q); break;
    case cg::MODEL_STW: r = cg::StateRef<cg::Stewart>::at(q); break;
    case cg::MODEL_GRD: r = cg::StateRef<cg::Grandi>::at(q); break;
    case cg::MODEL_TOR: r = cg::StateRef<cg::ToRORd>::at(q); break;
    case cg::MODEL_GAUR: r = cg::StateRef<cg::Gaur>::at(q); break;
    default: break;
    }
    return (double)(float)r;
}

// model time of a PDE time in seconds, as each wrapper's solveODE scales it
// (BuenoOrovio.C:107-108, AlievPanfilov.C:107-108)
double model_time(int model, double t_s)
{
    if (model == cg::MODEL_FDA) return t_s;  // monodomainFDAManufactured.C:107
    return model == cg::MODEL_AP ? t_s * 1000.0 / 12.9 : t_s * 1000.0;
}

// deltaT in model time units (BuenoOrovio.C:127, AlievPanfilov.C:127, ...)
double model_dt(int model, double dt_s)
{
    if (model == cg::MODEL_FDA) return dt_s;
    return model == cg::MODEL_AP ? dt_s * 1000.0 / 12.9 : dt_s * 1000.0;
}

// tStart and tEnd of one step as each solveODE forms them: scaled (t0 + dt) for the
// physiological models, tStart + deltaT for the manufactured one
void model_step_times(int model, double t0_s, double dt_s, double* t0, double* t1)
{
    *t0 = model_time(model, t0_s);
    *t1 = model == cg::MODEL_FDA ? *t0 + dt_s : model_time(model, t0_s + dt_s);
}

} // namespace

struct cg_ionic {
    int model, tissue, prec, solve_vm;
    ModelInfo mi;
    int64_t n;
    cg::ModelParams P;
    size_t elem;
    void* d_states = nullptr;
    void* d_gscale = nullptr;
    std::vector<double> h_gscale;     // host copy of the scales [4][n] (derived-variable calls)
    double* d_vm = nullptr;
    double* d_im = nullptr;
    hipStream_t stream = nullptr;
    cg::Stimulus stim{};
    // batched protocols (single-cell handles)
    cg::Stimulus* d_prot = nullptr;   // [n] per-cell protocol, or null
    int64_t* d_len = nullptr;         // [n] run length in steps
    int64_t* d_order = nullptr;       // lane -> cell, longest runs first
    int64_t max_len = 0;
    bool lens_set = false;
    int variant = 0;                  // kernel shape chosen by cg_batch_autotune
    // beat detection
    int max_beats = 0;
    double beat_thr = 0.0, beat_base_s = 0.0, beat_peak_s = 0.0;  // V/s, s, s
    double* d_det = nullptr;
    void* d_ring = nullptr;
    int ring_len = 0;
    double* d_beats = nullptr;
    int* d_nbeats = nullptr;
    double beats_dt = 0.0;            // dt of the run that filled the beat records
    // ODE integrator (ionicModel::odeSolver(), ionicModel.H:233-248)
    int ode_kind = CG_ODE_RUSH_LARSEN;
    double abs_tol = 1e-15, rel_tol = 1e-4;
    long long max_steps = 10000;
    double init_step = 0.0;           // model time units; <= 0: the first deltaT in seconds
    bool step_set = false;            // d_step holds the kept per-cell trial steps
    double* d_step = nullptr;         // [n]
    int* d_fail = nullptr;            // set by a cell whose integration failed
    // single-cell trace buffer, [steps][states] in double
    double* d_trace = nullptr;
    size_t trace_cap = 0;
};

struct cg_mesh {
    cg::HexGeom g;
    int64_t n;
    bool uniform = true;
    bool has_corr = false;
    double sig[9] = {0};
    double* d_od = nullptr;  // off-diagonal components (xy, xz, yz) [3][n]
    unsigned od_mask = 0;
    double* d_fx = nullptr;
    double* d_fy = nullptr;
    double* d_fz = nullptr;
    std::vector<double> h_sigma9;  // host copy for dt and implicit assembly
    cg::GeneralMesh* gm = nullptr; // set for a general (polyhedral) mesh; g is unused then
    double* d_sigma9 = nullptr;    // per-cell tensors on the device, made on demand:
    bool sigma_sym = false;        // [6][n] symmetric (xx xy xz yy yz zz), else [9][n]
    std::vector<double> h_centres; // hex block: the caller's cell centres [3n], or empty
    uint64_t centres_ver = 0;      // bumped by every cg_mesh_set_cell_centres
};

struct cg_tissue {
    cg_mesh* mesh;
    cg_ionic* ion;
    cg_tissue_config cfg;
    double* d_V[3] = {nullptr, nullptr, nullptr};
    int cur = 0;  // index of V^n; (cur+2)%3 holds V^{n-1}
    double* d_Iion = nullptr;
    uint8_t* d_flag = nullptr;
    double* d_tact = nullptr;
    int n_boxes = 0;
    double box_lo[cg::MAX_BOXES][3], box_hi[cg::MAX_BOXES][3];
    double box_t0[cg::MAX_BOXES], box_dur[cg::MAX_BOXES], box_I[cg::MAX_BOXES];
    // box membership of each cell from the mesh's own centres (hex block with centres):
    // bit b of d_box_mask[c] set when box b contains the centre; rebuilt when the boxes
    // or the centres change
    uint8_t* d_box_mask = nullptr;
    uint64_t boxes_ver = 1, mask_boxes_ver = 0, mask_centres_ver = 0;
    double t = 0.0;
    int64_t steps = 0;
    double act_v = 0.0;  // activation threshold (V)
    cg::ImplicitSolver* implicit = nullptr;
    cg::Profiler prof;
    // pseudo-ECG buffers: electrode positions, per-block partial sums, results
    double* d_ecg = nullptr;   // [ECG_MAX_ELECTRODES * (3 + ECG_BLOCKS + 1)]
    // general mesh, implicit: the lduMatrix solver, right-hand sides, controls, report
    cg::LduSolver* gldu = nullptr;
    cg::GeneralMesh* own_gm = nullptr;   // a hex tissue's block as a general mesh (diffusion entry)
    double* d_src = nullptr;             // caller's source field (diffusion entry)
    double* d_b0 = nullptr;
    double* d_b = nullptr;
    double g_c = 0.0;                 // ddt coefficient the solver's matrix holds (0: none)
    double g_tol = 0.0, g_rel = 0.0, g_rc = 0.0;
    int g_max_iter = 1000, g_outer = 1;
    int g_solves = 0;
    double g_r0 = 0.0, g_r1 = 0.0;
    int64_t g_iters = 0;
};

namespace {

// Integrator arguments of one ionic step from t0_s over dt_s seconds: tStart, tEnd and the
// clamp deltaT in model time as each solveODE scales them (BuenoOrovio.C:107-108,127,
// AlievPanfilov.C:107-108,127).  The first RKF45 step fills the per-cell trial steps with
// the handle's initial step, which by default is deltaT in seconds read as model time, as
// ionicModel's step_(num, initialDeltaT) is (ionicModel.C:38).
int ode_args(cg_ionic* h, double t0_s, double dt_s, cg::OdeArgs* o)
{
    *o = cg::OdeArgs{};
    o->method = h->ode_kind;   // CG_ODE_RUSH_LARSEN 0, CG_ODE_RKF45 1, CG_ODE_EULER 2
    if (!o->method) return CG_OK;
    const bool ap = h->model == cg::MODEL_AP;
    model_step_times(h->model, t0_s, dt_s, &o->t0, &o->t1);
    o->dt_clamp = model_dt(h->model, dt_s);
    o->abs_tol = h->abs_tol;
    o->rel_tol = h->rel_tol;
    o->max_steps = h->max_steps;
    o->t_s = t0_s;
    o->dt_s = dt_s;
    o->ap = ap;
    if (!h->step_set) {
        const double s0 = h->init_step > 0.0 ? h->init_step : dt_s;
        std::vector<double> v((size_t)h->n, s0);
        HIPCHK(hipMemcpyAsync(h->d_step, v.data(), sizeof(double) * h->n,
                              hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        h->step_set = true;
    }
    o->step = h->d_step;
    o->fail = h->d_fail;
    return CG_OK;
}

// ODESolver's fatal errors (step-size underflow, maxSteps) as an error code, read at the
// handle's synchronisation points
int check_ode(cg_ionic* h)
{
    if (h->ode_kind == CG_ODE_RUSH_LARSEN || !h->d_fail) return CG_OK;
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, h->d_fail, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!f) return CG_OK;
    HIPCHK(hipMemsetAsync(h->d_fail, 0, sizeof(int), h->stream));
    return fail(CG_ERR_ODE, "%s integration failed in at least one cell (step-size "
                            "underflow or more than maxSteps = %lld steps)",
                h->ode_kind == CG_ODE_EULER ? "Euler" : "RKF45", h->max_steps);
}

} // namespace

extern "C" {

const char* cg_last_error(void) { return g_err.c_str(); }
const char* cg_version(void) { return "cardiacgpu 0.1 gfx950"; }

int cg_model_id(const char* key)
{
    if (!key) return -1;
    for (int i = 0; i < kNumModels; ++i)
        if (!strcmp(key, kModelKeys[i])) return i;
    return -1;
}

int cg_tissue_flag(const char* t)
{
    if (!t) return -1;
    if (!strcmp(t, "epicardialCells")) return 1;
    if (!strcmp(t, "mCells")) return 2;
    if (!strcmp(t, "endocardialCells")) return 3;
    if (!strcmp(t, "myocyte")) return 4;
    // the manufactured model's `dimension` word
    if (!strcmp(t, "1D")) return 11;
    if (!strcmp(t, "2D")) return 12;
    if (!strcmp(t, "3D")) return 13;
    return -1;
}

int cg_model_nstates(int model)
{
    ModelInfo mi;
    return model_info(model, &mi) ? mi.ns : -1;
}

int cg_model_units(int model, double* vmul, double* voff, double* vscale, double* t_per_s)
{
    ModelInfo mi;
    if (!model_info(model, &mi)) return fail(CG_ERR_MODEL, "unknown model id %d", model);
    if (vmul) *vmul = model == cg::MODEL_FDA ? cg::FdaManufactured::VMUL : 1000.0;
    if (voff) *voff = mi.voff;
    if (vscale) *vscale = mi.vscale;
    if (t_per_s) *t_per_s = mi.tscale;
    return CG_OK;
}

int cg_model_state_name(int model, int i, char* buf, size_t len)
{
    ModelInfo mi;
    if (!model_info(model, &mi)) return fail(CG_ERR_MODEL, "unknown model id %d", model);
    if (i < 0 || i >= mi.ns || !buf || len == 0) return fail(CG_ERR_ARG, "bad state index %d", i);
    const char* nm = model == 0 ? kBONames[i] : model == 1 ? kTNNPNames[i]
                   : model == 2 ? kORdNames[i] : model == 3 ? kAPNames[i]
                   : model == 4 ? kFDANames[i] : model == 5 ? kCRNNames[i]
                   : model == 6 ? kPYSNames[i] : model == 7 ? kSTWNames[i]
                   : model == 8 ? kGRDNames[i] : model == 9 ? kTORNames[i]
                   : kGAURNames[i];
    snprintf(buf, len, "%s", nm);
    return CG_OK;
}

// ALGEBRAIC names in the reference's order (BuenoOrovio_2008.H:30-43,
// AlievPanfilov_1996.H:28-32) of the models whose algebraic list the device evaluates
static const char* const kBOAlgNames[] = {"Jfi", "Jso", "Jsi", "Jion", "Istim", "tauSo",
                                          "tauO", "tauVMinus", "tauWMinus", "tauS", "vInfty",
                                          "wInfty"};
static const char* const kAPAlgNames[] = {"AV_eps", "Istim", "Iion_cm"};

int cg_model_nalgebraic(int model)
{
    ModelInfo mi;
    if (!model_info(model, &mi)) return -1;
    return cg::model_nalgebraic(model);
}

int cg_model_algebraic_name(int model, int i, char* buf, size_t len)
{
    const int na = cg_model_nalgebraic(model);
    if (na < 0) return fail(CG_ERR_MODEL, "unknown model id %d", model);
    if (i < 0 || i >= na || !buf || len == 0) return fail(CG_ERR_ARG, "bad algebraic index %d", i);
    snprintf(buf, len, "%s", model == cg::MODEL_BO ? kBOAlgNames[i] : kAPAlgNames[i]);
    return CG_OK;
}

int cg_ionic_create(int model, int tissue, int64_t n, int solve_vm, int precision,
                    cg_ionic** out)
{
    if (!out) return fail(CG_ERR_ARG, "out is NULL");
    *out = nullptr;
    ModelInfo mi;
    if (!model_info(model, &mi)) return fail(CG_ERR_MODEL, "unknown ionicModel id %d", model);
    if (n <= 0) return fail(CG_ERR_ARG, "n_cells must be positive");
    if (precision != CG_FP64 && precision != CG_MIXED)
        return fail(CG_ERR_ARG, "unknown precision %d", precision);
    double p[cg::MAX_PARAMS] = {0}, y0[64] = {0};
    if (!model_host_init(model, tissue, p, y0))
        return fail(CG_ERR_TISSUE, "Unsupported tissue flag %d for ionicModel %s", tissue,
                    kModelKeys[model]);
    cg_ionic* h = new cg_ionic();
    h->model = model; h->tissue = tissue; h->prec = precision; h->solve_vm = solve_vm ? 1 : 0;
    h->mi = mi; h->n = n;
    memcpy(h->P.p, p, sizeof p);
    h->elem = precision == CG_FP64 ? sizeof(double) : sizeof(float);
    const size_t bytes = h->elem * (size_t)mi.ns * (size_t)n;
    if (hipMalloc(&h->d_states, bytes) != hipSuccess) {
        delete h;
        return fail(CG_ERR_NOMEM, "hipMalloc of %zu bytes for states failed", bytes);
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        hipFree(h->d_states);
        delete h;
        return fail(CG_ERR_HIP, "stream creation failed");
    }
    // initial states, one SoA plane per variable
    std::vector<char> plane(h->elem * (size_t)n);
    for (int k = 0; k < mi.ns; ++k) {
        if (precision == CG_FP64) {
            double* q = (double*)plane.data();
            for (int64_t i = 0; i < n; ++i) q[i] = y0[k];
        } else {
            // FP32 storage: offsets from the state's reference
            float* q = (float*)plane.data();
            const float w = (float)(y0[k] - state_ref(model, k));
            for (int64_t i = 0; i < n; ++i) q[i] = w;
        }
        hipMemcpy((char*)h->d_states + h->elem * (size_t)k * n, plane.data(), plane.size(),
                  hipMemcpyHostToDevice);
    }
    *out = h;
    return CG_OK;
}

int cg_ionic_destroy(cg_ionic* h)
{
    if (!h) return CG_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->d_states);
    if (h->d_gscale) hipFree(h->d_gscale);
    if (h->d_vm) hipFree(h->d_vm);
    if (h->d_im) hipFree(h->d_im);
    void* batch[] = {h->d_prot, h->d_len, h->d_order, h->d_det, h->d_ring, h->d_beats,
                     h->d_nbeats, h->d_step, h->d_fail, h->d_trace};
    for (void* q : batch)
        if (q) hipFree(q);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return CG_OK;
}

int cg_ionic_nstates(const cg_ionic* h) { return h ? h->mi.ns : -1; }

int cg_ionic_set_constant(cg_ionic* h, int k, double v)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (k < 0 || k >= h->mi.np) return fail(CG_ERR_ARG, "constant index %d out of range", k);
    h->P.p[k] = v;
    return CG_OK;
}

int cg_ionic_set_conductance_scale(cg_ionic* h, const double* s)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if ((h->model == cg::MODEL_BO || h->model == cg::MODEL_AP || h->model == cg::MODEL_FDA) && s)
        return fail(CG_ERR_STATE, "conductance scaling is not defined for ionicModel %s",
                    kModelKeys[h->model]);
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!s) {
        if (h->d_gscale) hipFree(h->d_gscale);
        h->d_gscale = nullptr;
        h->h_gscale.clear();
        return CG_OK;
    }
    const size_t cnt = 4 * (size_t)h->n;
    h->h_gscale.assign(s, s + cnt);
    if (!h->d_gscale && hipMalloc(&h->d_gscale, h->elem * cnt) != hipSuccess)
        return fail(CG_ERR_NOMEM, "hipMalloc for conductance scales failed");
    if (h->prec == CG_FP64) {
        HIPCHK(hipMemcpy(h->d_gscale, s, sizeof(double) * cnt, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(cnt);
        for (size_t i = 0; i < cnt; ++i) f[i] = (float)s[i];
        HIPCHK(hipMemcpy(h->d_gscale, f.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
    }
    return CG_OK;
}

int cg_ionic_set_stimulus(cg_ionic* h, const cg_stimulus* s)
{
    if (!h || !s) return fail(CG_ERR_ARG, "NULL argument");
    if (!h->solve_vm)
        return fail(CG_ERR_STATE,
                    "Detected active ionic-model stimulus protocol while using a "
                    "reaction-diffusion myocardium domain; use externalStimulus");
    if (s->stim_duration < 0.0) return fail(CG_ERR_STIMULUS, "stim_duration must be non-negative");
    h->stim = {s->stim_start, s->stim_period_s1, s->stim_duration, s->stim_amplitude,
               s->nstim1, s->stim_period_s2, s->nstim2};
    return CG_OK;
}

int cg_ionic_set_integrator(cg_ionic* h, int kind, double abs_tol, double rel_tol,
                            int64_t max_steps, double initial_step)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (kind != CG_ODE_RUSH_LARSEN && kind != CG_ODE_RKF45 && kind != CG_ODE_EULER)
        return fail(CG_ERR_ARG, "unknown ODE solver %d", kind);
    const bool adaptive = kind != CG_ODE_RUSH_LARSEN;
    if (adaptive &&
        (!(abs_tol >= 0.0) || !(rel_tol >= 0.0) || !(abs_tol + rel_tol > 0.0) || max_steps < 1))
        return fail(CG_ERR_ARG, "%s needs absTol, relTol >= 0 (not both 0) and maxSteps >= 1",
                    kind == CG_ODE_EULER ? "Euler" : "RKF45");
    // Gaur's FP64 tissue-mode RKF45 kernel (29 states, every stage array live) returns
    // non-finite states on physiological inputs in the gfx950 build, while the same source
    // compiled for the host (also under address and undefined-behaviour sanitizers) and the
    // FP32 and single-cell FP64 kernels agree with the oracle; the result depends on how
    // that one kernel is compiled, so it stays switched off until the cause is known.
    if (kind == CG_ODE_RKF45 && h->model == cg::MODEL_GAUR && h->prec == CG_FP64 && !h->solve_vm)
        return fail(CG_ERR_STATE, "RKF45 in tissue mode is not available for Gaur on FP64 "
                                  "handles (use Rush-Larsen, Euler or precision mixed)");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (adaptive && !h->d_step) {
        if (hipMalloc(&h->d_step, sizeof(double) * h->n) != hipSuccess ||
            hipMalloc(&h->d_fail, sizeof(int)) != hipSuccess)
            return fail(CG_ERR_NOMEM, "allocation of the per-cell ODE steps failed");
        HIPCHK(hipMemset(h->d_fail, 0, sizeof(int)));
    }
    h->ode_kind = kind;
    h->abs_tol = abs_tol;
    h->rel_tol = rel_tol;
    h->max_steps = max_steps;
    h->init_step = initial_step;
    h->step_set = false;
    return CG_OK;
}

int cg_ionic_step(cg_ionic* h, double t0, double dt, const double* Vm, double* Im, int flags)
{
    if (!h || !Vm || !Im) return fail(CG_ERR_ARG, "NULL argument");
    if (h->solve_vm) return fail(CG_ERR_STATE, "cg_ionic_step needs a tissue handle");
    if (!(dt > 0.0)) return fail(CG_ERR_ARG, "deltaT must be positive");
    cg::OdeArgs ode;
    if (int rc = ode_args(h, t0, dt, &ode)) return rc;
    const double* dv = Vm;
    double* di = Im;
    if (!(flags & CG_DEVICE_PTRS)) {
        if (!h->d_vm) {
            if (hipMalloc(&h->d_vm, sizeof(double) * h->n) != hipSuccess ||
                hipMalloc(&h->d_im, sizeof(double) * h->n) != hipSuccess)
                return fail(CG_ERR_NOMEM, "staging allocation failed");
        }
        HIPCHK(hipMemcpyAsync(h->d_vm, Vm, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
        dv = h->d_vm;
        di = h->d_im;
    }
    const double dtm = model_dt(h->model, dt);
    HIPCHK(cg::launch_ionic_tissue(h->model, h->prec, h->tissue, h->d_states, h->n, dv, di, dtm,
                                   h->P, h->d_gscale, ode, h->stream));
    if (!(flags & CG_DEVICE_PTRS))
        HIPCHK(hipMemcpyAsync(Im, di, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_ode(h);
}

int cg_ionic_run_cells(cg_ionic* h, double t0, double dt, int n_steps)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (!h->solve_vm) return fail(CG_ERR_STATE, "cg_ionic_run_cells needs a single-cell handle");
    if (n_steps < 0 || !(dt > 0.0)) return fail(CG_ERR_ARG, "bad step arguments");
    const double dtm = model_dt(h->model, dt);
    cg::OdeArgs ode;
    if (int rc = ode_args(h, t0, dt, &ode)) return rc;
    HIPCHK(cg::launch_ionic_cells(h->model, h->prec, h->tissue, h->d_states, h->n,
                                  model_time(h->model, t0), dtm, n_steps, h->P, h->d_gscale,
                                  h->stim, ode, h->stream));
    return CG_OK;
}

int cg_ionic_run_cells_trace(cg_ionic* h, double t0, double dt, int n_steps, int64_t cell,
                             double* trace)
{
    if (!h || !trace) return fail(CG_ERR_ARG, "NULL argument");
    if (!h->solve_vm) return fail(CG_ERR_STATE, "cg_ionic_run_cells needs a single-cell handle");
    if (n_steps < 0 || !(dt > 0.0)) return fail(CG_ERR_ARG, "bad step arguments");
    if (cell < 0 || cell >= h->n) return fail(CG_ERR_ARG, "traced cell %lld out of range",
                                              (long long)cell);
    if (n_steps == 0) return CG_OK;
    const size_t bytes = sizeof(double) * (size_t)n_steps * 2 * (size_t)h->mi.ns;
    if (bytes > h->trace_cap) {
        if (h->d_trace) hipFree(h->d_trace);
        h->d_trace = nullptr;
        h->trace_cap = 0;
        if (hipMalloc(&h->d_trace, bytes) != hipSuccess)
            return fail(CG_ERR_NOMEM, "trace allocation of %zu bytes failed", bytes);
        h->trace_cap = bytes;
    }
    const double dtm = model_dt(h->model, dt);
    cg::OdeArgs ode;
    if (int rc = ode_args(h, t0, dt, &ode)) return rc;
    HIPCHK(cg::launch_ionic_cells(h->model, h->prec, h->tissue, h->d_states, h->n,
                                  model_time(h->model, t0), dtm, n_steps, h->P, h->d_gscale,
                                  h->stim, ode, h->stream, h->d_trace, cell));
    HIPCHK(hipMemcpyAsync(trace, h->d_trace, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_ode(h);
}

// ----------------------------------------------------------- batched protocols
int cg_ionic_set_protocols(cg_ionic* h, const cg_stimulus* per_cell)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (!h->solve_vm) return fail(CG_ERR_STATE, "per-cell protocols need a single-cell handle");
    if (!per_cell) {
        if (h->d_prot) hipFree(h->d_prot);
        h->d_prot = nullptr;
        return CG_OK;
    }
    std::vector<cg::Stimulus> v((size_t)h->n);
    for (int64_t i = 0; i < h->n; ++i) {
        const cg_stimulus& q = per_cell[i];
        if (q.stim_duration < 0.0)
            return fail(CG_ERR_STIMULUS, "cell %lld: stim_duration must be non-negative",
                        (long long)i);
        v[i] = {q.stim_start, q.stim_period_s1, q.stim_duration, q.stim_amplitude, q.nstim1,
                q.stim_period_s2, q.nstim2};
    }
    if (!h->d_prot && hipMalloc(&h->d_prot, sizeof(cg::Stimulus) * h->n) != hipSuccess)
        return fail(CG_ERR_NOMEM, "protocol allocation failed");
    HIPCHK(hipMemcpy(h->d_prot, v.data(), sizeof(cg::Stimulus) * h->n, hipMemcpyHostToDevice));
    return CG_OK;
}

int cg_ionic_set_run_lengths(cg_ionic* h, const int64_t* steps, int64_t n_steps_all)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (!h->solve_vm) return fail(CG_ERR_STATE, "run lengths need a single-cell handle");
    std::vector<int64_t> len((size_t)h->n);
    for (int64_t i = 0; i < h->n; ++i) {
        len[i] = steps ? steps[i] : n_steps_all;
        if (len[i] < 0) return fail(CG_ERR_ARG, "negative run length for cell %lld", (long long)i);
    }
    // lanes in order of decreasing run length: the lanes of a wave finish together
    std::vector<int64_t> order((size_t)h->n);
    for (int64_t i = 0; i < h->n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return len[a] > len[b]; });
    if (!h->d_len && (hipMalloc(&h->d_len, sizeof(int64_t) * h->n) != hipSuccess ||
                      hipMalloc(&h->d_order, sizeof(int64_t) * h->n) != hipSuccess))
        return fail(CG_ERR_NOMEM, "run-length allocation failed");
    HIPCHK(hipMemcpy(h->d_len, len.data(), sizeof(int64_t) * h->n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_order, order.data(), sizeof(int64_t) * h->n, hipMemcpyHostToDevice));
    h->max_len = len.empty() ? 0 : len[order[0]];
    h->lens_set = true;
    return CG_OK;
}

int cg_ionic_enable_beats(cg_ionic* h, int max_beats, double dvdt_threshold, double baseline_s,
                          double peak_window_s)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (!h->solve_vm) return fail(CG_ERR_STATE, "beat detection needs a single-cell handle");
    if (max_beats < 0 || !(baseline_s > 0) || !(peak_window_s > 0))
        return fail(CG_ERR_ARG, "bad beat-detection arguments");
    auto drop = [&]() {
        void* q[] = {h->d_det, h->d_ring, h->d_beats, h->d_nbeats};
        for (void* x : q)
            if (x) hipFree(x);
        h->d_det = nullptr; h->d_ring = nullptr; h->d_beats = nullptr; h->d_nbeats = nullptr;
        h->ring_len = 0;
    };
    drop();
    h->max_beats = max_beats;
    if (max_beats == 0) return CG_OK;
    h->beat_thr = dvdt_threshold;
    h->beat_base_s = baseline_s;
    h->beat_peak_s = peak_window_s;
    const size_t n = (size_t)h->n;
    if (hipMalloc(&h->d_det, sizeof(double) * cg::DS_N * n) != hipSuccess ||
        hipMalloc(&h->d_beats, sizeof(double) * cg::BEAT_REC * max_beats * n) != hipSuccess ||
        hipMalloc(&h->d_nbeats, sizeof(int) * n) != hipSuccess) {
        drop();
        return fail(CG_ERR_NOMEM, "beat-detection allocation failed");
    }
    return CG_OK;
}

namespace {

// one chunk launch of handle h covering steps [s0, s0 + chunk)
hipError_t batch_launch(cg_ionic* h, double dtm, int64_t s0, int chunk, hipStream_t s)
{
    cg::BatchArgs a{};
    a.n = h->n;
    a.s0 = s0;
    a.chunk = chunk;
    a.dt = dtm;
    a.order = h->d_order;
    a.len = h->d_len;
    a.prot = h->d_prot;
    a.stim = h->stim;
    if (h->max_beats > 0) {
        const double per_s = model_time(h->model, 1.0);  // model time units per second
        a.det = h->d_det;
        a.ring = h->d_ring;
        a.ring_len = h->ring_len;
        a.beats = h->d_beats;
        a.nbeats = h->d_nbeats;
        a.max_beats = h->max_beats;
        a.thr = h->beat_thr * 1000.0 / per_s;  // V/s -> mV per model time unit
        a.baseline = h->beat_base_s * per_s;
        a.peak_window = h->beat_peak_s * per_s;
        a.vscale = h->mi.vscale;
        a.voff = h->mi.voff;
    }
    if (h->ode_kind != CG_ODE_RUSH_LARSEN) {
        a.method = h->ode_kind;
        a.abs_tol = h->abs_tol;
        a.rel_tol = h->rel_tol;
        a.max_steps = h->max_steps;
        a.step = h->d_step;
        a.fail = h->d_fail;
    }
    return h->prec == CG_FP64
        ? cg::launch_cells_protocol_fp64(h->model, h->tissue, h->variant, h->d_states, h->P,
                                         h->d_gscale, a, s)
        : cg::launch_cells_protocol_fp32(h->model, h->tissue, h->variant, h->d_states, h->P,
                                         h->d_gscale, a, s);
}

int batch_check(cg_ionic* const* hs, int nh, double dt_s)
{
    if (!hs || nh <= 0) return fail(CG_ERR_ARG, "no handles");
    if (!(dt_s > 0.0)) return fail(CG_ERR_ARG, "dt must be positive");
    for (int i = 0; i < nh; ++i) {
        if (!hs[i]) return fail(CG_ERR_ARG, "NULL handle %d", i);
        if (!hs[i]->solve_vm) return fail(CG_ERR_STATE, "handle %d is a tissue handle", i);
        if (!hs[i]->lens_set) return fail(CG_ERR_STATE, "handle %d: run lengths not set", i);
        if (hs[i]->model == cg::MODEL_FDA)
            return fail(CG_ERR_STATE, "handle %d: the manufactured model has no pacing protocol", i);
        // Gaur's FP64 RKF45 batch kernel failed (maxSteps) on cells whose oracle run
        // takes ordinary steps, as its tissue-mode twin returns non-finite states (see
        // cg_ionic_set_integrator); it stays refused
        if (hs[i]->ode_kind == CG_ODE_RKF45 && hs[i]->model == cg::MODEL_GAUR &&
            hs[i]->prec == CG_FP64)
            return fail(CG_ERR_STATE, "handle %d: RKF45 batches are not available for Gaur on "
                                      "FP64 handles", i);
        for (int j = 0; j < i; ++j)
            if (hs[j] == hs[i]) return fail(CG_ERR_ARG, "handle %d given twice", i);
    }
    return CG_OK;
}

} // namespace

int cg_batch_run(cg_ionic* const* hs, int nh, double dt_s, int chunk)
{
    int rc = batch_check(hs, nh, dt_s);
    if (rc) return rc;
    if (chunk <= 0) return fail(CG_ERR_ARG, "chunk must be positive");
    // beat detection state and the adaptive solvers' first trial steps, set outside the
    // graph
    for (int i = 0; i < nh; ++i) {
        cg_ionic* h = hs[i];
        HIPCHK(hipStreamSynchronize(h->stream));
        cg::OdeArgs o;
        if (int rc2 = ode_args(h, 0.0, dt_s, &o)) return rc2;
        if (h->max_beats <= 0) continue;
        const double dtm = model_dt(h->model, dt_s);
        const int need = (int)std::ceil(h->beat_base_s * model_time(h->model, 1.0) / dtm) + 4;
        if (need != h->ring_len) {
            if (h->d_ring) hipFree(h->d_ring);
            h->d_ring = nullptr;
            if (hipMalloc(&h->d_ring, h->elem * (size_t)need * h->n) != hipSuccess)
                return fail(CG_ERR_NOMEM, "beat ring allocation failed");
            h->ring_len = need;
        }
        HIPCHK(hipMemset(h->d_det, 0, sizeof(double) * cg::DS_N * h->n));
        HIPCHK(hipMemset(h->d_nbeats, 0, sizeof(int) * h->n));
        h->beats_dt = dt_s;
    }
    // capture: fork every handle's stream off the first, chunk launches per handle, join
    hipStream_t s0 = hs[0]->stream;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;
    auto cleanup = [&]() {
        for (hipEvent_t e : ev) hipEventDestroy(e);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    };
    HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    hipEvent_t fork;
    if ((err = hipEventCreateWithFlags(&fork, hipEventDisableTiming)) == hipSuccess) {
        ev.push_back(fork);
        err = hipEventRecord(fork, s0);
    }
    for (int i = 1; i < nh && err == hipSuccess; ++i) err = hipStreamWaitEvent(hs[i]->stream, fork, 0);
    for (int i = 0; i < nh && err == hipSuccess; ++i) {
        cg_ionic* h = hs[i];
        const double dtm = model_dt(h->model, dt_s);
        for (int64_t st = 0; st < h->max_len && err == hipSuccess; st += chunk)
            err = batch_launch(h, dtm, st, chunk, h->stream);
    }
    for (int i = 1; i < nh && err == hipSuccess; ++i) {
        hipEvent_t join;
        if ((err = hipEventCreateWithFlags(&join, hipEventDisableTiming)) != hipSuccess) break;
        ev.push_back(join);
        if ((err = hipEventRecord(join, hs[i]->stream)) == hipSuccess)
            err = hipStreamWaitEvent(s0, join, 0);
    }
    const hipError_t ec = hipStreamEndCapture(s0, &graph);
    if (err == hipSuccess) err = ec;
    if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (err == hipSuccess) err = hipGraphLaunch(exec, s0);
    if (err == hipSuccess) err = hipStreamSynchronize(s0);
    cleanup();
    if (err != hipSuccess) return fail(CG_ERR_HIP, "batched run: %s", hipGetErrorString(err));
    for (int i = 0; i < nh; ++i)
        if (int rc2 = check_ode(hs[i])) return rc2;
    return CG_OK;
}

int cg_batch_autotune(cg_ionic* const* hs, int nh, double dt_s, int trial_steps, char* report,
                      size_t len)
{
    int rc = batch_check(hs, nh, dt_s);
    if (rc) return rc;
    if (trial_steps <= 0) return fail(CG_ERR_ARG, "trial_steps must be positive");
    for (int i = 0; i < nh; ++i)
        if (hs[i]->ode_kind != CG_ODE_RUSH_LARSEN)
            return fail(CG_ERR_STATE, "handle %d: the adaptive solvers run one kernel shape; "
                                      "autotuning is for Rush-Larsen handles", i);
    std::string rep;
    for (int i = 0; i < nh; ++i) {
        cg_ionic* h = hs[i];
        const double dtm = model_dt(h->model, dt_s);
        const size_t bytes = h->elem * (size_t)h->mi.ns * h->n;
        void* save = nullptr;
        if (hipMalloc(&save, bytes) != hipSuccess) return fail(CG_ERR_NOMEM, "autotune scratch");
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipError_t err = hipMemcpyAsync(save, h->d_states, bytes, hipMemcpyDeviceToDevice, h->stream);
        const int keep_beats = h->max_beats;
        h->max_beats = 0;  // trials without beat detection
        float best = 0.f;
        int best_v = 0;
        char line[160];
        for (int v = 0; v < cg::BATCH_VARIANTS && err == hipSuccess; ++v) {
            h->variant = v;
            err = batch_launch(h, dtm, 0, trial_steps, h->stream);  // warm-up
            if (err == hipSuccess) err = hipEventRecord(e0, h->stream);
            if (err == hipSuccess) err = batch_launch(h, dtm, 0, trial_steps, h->stream);
            if (err == hipSuccess) err = hipEventRecord(e1, h->stream);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (err != hipSuccess) break;
            snprintf(line, sizeof line, "%s/%s/%s: %s %.3f ms\n", kModelKeys[h->model],
                     h->prec == CG_FP64 ? "fp64" : "mixed", "variant", cg::batch_variant_name(v), ms);
            rep += line;
            if (v == 0 || ms < best) { best = ms; best_v = v; }
        }
        h->max_beats = keep_beats;
        if (err == hipSuccess)
            err = hipMemcpyAsync(h->d_states, save, bytes, hipMemcpyDeviceToDevice, h->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(h->stream);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipFree(save);
        if (err != hipSuccess) return fail(CG_ERR_HIP, "autotune: %s", hipGetErrorString(err));
        h->variant = best_v;
        snprintf(line, sizeof line, "%s/%s chosen: %s\n", kModelKeys[h->model],
                 h->prec == CG_FP64 ? "fp64" : "mixed", cg::batch_variant_name(best_v));
        rep += line;
    }
    if (report && len) snprintf(report, len, "%s", rep.c_str());
    return CG_OK;
}

int cg_ionic_get_beats(cg_ionic* h, int32_t* counts, double* records)
{
    if (!h) return fail(CG_ERR_ARG, "NULL handle");
    if (h->max_beats <= 0 || !h->d_nbeats) return fail(CG_ERR_STATE, "beat detection is off");
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->n;
    std::vector<int> cnt(n);
    HIPCHK(hipMemcpy(cnt.data(), h->d_nbeats, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (counts)
        for (size_t i = 0; i < n; ++i) counts[i] = cnt[i];
    if (records) {
        const size_t nr = n * h->max_beats * cg::BEAT_REC;
        HIPCHK(hipMemcpy(records, h->d_beats, sizeof(double) * nr, hipMemcpyDeviceToHost));
        // model units -> s and V; entries past each cell's count are NaN
        const double per_s = model_time(h->model, 1.0);
        for (size_t i = 0; i < n; ++i)
            for (int b = 0; b < h->max_beats; ++b) {
                double* r = records + (i * h->max_beats + b) * cg::BEAT_REC;
                if (b >= cnt[i]) {
                    for (int k = 0; k < cg::BEAT_REC; ++k) r[k] = NAN;
                    continue;
                }
                r[0] /= per_s;
                r[1] *= 1e-3;
                r[2] *= 1e-3;
                r[3] *= 1e-3;
                r[4] /= per_s;
            }
    }
    return CG_OK;
}

int cg_ionic_eval(cg_ionic* h, int kind, int64_t n, const double* states,
                  const double* t_model, double* out)
{
    if (!h || !states || !out || (h->solve_vm && !t_model)) return fail(CG_ERR_ARG, "NULL argument");
    if (n < 1) return fail(CG_ERR_ARG, "no rows");
    if (kind != CG_RATES && kind != CG_ALGEBRAIC) return fail(CG_ERR_ARG, "unknown kind %d", kind);
    const int nv = kind == CG_RATES ? h->mi.ns : cg::model_nalgebraic(h->model);
    if (nv <= 0)
        return fail(CG_ERR_MODEL, "the algebraic variables of ionicModel %s are not evaluated "
                    "on the device", kModelKeys[h->model]);
    const int ns = h->mi.ns;
    const bool scales = n == h->n && !h->h_gscale.empty();
    double *d_st = nullptr, *d_t = nullptr, *d_g = nullptr, *d_out = nullptr;
    bool ok = hipMalloc(&d_st, sizeof(double) * ns * n) == hipSuccess &&
              hipMalloc(&d_t, sizeof(double) * n) == hipSuccess &&
              hipMalloc(&d_out, sizeof(double) * nv * n) == hipSuccess &&
              (!scales || hipMalloc(&d_g, sizeof(double) * 4 * n) == hipSuccess);
    hipError_t e = ok ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = hipMemcpy(d_st, states, sizeof(double) * ns * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (t_model) e = hipMemcpy(d_t, t_model, sizeof(double) * n, hipMemcpyHostToDevice);
        else e = hipMemset(d_t, 0, sizeof(double) * n);
    }
    if (e == hipSuccess && scales)
        e = hipMemcpy(d_g, h->h_gscale.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = cg::launch_ionic_eval(h->model, h->tissue, n, d_st, d_t, kind, h->solve_vm, h->stim,
                                  h->P, d_g, d_out, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d_out, sizeof(double) * nv * n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d_st); hipFree(d_t); hipFree(d_g); hipFree(d_out);
    if (e == hipErrorOutOfMemory) return fail(CG_ERR_NOMEM, "eval buffers");
    if (e != hipSuccess) return fail(CG_ERR_HIP, "eval: %s", hipGetErrorString(e));
    return CG_OK;
}

int cg_ionic_get_restitution(cg_ionic* h, double* apd_di, uint8_t* flags)
{
    if (!h || !apd_di) return fail(CG_ERR_ARG, "NULL argument");
    if (h->max_beats <= 0 || !h->d_nbeats) return fail(CG_ERR_STATE, "beat detection is off");
    const size_t n = (size_t)h->n;
    double* d_out = nullptr;
    uint8_t* d_flags = nullptr;
    if (hipMalloc(&d_out, sizeof(double) * 2 * n) != hipSuccess ||
        (flags && hipMalloc(&d_flags, n * h->max_beats) != hipSuccess)) {
        hipFree(d_out);
        return fail(CG_ERR_NOMEM, "restitution output allocation failed");
    }
    const double inv_per_s = 1.0 / model_time(h->model, 1.0);
    hipError_t e = cg::launch_restitution(h->d_beats, h->d_nbeats, h->n, h->max_beats, inv_per_s,
                                          d_out, d_flags, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(apd_di, d_out, sizeof(double) * 2 * n, hipMemcpyDeviceToHost,
                           h->stream);
    if (e == hipSuccess && flags)
        e = hipMemcpyAsync(flags, d_flags, n * h->max_beats, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d_out);
    hipFree(d_flags);
    if (e != hipSuccess) return fail(CG_ERR_HIP, "restitution: %s", hipGetErrorString(e));
    return CG_OK;
}

static int states_io(cg_ionic* h, int first, int nv, double* host, const double* hin)
{
    if (!h || (!host && !hin)) return fail(CG_ERR_ARG, "NULL argument");
    if (first < 0 || nv < 0 || first + nv > h->mi.ns) return fail(CG_ERR_ARG, "state range");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (int rc = check_ode(h)) return rc;
    const size_t cnt = (size_t)nv * h->n;
    char* dsrc = (char*)h->d_states + h->elem * (size_t)first * h->n;
    if (h->prec == CG_FP64) {
        if (host) HIPCHK(hipMemcpy(host, dsrc, sizeof(double) * cnt, hipMemcpyDeviceToHost));
        else HIPCHK(hipMemcpy(dsrc, hin, sizeof(double) * cnt, hipMemcpyHostToDevice));
    } else {
        // FP32 storage holds offsets from each state's reference
        std::vector<float> f(cnt);
        if (host) {
            HIPCHK(hipMemcpy(f.data(), dsrc, sizeof(float) * cnt, hipMemcpyDeviceToHost));
            for (int v = 0; v < nv; ++v) {
                const double ref = state_ref(h->model, first + v);
                for (int64_t i = 0; i < h->n; ++i)
                    host[v * h->n + i] = ref + (double)f[v * h->n + i];
            }
        } else {
            for (int v = 0; v < nv; ++v) {
                const double ref = state_ref(h->model, first + v);
                for (int64_t i = 0; i < h->n; ++i)
                    f[v * h->n + i] = (float)(hin[v * h->n + i] - ref);
            }
            HIPCHK(hipMemcpy(dsrc, f.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
        }
    }
    return CG_OK;
}

int cg_ionic_get_states(cg_ionic* h, int first, int nv, double* host)
{
    return states_io(h, first, nv, host, nullptr);
}

int cg_ionic_set_states(cg_ionic* h, int first, int nv, const double* host)
{
    return states_io(h, first, nv, nullptr, host);
}

// ------------------------------------------------------------------------ mesh
int cg_mesh_create_hex(int nx, int ny, int nz, double hx, double hy, double hz, cg_mesh** out)
{
    if (!out) return fail(CG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1 || !(hx > 0) || !(hy > 0) || !(hz > 0))
        return fail(CG_ERR_ARG, "bad hex block dimensions");
    if ((int64_t)nx * ny * nz >= ((int64_t)1 << 31))
        return fail(CG_ERR_ARG, "hex block of %lld cells exceeds 2^31 - 1", (long long)nx * ny * nz);
    cg_mesh* m = new cg_mesh();
    m->g = {nx, ny, nz, hx, hy, hz};
    m->n = (int64_t)nx * ny * nz;
    *out = m;
    return CG_OK;
}

int cg_mesh_create_general(int64_t n_cells, int64_t n_faces, const int32_t* owner,
                           const int32_t* neighbour, const double* Sf, const double* Cf,
                           const double* C, const double* V, int64_t n_boundary_faces,
                           const int32_t* boundary_owner, const double* boundary_Sf,
                           cg_mesh** out)
{
    if (!out) return fail(CG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!C || !V || (n_faces > 0 && (!owner || !neighbour || !Sf || !Cf)) ||
        (n_boundary_faces > 0 && (!boundary_owner || !boundary_Sf)))
        return fail(CG_ERR_ARG, "NULL mesh array");
    cg::GeneralMesh* G = nullptr;
    const int rc = cg::general_create(&G, n_cells, n_faces, owner, neighbour, Sf, Cf, C, V,
                                      n_boundary_faces, boundary_owner, boundary_Sf, &g_err);
    if (rc) return rc == 5 ? CG_ERR_NOMEM : CG_ERR_ARG;
    cg_mesh* m = new cg_mesh();
    m->gm = G;
    m->n = n_cells;
    *out = m;
    return CG_OK;
}

int cg_mesh_match_ldu(int nx, int ny, int nz, int64_t nf, const int32_t* owner,
                      const int32_t* neighbour)
{
    if (!owner || !neighbour) return fail(CG_ERR_ARG, "NULL addressing");
    int64_t f = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int32_t c = i + nx * (j + ny * k);
                const int32_t nb[3] = {c + 1, c + nx, c + nx * ny};
                const bool has[3] = {i < nx - 1, j < ny - 1, k < nz - 1};
                for (int d = 0; d < 3; ++d) {
                    if (!has[d]) continue;
                    if (f >= nf || owner[f] != c || neighbour[f] != nb[d])
                        return fail(CG_ERR_ARG, "addressing differs from the hex block at face %lld",
                                    (long long)f);
                    ++f;
                }
            }
    if (f != nf) return fail(CG_ERR_ARG, "face count %lld != %lld", (long long)nf, (long long)f);
    return CG_OK;
}

int cg_mesh_destroy(cg_mesh* m)
{
    if (!m) return CG_OK;
    hipFree(m->d_od); hipFree(m->d_fx); hipFree(m->d_fy); hipFree(m->d_fz);
    hipFree(m->d_sigma9);
    cg::general_destroy(m->gm);
    delete m;
    return CG_OK;
}

int cg_mesh_set_cell_centres(cg_mesh* m, const double* C)
{
    if (!m || !C) return fail(CG_ERR_ARG, "NULL argument");
    if (m->gm) return fail(CG_ERR_STATE, "a general mesh carries its centres from cg_mesh_create_general");
    for (int64_t q = 0; q < 3 * m->n; ++q)
        if (!std::isfinite(C[q])) return fail(CG_ERR_ARG, "cell centre %lld is not finite", (long long)(q / 3));
    m->h_centres.assign(C, C + 3 * m->n);
    ++m->centres_ver;
    return CG_OK;
}

int64_t cg_mesh_ncells(const cg_mesh* m) { return m ? m->n : -1; }

int cg_mesh_set_conductivity(cg_mesh* m, const double* s, int uniform)
{
    if (!m || !s) return fail(CG_ERR_ARG, "NULL argument");
    const cg::HexGeom& g = m->g;
    const int64_t n = m->n;
    hipFree(m->d_od); hipFree(m->d_fx); hipFree(m->d_fy); hipFree(m->d_fz);
    hipFree(m->d_sigma9);
    m->d_od = m->d_fx = m->d_fy = m->d_fz = m->d_sigma9 = nullptr;
    m->h_sigma9.assign(9 * (size_t)n, 0.0);
    for (int64_t c = 0; c < n; ++c)
        for (int q = 0; q < 9; ++q) m->h_sigma9[9 * c + q] = uniform ? s[q] : s[9 * c + q];
    bool offdiag = false;
    for (int64_t c = 0; c < n && !offdiag; ++c)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (a != b && m->h_sigma9[9 * c + 3 * a + b] != 0.0) offdiag = true;
    m->has_corr = offdiag;
    m->uniform = uniform != 0;
    if (m->gm) {
        if (uniform) memcpy(m->sig, s, sizeof m->sig);
        if (cg::general_set_sigma(m->gm, m->h_sigma9.data(), &g_err)) return CG_ERR_HIP;
        return CG_OK;
    }
    if (uniform) {
        memcpy(m->sig, s, sizeof m->sig);
    } else {
        // per-face orthogonal coefficients: a_d = sigma_dd,f / h_d^2 (face area/(h V))
        std::vector<double> fx(n, 0.0), fy(n, 0.0), fz(n, 0.0);
        const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
        for (int64_t c = 0; c < n; ++c) {
            const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / sz);
            const double* a = &m->h_sigma9[9 * c];
            if (i < g.nx - 1) fx[c] = 0.5 * (a[0] + m->h_sigma9[9 * (c + 1) + 0]) / (g.hx * g.hx);
            if (j < g.ny - 1) fy[c] = 0.5 * (a[4] + m->h_sigma9[9 * (c + sy) + 4]) / (g.hy * g.hy);
            if (k < g.nz - 1) fz[c] = 0.5 * (a[8] + m->h_sigma9[9 * (c + sz) + 8]) / (g.hz * g.hz);
        }
        if (hipMalloc(&m->d_fx, sizeof(double) * n) != hipSuccess ||
            hipMalloc(&m->d_fy, sizeof(double) * n) != hipSuccess ||
            hipMalloc(&m->d_fz, sizeof(double) * n) != hipSuccess)
            return fail(CG_ERR_NOMEM, "face coefficient allocation failed");
        HIPCHK(hipMemcpy(m->d_fx, fx.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->d_fy, fy.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->d_fz, fz.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (offdiag) {
        std::vector<double> od(3 * (size_t)n);
        m->od_mask = 0;
        for (int64_t c = 0; c < n; ++c) {
            od[c] = m->h_sigma9[9 * c + 1];
            od[n + c] = m->h_sigma9[9 * c + 2];
            od[2 * n + c] = m->h_sigma9[9 * c + 5];
            for (int q = 0; q < 3; ++q)
                if (od[q * n + c] != 0.0) m->od_mask |= 1u << q;
        }
        if (hipMalloc(&m->d_od, sizeof(double) * 3 * n) != hipSuccess)
            return fail(CG_ERR_NOMEM, "conductivity allocation failed");
        HIPCHK(hipMemcpy(m->d_od, od.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    }
    return CG_OK;
}

int cg_mesh_explicit_dt(const cg_mesh* m, double chi, double cm, double maxco, double* dt)
{
    if (!m || !dt) return fail(CG_ERR_ARG, "NULL argument");
    if (m->h_sigma9.empty()) return fail(CG_ERR_STATE, "conductivity not set");
    if (m->gm) {
        if (cg::general_faces(m->gm) == 0) return fail(CG_ERR_STATE, "no internal faces");
        *dt = cg::general_explicit_dt(m->gm, chi * cm, maxco);
        return CG_OK;
    }
    const cg::HexGeom& g = m->g;
    const double chiCm = chi * cm;
    double best = INFINITY;
    const int64_t st[3] = {1, g.nx, (int64_t)g.nx * g.ny};
    const double h[3] = {g.hx, g.hy, g.hz};
    for (int64_t c = 0; c < m->n; ++c) {
        const int idx[3] = {(int)(c % g.nx), (int)((c / g.nx) % g.ny), (int)(c / st[2])};
        const int nn[3] = {g.nx, g.ny, g.nz};
        for (int d = 0; d < 3; ++d) {
            if (idx[d] >= nn[d] - 1) continue;
            const double snn = 0.5 * (m->h_sigma9[9 * c + 4 * d] + m->h_sigma9[9 * (c + st[d]) + 4 * d]);
            const double q = h[d] * h[d] / (snn / chiCm);
            if (q < best) best = q;
        }
    }
    if (!std::isfinite(best)) return fail(CG_ERR_STATE, "no internal faces");
    *dt = maxco * best;
    return CG_OK;
}

static cg::HexCoefs mesh_coefs(const cg_mesh* m)
{
    cg::HexCoefs k{};
    if (m->uniform) {
        k.ax = m->sig[0] / (m->g.hx * m->g.hx);
        k.ay = m->sig[4] / (m->g.hy * m->g.hy);
        k.az = m->sig[8] / (m->g.hz * m->g.hz);
    } else {
        k.fx = m->d_fx; k.fy = m->d_fy; k.fz = m->d_fz;
    }
    return k;
}

int cg_mesh_laplacian(cg_mesh* m, const double* V, double* out)
{
    if (!m || !V || !out) return fail(CG_ERR_ARG, "NULL argument");
    if (m->h_sigma9.empty()) return fail(CG_ERR_STATE, "conductivity not set");
    if (m->gm) {
        HIPCHK(cg::general_laplacian(m->gm, V, out, nullptr));
        HIPCHK(hipDeviceSynchronize());
        return CG_OK;
    }
    HIPCHK(cg::launch_hex_laplacian(m->g, mesh_coefs(m), V, m->d_od, m->od_mask, out, nullptr));
    HIPCHK(hipDeviceSynchronize());
    return CG_OK;
}

// ---------------------------------------------------------------------- tissue
int cg_tissue_create(cg_mesh* m, cg_ionic* ion, const cg_tissue_config* cfg, cg_tissue** out)
{
    if (!out || !m || !ion || !cfg) return fail(CG_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (ion->solve_vm) return fail(CG_ERR_STATE, "tissue needs an ionic handle with solve_vm = 0");
    if (ion->n != m->n) return fail(CG_ERR_ARG, "ionic cells %lld != mesh cells %lld",
                                    (long long)ion->n, (long long)m->n);
    if (m->h_sigma9.empty()) return fail(CG_ERR_STATE, "conductivity not set");
    if (!(cfg->dt > 0) || !(cfg->chi > 0) || !(cfg->cm > 0)) return fail(CG_ERR_ARG, "bad dt/chi/cm");
    if (cfg->algorithm != CG_EXPLICIT && cfg->algorithm != CG_IMPLICIT)
        return fail(CG_ERR_ARG, "unknown solutionAlgorithm %d", cfg->algorithm);
    if (cfg->ddt != CG_DDT_EULER && cfg->ddt != CG_DDT_BACKWARD)
        return fail(CG_ERR_ARG, "unknown ddt scheme %d", cfg->ddt);
    cg_tissue* t = new cg_tissue();
    t->mesh = m; t->ion = ion; t->cfg = *cfg;
    t->prof.stream = ion->stream;
    const int64_t n = m->n;
    bool ok = true;
    for (int q = 0; q < 3; ++q) ok &= hipMalloc(&t->d_V[q], sizeof(double) * n) == hipSuccess;
    ok &= hipMalloc(&t->d_Iion, sizeof(double) * n) == hipSuccess;
    ok &= hipMalloc(&t->d_flag, n) == hipSuccess;
    ok &= hipMalloc(&t->d_tact, sizeof(double) * n) == hipSuccess;
    if (!ok) {
        cg_tissue_destroy(t);
        return fail(CG_ERR_NOMEM, "tissue allocation failed");
    }
    std::vector<double> v(n, cfg->vm0);
    for (int q = 0; q < 3; ++q)
        HIPCHK(hipMemcpy(t->d_V[q], v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(t->d_Iion, 0, sizeof(double) * n));
    HIPCHK(hipMemset(t->d_flag, 1, n));
    HIPCHK(hipMemset(t->d_tact, 0, sizeof(double) * n));
    t->g_tol = cfg->tol;
    t->g_max_iter = cfg->max_iter;
    if (cfg->algorithm == CG_IMPLICIT && m->gm) {
        const int rc = cg::ldu_create(&t->gldu, n, cg::general_faces(m->gm),
                                      cg::general_owner(m->gm), cg::general_neighbour(m->gm),
                                      &g_err);
        if (rc || hipMalloc(&t->d_b0, sizeof(double) * n) != hipSuccess ||
            hipMalloc(&t->d_b, sizeof(double) * n) != hipSuccess) {
            cg_tissue_destroy(t);
            return rc == 1 ? CG_ERR_ARG : CG_ERR_NOMEM;
        }
    } else if (cfg->algorithm == CG_IMPLICIT) {
        int rc = cg::implicit_create(&t->implicit, m->g, m->h_sigma9.data(), cfg->chi * cfg->cm,
                                     cfg->dt, ion->prec, ion->stream, &g_err);
        if (rc != 0) {
            cg_tissue_destroy(t);
            return CG_ERR_NOMEM;
        }
    }
    *out = t;
    return CG_OK;
}

int cg_tissue_destroy(cg_tissue* t)
{
    if (!t) return CG_OK;
    // the ionic handle may already be gone: a destroy touches its own handle only
    hipDeviceSynchronize();
    for (int q = 0; q < 3; ++q) hipFree(t->d_V[q]);
    hipFree(t->d_Iion); hipFree(t->d_flag); hipFree(t->d_tact);
    if (t->d_box_mask) hipFree(t->d_box_mask);
    if (t->d_ecg) hipFree(t->d_ecg);
    if (t->gldu) cg::ldu_destroy(t->gldu);
    cg::general_destroy(t->own_gm);
    if (t->d_src) hipFree(t->d_src);
    if (t->d_b0) hipFree(t->d_b0);
    if (t->d_b) hipFree(t->d_b);
    if (t->implicit) cg::implicit_destroy(t->implicit);
    delete t;
    return CG_OK;
}

int cg_tissue_set_stimulus_boxes(cg_tissue* t, int n, const double* lo, const double* hi,
                                 const double* t0, const double* dur, const double* I)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (n < 0 || n > cg::MAX_BOXES) return fail(CG_ERR_ARG, "at most %d stimulus boxes", cg::MAX_BOXES);
    if (n > 0 && (!lo || !hi || !t0 || !dur || !I)) return fail(CG_ERR_ARG, "NULL box array");
    for (int b = 0; b < n; ++b) {
        if (dur[b] < 0.0)
            return fail(CG_ERR_STIMULUS, "stimulusDuration must be non-negative. Found value %g at index %d.",
                        dur[b], b);
        for (int d = 0; d < 3; ++d) { t->box_lo[b][d] = lo[3 * b + d]; t->box_hi[b][d] = hi[3 * b + d]; }
        t->box_t0[b] = t0[b]; t->box_dur[b] = dur[b]; t->box_I[b] = I[b];
    }
    t->n_boxes = n;
    ++t->boxes_ver;
    return CG_OK;
}

// The membership mask of the boxes over the mesh's own centres (hex block given them):
// myocardiumDomain.C:307-342 tests mesh().C() with the inclusive boundBox::contains.
static int ensure_box_mask(cg_tissue* t)
{
    const cg_mesh* m = t->mesh;
    if (m->gm || m->h_centres.empty()) return CG_OK;
    if (t->d_box_mask && t->mask_boxes_ver == t->boxes_ver && t->mask_centres_ver == m->centres_ver)
        return CG_OK;
    const int64_t n = m->n;
    std::vector<uint8_t> mask((size_t)n, 0);
    const double* C = m->h_centres.data();
    for (int64_t c = 0; c < n; ++c) {
        unsigned bits = 0;
        for (int b = 0; b < t->n_boxes; ++b) {
            bool in = true;
            for (int d = 0; d < 3; ++d)
                in = in && C[3 * c + d] >= t->box_lo[b][d] && C[3 * c + d] <= t->box_hi[b][d];
            if (in) bits |= 1u << b;
        }
        mask[(size_t)c] = (uint8_t)bits;
    }
    if (!t->d_box_mask && hipMalloc(&t->d_box_mask, (size_t)n) != hipSuccess) {
        t->d_box_mask = nullptr;
        return fail(CG_ERR_NOMEM, "box mask allocation failed");
    }
    HIPCHK(hipStreamSynchronize(t->ion->stream));
    HIPCHK(hipMemcpy(t->d_box_mask, mask.data(), (size_t)n, hipMemcpyHostToDevice));
    t->mask_boxes_ver = t->boxes_ver;
    t->mask_centres_ver = m->centres_ver;
    return CG_OK;
}

static cg::ActiveBoxes active_boxes(const cg_tissue* t, double t0)
{
    cg::ActiveBoxes a{};
    const bool use_mask = !t->mesh->gm && !t->mesh->h_centres.empty();
    a.mask = use_mask ? t->d_box_mask : nullptr;
    for (int b = 0; b < t->n_boxes; ++b) {
        const double tS = t->box_t0[b], tE = tS + t->box_dur[b];
        if (t0 < tS || t0 > tE) continue;
        for (int d = 0; d < 3; ++d) { a.lo[a.n][d] = t->box_lo[b][d]; a.hi[a.n][d] = t->box_hi[b][d]; }
        a.intensity[a.n] = t->box_I[b];
        a.bit[a.n] = (uint8_t)b;
        ++a.n;
    }
    return a;
}

// The step on a general mesh (myocardiumDomain.C:487-529 with monodomainSolver.C:103-150):
// the ionic kernel of cg_ionic_step on the device Vm, then the explicit update, or the
// fvMatrix solved by the lduMatrix PCG with OpenFOAM's norm inside the PIMPLE corrector
// loop (the tensor and non-orthogonal correction re-evaluated from the latest Vm).
static int general_advance(cg_tissue* t, cg::GeneralMesh* G, int n_steps, bool ionic,
                           const double* src_dev)
{
    cg_ionic* ion = t->ion;
    cg_mesh* m = t->mesh;
    hipStream_t s = ion->stream;
    const double dt = t->cfg.dt, chiCm = t->cfg.chi * t->cfg.cm;
    const double dtm = model_dt(ion->model, dt);
    cg::OdeArgs ode0;
    if (int rc = ode_args(ion, 0.0, dt, &ode0)) return rc;
    std::vector<double> diag, upper;
    for (int q = 0; q < n_steps; ++q) {
        const double tn = t->t + dt;
        const double t0 = tn - dt;
        const bool bdf2 = t->cfg.ddt == CG_DDT_BACKWARD && t->steps >= 1;
        cg::GeneralStep st{};
        st.dt = dt; st.chiCm = chiCm; st.t_old = t0; st.act_v = t->act_v;
        st.c = bdf2 ? 1.5 : 1.0; st.c0 = bdf2 ? 2.0 : 1.0; st.c00 = bdf2 ? 0.5 : 0.0;
        st.boxes = active_boxes(t, t0);
        st.src = src_dev;
        cg::OdeArgs ode = ode0;
        if (ode.method) model_step_times(ion->model, t0, dt, &ode.t0, &ode.t1);
        double* Vn = t->d_V[t->cur];
        double* Vm1 = t->d_V[(t->cur + 2) % 3];
        double* Vnew = t->d_V[(t->cur + 1) % 3];
        cg::Profiler* prof = t->prof.enabled ? &t->prof : nullptr;
        int tk = -1;
        if (ionic) {
            tk = prof ? prof->begin(cg::K_IONIC) : -1;
            HIPCHK(cg::launch_ionic_tissue(ion->model, ion->prec, ion->tissue, ion->d_states,
                                           ion->n, Vn, t->d_Iion, dtm, ion->P, ion->d_gscale,
                                           ode, s));
            if (prof) prof->end(tk);
        }
        if (t->cfg.algorithm == CG_EXPLICIT) {
            tk = prof ? prof->begin(cg::K_APPLY) : -1;
            HIPCHK(cg::general_explicit(G, Vn, Vm1, Vnew, t->d_Iion, t->d_flag, t->d_tact,
                                        st, s));
            if (prof) prof->end(tk);
        } else {
            if (t->g_c != st.c) {
                diag.resize(m->n);
                upper.resize(cg::general_faces(G));
                cg::general_matrix(G, chiCm, st.c, dt, diag.data(), upper.data());
                if (cg::ldu_set_coeffs(t->gldu, diag.data(), upper.data(), false))
                    return fail(CG_ERR_HIP, "matrix upload failed");
                t->g_c = st.c;
            }
            HIPCHK(cg::general_rhs0(G, Vn, Vm1, t->d_Iion, t->d_b0, st, s));
            // pimpleLoop: after the first corrector, an initial residual of the previous
            // solve below residualControl makes the current corrector the last
            const int n_outer = t->g_outer > 1 ? t->g_outer : 1;
            // the solve starts from Vn, as the reference's; with a single corrector (no
            // residual-controlled loop that its initial residual would steer) from the
            // extrapolation 2 Vn - Vn-1, which changes the iterations, not the solution
            HIPCHK(cg::general_guess(G, Vn, Vm1, Vnew, n_outer == 1 && t->steps >= 1, s));
            double prev = 0.0;
            int solves = 0;
            for (int outer = 1;; ++outer) {
                const bool last = n_outer == 1 || outer >= n_outer ||
                                  (outer > 1 && prev < t->g_rc);
                // the explicit correction of the first corrector is the one of Vn (the
                // field the equation is assembled from), later ones of the latest solution
                HIPCHK(cg::general_rhs_corr(G, outer == 1 ? Vn : Vnew, t->d_b0, t->d_b, s));
                HIPCHK(hipStreamSynchronize(s));
                int it = 0;
                double r0 = 0.0, r1 = 0.0;
                if (cg::ldu_solve(t->gldu, Vnew, t->d_b, true, t->g_tol, t->g_rel,
                                  t->g_max_iter, &it, &r0, &r1))
                    return fail(CG_ERR_HIP, "lduMatrix solve failed");
                t->g_iters += it;
                t->g_r0 = r0;
                t->g_r1 = r1;
                prev = r0;
                ++solves;
                if (last) break;
            }
            t->g_solves = solves;
            HIPCHK(cg::general_activation(G, Vn, Vnew, t->d_flag, t->d_tact, t0, dt, t->act_v, s));
        }
        t->cur = (t->cur + 1) % 3;
        t->t = tn;
        t->steps++;
    }
    return CG_OK;
}

int cg_tissue_diffusion_step(cg_tissue* t, const double* Iion, const double* source, int flags)
{
    if (!t || !Iion) return fail(CG_ERR_ARG, "NULL argument");
    if (int rc = ensure_box_mask(t)) return rc;
    cg_mesh* m = t->mesh;
    hipStream_t s = t->ion->stream;
    const size_t bytes = sizeof(double) * m->n;
    cg::GeneralMesh* G = m->gm;
    if (!G) {
        if (!t->own_gm) {
            const int rc = cg::general_from_hex(&t->own_gm, m->g, &g_err);
            if (rc) return rc == 5 ? CG_ERR_NOMEM : CG_ERR_ARG;
            if (cg::general_set_sigma(t->own_gm, m->h_sigma9.data(), &g_err)) return CG_ERR_HIP;
        }
        G = t->own_gm;
    }
    if (t->cfg.algorithm == CG_IMPLICIT && !t->gldu) {
        const int rc = cg::ldu_create(&t->gldu, m->n, cg::general_faces(G), cg::general_owner(G),
                                      cg::general_neighbour(G), &g_err);
        if (rc) return rc == 1 ? CG_ERR_ARG : CG_ERR_NOMEM;
        if (hipMalloc(&t->d_b0, bytes) != hipSuccess || hipMalloc(&t->d_b, bytes) != hipSuccess)
            return fail(CG_ERR_NOMEM, "right-hand side allocation failed");
    }
    if (source && !t->d_src && hipMalloc(&t->d_src, bytes) != hipSuccess)
        return fail(CG_ERR_NOMEM, "source allocation failed");
    const hipMemcpyKind k = (flags & CG_DEVICE_PTRS) ? hipMemcpyDeviceToDevice
                                                    : hipMemcpyHostToDevice;
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpyAsync(t->d_Iion, Iion, bytes, k, s));
    if (source) HIPCHK(hipMemcpyAsync(t->d_src, source, bytes, k, s));
    if (int rc = general_advance(t, G, 1, false, source ? t->d_src : nullptr)) return rc;
    HIPCHK(hipStreamSynchronize(s));
    return CG_OK;
}

int cg_tissue_step(cg_tissue* t, int n_steps)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (t->mesh->gm) return general_advance(t, t->mesh->gm, n_steps, true, nullptr);
    if (int rc = ensure_box_mask(t)) return rc;
    cg_ionic* ion = t->ion;
    cg_mesh* m = t->mesh;
    const double dt = t->cfg.dt, chiCm = t->cfg.chi * t->cfg.cm;
    const double dtm = model_dt(ion->model, dt);
    const cg::HexCoefs coefs = mesh_coefs(m);
    // constants of the step that starts at accumulated time t_before after steps_before
    // steps: t += dt; t0 = t - dt; prepareTimeStep(t0) (myocardiumDomain.C:487-495)
    cg::OdeArgs ode0;
    if (int rc = ode_args(ion, 0.0, dt, &ode0)) return rc;  // fills the trial steps once
    auto make_step = [&](double t_before, int64_t steps_before) {
        const double tn = t_before + dt;
        const double t0 = tn - dt;
        const bool bdf2 = t->cfg.ddt == CG_DDT_BACKWARD && steps_before >= 1;
        cg::ExplicitStep st{};
        st.dt = dt; st.dt_model = dtm; st.chiCm = chiCm;
        st.c = bdf2 ? 1.5 : 1.0; st.c0 = bdf2 ? 2.0 : 1.0; st.c00 = bdf2 ? 0.5 : 0.0;
        st.t_old = t0;
        st.act_v = t->act_v;
        st.boxes = active_boxes(t, t0);
        st.od = m->d_od;
        st.od_mask = m->od_mask;
        st.ode = ode0;
        if (ode0.method) model_step_times(ion->model, t0, dt, &st.ode.t0, &st.ode.t1);
        return st;
    };
    bool ionic_queued = false;  // this step's ionic launch was queued by the previous one
    for (int s = 0; s < n_steps; ++s) {
        const double tn = t->t + dt;
        const cg::ExplicitStep st = make_step(t->t, t->steps);
        double* Vn = t->d_V[t->cur];
        double* Vm1 = t->d_V[(t->cur + 2) % 3];
        double* Vnew = t->d_V[(t->cur + 1) % 3];
        cg::Profiler* prof = t->prof.enabled ? &t->prof : nullptr;
        if (t->cfg.algorithm == CG_EXPLICIT) {
            const int tk = prof ? prof->begin(cg::K_IONIC) : -1;
            HIPCHK(cg::launch_explicit_hex(ion->model, ion->prec, ion->tissue, m->g, coefs,
                                           ion->d_states, Vn, Vm1, Vnew, t->d_Iion, t->d_flag,
                                           t->d_tact, st, ion->P, ion->stream));
            if (prof) prof->end(tk);
        } else {
            cg::NextIonic nx;
            const bool more = s + 1 < n_steps;
            if (more) nx.st = make_step(tn, t->steps + 1);
            int rc = cg::implicit_step(t->implicit, ion->model, ion->prec, ion->tissue,
                                       ion->d_states, ion->P, Vn, Vm1, Vnew, t->d_Iion,
                                       t->d_flag, t->d_tact, st, t->cfg.tol, t->cfg.max_iter,
                                       prof, ionic_queued, more ? &nx : nullptr,
                                       t->steps >= 1, &g_err);
            if (rc != 0) return rc;
            ionic_queued = more;
        }
        t->cur = (t->cur + 1) % 3;
        t->t = tn;
        t->steps++;
    }
    return CG_OK;
}

int cg_tissue_set_activation_threshold(cg_tissue* t, double v_threshold)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (!(std::fabs(v_threshold) < 1.0)) return fail(CG_ERR_ARG, "threshold out of range");
    t->act_v = v_threshold;
    return CG_OK;
}

int cg_tissue_sync(cg_tissue* t)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    HIPCHK(hipStreamSynchronize(t->ion->stream));
    return check_ode(t->ion);
}

static double* field_ptr(cg_tissue* t, int which)
{
    switch (which) {
    case CG_FIELD_VM: return t->d_V[t->cur];
    case CG_FIELD_IION: return t->d_Iion;
    case CG_FIELD_ACTIVATION: return t->d_tact;
    case CG_FIELD_VM_OLD: return t->d_V[(t->cur + 2) % 3];
    default: return nullptr;
    }
}

int cg_tissue_get_field(cg_tissue* t, int which, double* host)
{
    if (!t || !host) return fail(CG_ERR_ARG, "NULL argument");
    double* d = field_ptr(t, which);
    if (!d) return fail(CG_ERR_ARG, "unknown field %d", which);
    HIPCHK(hipStreamSynchronize(t->ion->stream));
    if (int rc = check_ode(t->ion)) return rc;
    HIPCHK(hipMemcpy(host, d, sizeof(double) * t->mesh->n, hipMemcpyDeviceToHost));
    return CG_OK;
}

// Per-cell conductivity of a non-uniform mesh on the device, component-major ([6][n]
// symmetric or [9][n]) so that lanes read consecutive cells; made on the first ECG call.
static int ensure_sigma_dev(cg_mesh* m)
{
    if (m->uniform || m->d_sigma9) return CG_OK;
    const std::vector<double>& s9 = m->h_sigma9;
    bool sym = true;
    for (int64_t c = 0; c < m->n && sym; ++c) {
        const double* q = &s9[9 * c];
        sym = q[1] == q[3] && q[2] == q[6] && q[5] == q[7];
    }
    static const int k6[6] = {0, 1, 2, 4, 5, 8};
    const int nc = sym ? 6 : 9;
    std::vector<double> soa((size_t)nc * m->n);
    for (int q = 0; q < nc; ++q)
        for (int64_t c = 0; c < m->n; ++c)
            soa[(size_t)q * m->n + c] = s9[9 * c + (sym ? k6[q] : q)];
    const size_t bytes = sizeof(double) * soa.size();
    if (hipMalloc(&m->d_sigma9, bytes) != hipSuccess) {
        m->d_sigma9 = nullptr;
        return fail(CG_ERR_NOMEM, "conductivity allocation failed");
    }
    HIPCHK(hipMemcpy(m->d_sigma9, soa.data(), bytes, hipMemcpyHostToDevice));
    m->sigma_sym = sym;
    return CG_OK;
}

static int check_ecg_args(const cg_tissue* t, int n_electrodes, const double* positions,
                          const double* values, int max_electrodes)
{
    if (!t || (n_electrodes > 0 && (!positions || !values)))
        return fail(CG_ERR_ARG, "NULL argument");
    if (n_electrodes < 0 || n_electrodes > max_electrodes)
        return fail(CG_ERR_ARG, "between 0 and %d electrodes", max_electrodes);
    for (int i = 0; i < 3 * n_electrodes; ++i)
        if (!std::isfinite(positions[i])) return fail(CG_ERR_ARG, "electrode position not finite");
    return CG_OK;
}

// The ECG kernels on the tissue's current Vm with the given device buffers: positions
// are copied into pos, the sums come back through out; synchronous.
static int run_pseudo_ecg(cg_tissue* t, int n_electrodes, const double* positions,
                          double* values, double* pos, double* out, double* partial)
{
    cg_mesh* m = t->mesh;
    hipStream_t s = t->ion->stream;
    HIPCHK(hipStreamSynchronize(s));
    if (int rc = check_ode(t->ion)) return rc;
    HIPCHK(hipMemcpyAsync(pos, positions, sizeof(double) * 3 * n_electrodes,
                          hipMemcpyHostToDevice, s));
    cg::Profiler* prof = t->prof.enabled ? &t->prof : nullptr;
    const int tk = prof ? prof->begin(cg::K_ECG) : -1;
    if (m->gm)
        HIPCHK(cg::general_pseudo_ecg(m->gm, t->d_V[t->cur], m->uniform ? nullptr : m->d_sigma9,
                                      m->sigma_sym ? 1 : 0, m->uniform ? m->sig : nullptr,
                                      n_electrodes, pos, cg::ECG_BLOCKS, partial, out, s));
    else
        HIPCHK(cg::launch_pseudo_ecg(m->g, t->d_V[t->cur], m->uniform ? nullptr : m->d_sigma9,
                                     m->sigma_sym ? 1 : 0, m->uniform ? m->sig : nullptr,
                                     n_electrodes, pos, partial, out, s));
    if (prof) prof->end(tk);
    HIPCHK(hipMemcpyAsync(values, out, sizeof(double) * n_electrodes, hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    return CG_OK;
}

int cg_tissue_pseudo_ecg(cg_tissue* t, int n_electrodes, const double* positions,
                         double* values)
{
    if (int rc = check_ecg_args(t, n_electrodes, positions, values, cg::ECG_MAX_ELECTRODES))
        return rc;
    if (n_electrodes == 0) return CG_OK;
    if (int rc = ensure_sigma_dev(t->mesh)) return rc;
    const size_t per = 3 + cg::ECG_BLOCKS + 1;
    if (!t->d_ecg && hipMalloc(&t->d_ecg, sizeof(double) * per * cg::ECG_MAX_ELECTRODES) !=
                         hipSuccess) {
        t->d_ecg = nullptr;
        return fail(CG_ERR_NOMEM, "ECG buffer allocation failed");
    }
    double* pos = t->d_ecg;
    double* out = pos + 3 * cg::ECG_MAX_ELECTRODES;
    double* partial = out + cg::ECG_MAX_ELECTRODES;
    return run_pseudo_ecg(t, n_electrodes, positions, values, pos, out, partial);
}

int cg_tissue_pseudo_ecg_noalloc(cg_tissue* t, int n_electrodes, const double* positions,
                                 double* values)
{
    if (int rc = check_ecg_args(t, n_electrodes, positions, values,
                                cg::ECG_NOALLOC_ELECTRODES))
        return rc;
    if (n_electrodes == 0) return CG_OK;
    if (!t->mesh->uniform && !t->mesh->d_sigma9)
        return fail(CG_ERR_STATE, "the per-cell conductivity is not on the device yet: one "
                    "cg_tissue_pseudo_ecg call (which uploads it) must come first");
    double* scr = nullptr;
    if (hipError_t e = cg::ecg_scratch(&scr))
        return fail(CG_ERR_HIP, "scratch address lookup failed: %s", hipGetErrorString(e));
    double* pos = scr;
    double* out = pos + 3 * cg::ECG_NOALLOC_ELECTRODES;
    double* partial = out + cg::ECG_NOALLOC_ELECTRODES;
    return run_pseudo_ecg(t, n_electrodes, positions, values, pos, out, partial);
}

int cg_tissue_set_field(cg_tissue* t, int which, const double* host)
{
    if (!t || !host) return fail(CG_ERR_ARG, "NULL argument");
    double* d = field_ptr(t, which);
    if (!d) return fail(CG_ERR_ARG, "unknown field %d", which);
    HIPCHK(hipStreamSynchronize(t->ion->stream));
    HIPCHK(hipMemcpy(d, host, sizeof(double) * t->mesh->n, hipMemcpyHostToDevice));
    return CG_OK;
}

int cg_tissue_time(const cg_tissue* t, double* time_s, int64_t* steps)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (time_s) *time_s = t->t;
    if (steps) *steps = t->steps;
    return CG_OK;
}

int cg_tissue_set_profiling(cg_tissue* t, int on)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    t->prof.reset();
    t->prof.enabled = on != 0;
    return CG_OK;
}

int cg_tissue_kernel_times(cg_tissue* t, int max, char* names, double* total_ms,
                           int64_t* launches, int* n_out)
{
    if (!t || !n_out) return fail(CG_ERR_ARG, "NULL argument");
    t->prof.resolve(true);
    int k = 0;
    for (int id = 0; id < cg::K_COUNT && k < max; ++id) {
        if (t->prof.launches[id] == 0) continue;
        if (names) snprintf(names + 32 * k, 32, "%s", cg::kernel_name(id));
        if (total_ms) total_ms[k] = t->prof.total_ms[id];
        if (launches) launches[k] = t->prof.launches[id];
        ++k;
    }
    *n_out = k;
    return CG_OK;
}

int cg_tissue_solver_stats(const cg_tissue* t, int64_t* iters, double* res)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (t->gldu) {
        if (iters) *iters = t->g_iters;
        if (res) *res = t->g_r1;
        return CG_OK;
    }
    if (!t->implicit) return fail(CG_ERR_STATE, "explicit tissue has no linear solver");
    cg::implicit_stats(t->implicit, iters, res);
    return CG_OK;
}

int cg_tissue_set_solver_controls(cg_tissue* t, int norm, double tolerance, double rel_tol,
                                  int max_iter, int n_outer_correctors,
                                  double residual_control)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (!t->implicit && !t->gldu) return fail(CG_ERR_STATE, "explicit tissue has no linear solver");
    if (norm != CG_NORM_L2 && norm != CG_NORM_OPENFOAM)
        return fail(CG_ERR_ARG, "unknown residual norm %d", norm);
    if (!(tolerance >= 0.0) || !(rel_tol >= 0.0) || max_iter < 1 || !(residual_control >= 0.0))
        return fail(CG_ERR_ARG, "bad solver controls");
    if (t->mesh->gm && norm != CG_NORM_OPENFOAM)
        return fail(CG_ERR_ARG, "a general mesh solves in OpenFOAM's residual norm only");
    if (norm == CG_NORM_OPENFOAM) {  // also the controls of cg_tissue_diffusion_step
        t->g_tol = tolerance;
        t->g_rel = rel_tol;
        t->g_max_iter = max_iter;
        t->g_outer = n_outer_correctors > 1 ? n_outer_correctors : 1;
        t->g_rc = residual_control;
    }
    if (t->mesh->gm) return CG_OK;
    if (norm == CG_NORM_L2 && !(tolerance > 0.0))
        return fail(CG_ERR_ARG, "the relative L2 norm needs a positive tolerance");
    HIPCHK(hipStreamSynchronize(t->ion->stream));
    cg::SolveControls c;
    c.norm = norm;
    c.rel_tol = rel_tol;
    c.n_outer = n_outer_correctors;
    c.residual_control = residual_control;
    const int rc = cg::implicit_set_controls(t->implicit, c);
    if (rc == 5) return fail(CG_ERR_NOMEM, "allocation for the outer correctors failed");
    if (rc) return fail(CG_ERR_ARG, "bad solver controls");
    t->cfg.tol = tolerance;
    t->cfg.max_iter = max_iter;
    return CG_OK;
}

int cg_tissue_solver_report(const cg_tissue* t, int* solves, double* init_res, double* final_res)
{
    if (!t) return fail(CG_ERR_ARG, "NULL handle");
    if (t->gldu) {
        if (solves) *solves = t->g_solves;
        if (init_res) *init_res = t->g_r0;
        if (final_res) *final_res = t->g_r1;
        return CG_OK;
    }
    if (!t->implicit) return fail(CG_ERR_STATE, "explicit tissue has no linear solver");
    cg::implicit_report(t->implicit, solves, init_res, final_res);
    return CG_OK;
}

// --------------------------------------------------------------- lduMatrix
struct cg_ldu {
    cg::LduSolver* L = nullptr;
    bool coeffs = false;
};

int cg_ldu_create(int64_t n_cells, int64_t n_faces, const int32_t* lower, const int32_t* upper,
                  cg_ldu** out)
{
    if (!out || (n_faces > 0 && (!lower || !upper))) return fail(CG_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (n_cells < 1 || n_faces < 0 || n_cells >= ((int64_t)1 << 31))
        return fail(CG_ERR_ARG, "bad lduAddressing sizes");
    cg_ldu* h = new cg_ldu();
    const int rc = cg::ldu_create(&h->L, n_cells, n_faces, lower, upper, &g_err);
    if (rc) {
        delete h;
        return rc == 5 ? CG_ERR_NOMEM : CG_ERR_ARG;
    }
    *out = h;
    return CG_OK;
}

int cg_ldu_destroy(cg_ldu* h)
{
    if (!h) return CG_OK;
    cg::ldu_destroy(h->L);
    delete h;
    return CG_OK;
}

int cg_ldu_width(const cg_ldu* h) { return h ? cg::ldu_width(h->L) : -1; }

int cg_ldu_set_coeffs(cg_ldu* h, const double* diag, const double* upper, int flags)
{
    if (!h || !diag) return fail(CG_ERR_ARG, "NULL argument");
    if (cg::ldu_set_coeffs(h->L, diag, upper, flags & CG_DEVICE_PTRS))
        return fail(CG_ERR_HIP, "coefficient upload failed");
    h->coeffs = true;
    return CG_OK;
}

int cg_ldu_amul(cg_ldu* h, const double* x, double* y, int flags)
{
    if (!h || !x || !y) return fail(CG_ERR_ARG, "NULL argument");
    if (!h->coeffs) return fail(CG_ERR_STATE, "coefficients not set");
    if (cg::ldu_amul(h->L, x, y, flags & CG_DEVICE_PTRS, nullptr, nullptr))
        return fail(CG_ERR_HIP, "lduMatrix Amul failed");
    return CG_OK;
}

int cg_ldu_amul_timed(cg_ldu* h, const double* x, double* y, float* ms)
{
    if (!h || !x || !y || !ms) return fail(CG_ERR_ARG, "NULL argument");
    if (!h->coeffs) return fail(CG_ERR_STATE, "coefficients not set");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int rc = cg::ldu_amul(h->L, x, y, true, e0, e1);
    if (rc == 0) hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc ? fail(CG_ERR_HIP, "lduMatrix Amul failed") : CG_OK;
}

int cg_ldu_solve(cg_ldu* h, double* psi, const double* source, double tol, double rel_tol,
                 int max_iter, int flags, int* iters, double* r0, double* r1)
{
    if (!h || !psi || !source) return fail(CG_ERR_ARG, "NULL argument");
    if (!h->coeffs) return fail(CG_ERR_STATE, "coefficients not set");
    if (!(tol >= 0.0) || !(rel_tol >= 0.0) || max_iter < 0)
        return fail(CG_ERR_ARG, "bad solver controls");
    if (cg::ldu_solve(h->L, psi, source, flags & CG_DEVICE_PTRS, tol, rel_tol, max_iter, iters,
                      r0, r1))
        return fail(CG_ERR_HIP, "lduMatrix PCG failed");
    return CG_OK;
}

// ------------------------------------------------------------------ adjoint
struct cg_adjoint {
    cg_tissue* t;
    cg::Adjoint* A;
    double t0;  // tissue time at creation
    int n_probes;
    int64_t generation = 0;  // number of the latest forward run (0: none yet)
    int last_steps = 0;      // its step count
};

int cg_adjoint_create(cg_tissue* t, int n_probes, const int64_t* probe_cells,
                      int checkpoint_every, cg_adjoint** out)
{
    if (!t || !out || (n_probes > 0 && !probe_cells)) return fail(CG_ERR_ARG, "NULL argument");
    *out = nullptr;
    cg_ionic* ion = t->ion;
    cg_mesh* m = t->mesh;
    if (m->gm) return fail(CG_ERR_STATE, "the adjoint covers the hex block");
    if (t->cfg.ddt != CG_DDT_EULER)
        return fail(CG_ERR_STATE, "the adjoint covers the Euler-ddt tissue step");
    if (ion->model != cg::MODEL_BO || ion->prec != CG_FP64)
        return fail(CG_ERR_STATE, "the adjoint covers BuenoOrovio in FP64");
    if (t->ion->ode_kind != CG_ODE_RUSH_LARSEN)
        return fail(CG_ERR_STATE, "the adjoint differentiates the Rush-Larsen ionic step");
    if (int rc = ensure_box_mask(t)) return rc;
    if (!m->uniform) return fail(CG_ERR_STATE, "the adjoint scales a uniform conductivity");
    if (m->sig[1] != 0 || m->sig[2] != 0 || m->sig[5] != 0 || m->sig[3] != 0 ||
        m->sig[6] != 0 || m->sig[7] != 0)
        return fail(CG_ERR_STATE, "the adjoint needs a diagonal conductivity");
    if (n_probes < 1 || checkpoint_every < 1) return fail(CG_ERR_ARG, "bad probe count or interval");
    HIPCHK(hipStreamSynchronize(ion->stream));
    const double diag[3] = {m->sig[0], m->sig[4], m->sig[8]};
    cg::ActiveBoxesAll bx{};
    bx.n = t->n_boxes;
    bx.mask = m->h_centres.empty() ? nullptr : t->d_box_mask;
    for (int b = 0; b < t->n_boxes; ++b) {
        for (int d = 0; d < 3; ++d) { bx.lo[b][d] = t->box_lo[b][d]; bx.hi[b][d] = t->box_hi[b][d]; }
        bx.t0[b] = t->box_t0[b];
        bx.dur[b] = t->box_dur[b];
        bx.I[b] = t->box_I[b];
    }
    std::string err;
    cg::Adjoint* A = nullptr;
    const int rc = cg::adjoint_create(&A, m->g, diag, t->cfg.dt, t->cfg.chi * t->cfg.cm, ion->P,
                                      t->d_V[t->cur], ion->d_states, bx, checkpoint_every,
                                      n_probes, probe_cells, t->cfg.algorithm == CG_IMPLICIT,
                                      ion->stream, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    cg_adjoint* a = new cg_adjoint{t, A, t->t, n_probes};
    *out = a;
    return CG_OK;
}

int cg_adjoint_destroy(cg_adjoint* a)
{
    if (!a) return CG_OK;
    hipDeviceSynchronize();   // the tissue and ionic handles may already be gone
    cg::adjoint_destroy(a->A);
    delete a;
    return CG_OK;
}

int cg_adjoint_forward(cg_adjoint* a, const double* kappa_dev, const double* theta, int n_steps,
                       double* probes_dev, int64_t* generation)
{
    if (!a || !kappa_dev || !probes_dev) return fail(CG_ERR_ARG, "NULL argument");
    if (n_steps < 1) return fail(CG_ERR_ARG, "n_steps must be positive");
    // whatever happens below, the previous run's checkpoints are no longer valid
    ++a->generation;
    a->last_steps = 0;
    // start time of each step, with the tissue's time accumulation
    std::vector<double> t0((size_t)n_steps);
    double tt = a->t0;
    for (int s = 0; s < n_steps; ++s) {
        const double tn = tt + a->t->cfg.dt;
        t0[s] = tn - a->t->cfg.dt;
        tt = tn;
    }
    const int idx[2] = {cg::BuenoOrovio::tfi, cg::BuenoOrovio::tsi};
    std::string err;
    const size_t budget = (size_t)64 << 30;  // every step kept up to 64 GiB of HBM
    const int rc = cg::adjoint_forward(a->A, kappa_dev, theta, idx, theta ? 2 : 0, n_steps, t0,
                                       probes_dev, budget, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    a->last_steps = n_steps;
    if (generation) *generation = a->generation;
    return CG_OK;
}

int cg_adjoint_backward(cg_adjoint* a, int64_t generation, int n_steps,
                        const double* dprobes_dev, double* dkappa_dev, double* dtheta)
{
    if (!a || !dprobes_dev || !dkappa_dev || !dtheta) return fail(CG_ERR_ARG, "NULL argument");
    if (a->last_steps == 0 || generation != a->generation)
        return fail(CG_ERR_STATE, "backward of forward run %lld, but the latest completed run "
                    "is %lld: its checkpoints were replaced", (long long)generation,
                    a->last_steps ? (long long)a->generation : 0LL);
    if (n_steps != a->last_steps)
        return fail(CG_ERR_STATE, "backward over %d steps of a forward run of %d steps", n_steps,
                    a->last_steps);
    std::string err;
    const int rc = cg::adjoint_backward(a->A, dprobes_dev, dkappa_dev, dtheta, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    return CG_OK;
}

int cg_tissue_apply_operator(cg_tissue* t, double ddt_coeff, int precision, const void* x_dev,
                             void* y_dev)
{
    if (!t || !x_dev || !y_dev) return fail(CG_ERR_ARG, "NULL argument");
    if (t->mesh->gm) return fail(CG_ERR_STATE, "the operator entry covers the hex block");
    if (!t->implicit) return fail(CG_ERR_STATE, "explicit tissue has no implicit operator");
    if (precision != CG_FP64 && precision != CG_MIXED)
        return fail(CG_ERR_ARG, "unknown precision %d", precision);
    if (!(ddt_coeff >= 0)) return fail(CG_ERR_ARG, "ddt coefficient must be >= 0");
    cg::Profiler* prof = t->prof.enabled ? &t->prof : nullptr;
    const int rc = cg::implicit_apply(t->implicit, precision, ddt_coeff, x_dev, y_dev, prof);
    if (rc == 9) return fail(CG_ERR_STATE, "fp32 operator exists on a mixed-precision tissue only");
    if (rc != 0) return fail(CG_ERR_HIP, "operator launch failed");
    return CG_OK;
}

// ---------------------------------------------------------------- ASCII formats
int cg_format_scalars(const double* values, int64_t n, int precision, int fixed, int64_t row,
                      char* buf, int64_t cap, int64_t* len)
{
    if (!len || (n > 0 && (!values || !buf))) return fail(CG_ERR_ARG, "NULL argument");
    if (n < 0 || cap < 0 || row < 1) return fail(CG_ERR_ARG, "bad size");
    if (precision < 0 || precision > 30) return fail(CG_ERR_ARG, "precision %d out of range",
                                                     precision);
    *len = 0;
    if (!cg::format_scalars(values, n, precision, fixed != 0, row, buf, cap, len))
        return fail(CG_ERR_ARG, "output buffer of %lld bytes is too small", (long long)cap);
    return CG_OK;
}

int cg_parse_scalars(const char* text, int64_t len, double* out, int64_t n, int64_t* consumed)
{
    if (!consumed || (n > 0 && (!text || !out))) return fail(CG_ERR_ARG, "NULL argument");
    if (n < 0 || len < 0) return fail(CG_ERR_ARG, "negative size");
    *consumed = 0;
    const int rc = cg::parse_scalars(text, len, out, n, consumed);
    if (rc == 1) return fail(CG_ERR_ARG, "malformed number in list");
    if (rc == 2) return fail(CG_ERR_ARG, "list holds fewer than %lld numbers", (long long)n);
    return CG_OK;
}

} // extern "C"
