// FP64 instantiations of the model kernels (IEEE division, exp and log): the plain
// double-precision path of the tissue step and of the single-cell runs.
#include "model_kernels.hpp"

namespace cg {
CG_DEFINE_MODEL_ENTRIES(double, fp64)
} // namespace cg

#include "batch_kernels.hpp"
namespace cg {
CG_DEFINE_BATCH_ENTRY(double, fp64)

// Early-stroke classification and APD90/DI of each cell's beat records
// (classify_early_strokes and compute_apd_di of postProcessing_restCurves.py:203-297):
// a beat is early when the next upstroke comes before its 90 % repolarisation or it never
// repolarised; early beats lose their APD; a beat's DI is valid when the beat before it
// was not early and it repolarised (decided after the early pass, as the reference does,
// so an interrupted beat that repolarised keeps it).  The last two beats are the S1-S2 pair:
// (APD90 of S2, DI = S2 upstroke - S1 repolarisation), NaN when S1 is early, an APD or
// S2's DI is not valid, or either is not positive; also NaN when the cell detected more
// beats than its records hold (the last two are then not stored).  One lane per cell
// walks its records (model units; times scaled to s by inv_per_s).
__global__ __launch_bounds__(256) void k_restitution(const double* __restrict__ beats,
                                                     const int* __restrict__ nbeats,
                                                     int64_t n, int max_beats,
                                                     double inv_per_s, double* __restrict__ out,
                                                     uint8_t* __restrict__ flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nb_all = nbeats[i];
    const bool truncated = nb_all > max_beats;
    const int nb = truncated ? max_beats : nb_all;
    const double* r = beats + (size_t)i * max_beats * BEAT_REC;
    auto tup = [&](int k) { return r[k * BEAT_REC + 0]; };
    auto trep = [&](int k) { return r[k * BEAT_REC + 4]; };
    auto found = [&](int k) { return trep(k) == trep(k); };   // not NaN
    bool prev_early = false, s1_early = false, s1_apd = false;
    unsigned f = 0;
    for (int k = 0; k < nb; ++k) {
        const bool early = k + 1 < nb && (!found(k) || tup(k + 1) < trep(k));
        const bool valid_apd = found(k) && !early;
        const bool valid_di = k > 0 && !prev_early && found(k);
        f = (early ? RB_EARLY : 0u) | (valid_apd ? RB_VALID_APD : 0u) |
            (valid_di ? RB_VALID_DI : 0u) | (truncated ? RB_TRUNCATED : 0u);
        if (flags) flags[(size_t)i * max_beats + k] = (uint8_t)f;
        if (k == nb - 2) {
            s1_early = early;
            s1_apd = valid_apd;
        }
        prev_early = early;
    }
    if (flags)
        for (int k = nb; k < max_beats; ++k) flags[(size_t)i * max_beats + k] = 0;
    double apd = __builtin_nan(""), di = __builtin_nan("");
    if (!truncated && nb >= 2 && !s1_early && s1_apd && (f & RB_VALID_APD) &&
        (f & RB_VALID_DI)) {
        const double d = (tup(nb - 1) - trep(nb - 2)) * inv_per_s;
        const double a = (trep(nb - 1) - tup(nb - 1)) * inv_per_s;
        if (d > 0.0 && a > 0.0) {
            apd = a;
            di = d;
        }
    }
    out[2 * i] = apd;
    out[2 * i + 1] = di;
}

hipError_t launch_restitution(const double* beats, const int* nbeats, int64_t n, int max_beats,
                              double inv_per_s, double* out, uint8_t* flags, hipStream_t s)
{
    hipLaunchKernelGGL(k_restitution, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       beats, nbeats, n, max_beats, inv_per_s, out, flags);
    return hipGetLastError();
}

template <class M>
static hipError_t eval_of(int tissue, int64_t n, const double* st, const double* t, int kind,
                          int solve_v, const Stimulus& stim, const ModelParams& P,
                          const double* gsc, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_ionic_eval<M>, dim3((unsigned)((n + MBS - 1) / MBS)), dim3(MBS), 0, s,
                       tissue, n, st, t, kind, solve_v, stim, kparams<double>(P), gsc, out);
    return hipGetLastError();
}

int model_nalgebraic(int model)
{
    switch (model) {
    case MODEL_BO: return Algebraic<BuenoOrovio>::NA;
    case MODEL_AP: return Algebraic<AlievPanfilov>::NA;
    default: return 0;
    }
}

hipError_t launch_ionic_eval(int model, int tissue, int64_t n, const double* st, const double* t,
                             int kind, int solve_v, const Stimulus& stim, const ModelParams& P,
                             const double* gsc, double* out, hipStream_t s)
{
#define CG_EVAL(M) eval_of<M>(tissue, n, st, t, kind, solve_v, stim, P, gsc, out, s)
    switch (model) {
    case MODEL_BO: return CG_EVAL(BuenoOrovio);
    case MODEL_TNNP: return CG_EVAL(TNNP);
    case MODEL_ORD: return CG_EVAL(ORd);
    case MODEL_AP: return CG_EVAL(AlievPanfilov);
    case MODEL_FDA: return CG_EVAL(FdaManufactured);
    case MODEL_CRN: return CG_EVAL(Courtemanche);
    case MODEL_PYS: return CG_EVAL(PerisYague);
    case MODEL_STW: return CG_EVAL(Stewart);
    case MODEL_GRD: return CG_EVAL(Grandi);
    case MODEL_TOR: return CG_EVAL(ToRORd);
    case MODEL_GAUR: return CG_EVAL(Gaur);
    default: return hipErrorInvalidValue;
    }
#undef CG_EVAL
}

const char* batch_variant_name(int v)
{
    static const char* names[BATCH_VARIANTS] = {"256 threads", "64 threads",
                                                "256 threads, >=2 waves/SIMD",
                                                "64 threads, >=2 waves/SIMD",
                                                "256 threads, >=3 waves/SIMD", "128 threads"};
    return v >= 0 && v < BATCH_VARIANTS ? names[v] : "?";
}
} // namespace cg
