// Batched single-cell protocols: every lane runs one independent cell (Vm inside the
// ODE, SingleCellSolver::evolve, singleCellSolver.C:179-229) with its own S1/S2 stimulus
// protocol (stimulusIO.C:238-286), its own run length and its own conductance scales,
// and detects beats and APD90 on the fly.  A launch advances every lane by up to
// `chunk` steps; the host captures the chain of chunk launches of one or more models in
// one hipGraph.
//
// Time is the step count times dt (model units, FP64), never an accumulated sum.
//
// Beat detection restates detect_beats of postProcessing_restCurves.py:115-200 on the
// trace (t_j = (j+1) dt, V_j after step j+1) as a stream:
//   dv_j = (V_{j+1} - V_{j-1}) / (t_{j+1} - t_{j-1}) (centred), copies at both ends;
//   upstroke t_u where dv crosses `thr` upward (linear interpolation);
//   base = min V over [t_u - baseline, t_u) from a per-lane ring of recent samples;
//   peak = first max of V over [t_u, min(t_u + peak_window, next t_u)];
//   V90 = base + 0.1 (peak - base); t_repol90 = first downward crossing of V90 after
//   the peak inside the same window.
// A sample is assigned to beats two samples after it arrives, when every upstroke before
// it is known; a new running maximum restarts the search for the crossing, so the
// result equals the offline two-pass definition.
#pragma once
#include "model_kernels.hpp"

namespace cg {

struct Beat {
    double tu, tend, base, peak, jp, v90, trep;
    unsigned flags;
};

CG_DEV void beat_load(Beat& b, const double* det, int64_t n, int64_t lane, int f0)
{
    b.tu = det[(f0 + BF_TU) * n + lane];
    b.tend = det[(f0 + BF_TEND) * n + lane];
    b.base = det[(f0 + BF_BASE) * n + lane];
    b.peak = det[(f0 + BF_PEAK) * n + lane];
    b.jp = det[(f0 + BF_JP) * n + lane];
    b.v90 = det[(f0 + BF_V90) * n + lane];
    b.trep = det[(f0 + BF_TREP) * n + lane];
    b.flags = (unsigned)det[(f0 + BF_FLAGS) * n + lane];
}
CG_DEV void beat_store(const Beat& b, double* det, int64_t n, int64_t lane, int f0)
{
    det[(f0 + BF_TU) * n + lane] = b.tu;
    det[(f0 + BF_TEND) * n + lane] = b.tend;
    det[(f0 + BF_BASE) * n + lane] = b.base;
    det[(f0 + BF_PEAK) * n + lane] = b.peak;
    det[(f0 + BF_JP) * n + lane] = b.jp;
    det[(f0 + BF_V90) * n + lane] = b.v90;
    det[(f0 + BF_TREP) * n + lane] = b.trep;
    det[(f0 + BF_FLAGS) * n + lane] = (double)b.flags;
}

// sample p (time tp, value vp, predecessor vpm at tpm) into beat b
CG_DEV void beat_sample(Beat& b, int64_t p, double tp, double vp, double tpm, double vpm)
{
    if (!(b.flags & BEAT_OPEN) || tp < b.tu || tp > b.tend) return;
    b.flags |= BEAT_HAS;
    if (vp > b.peak) {
        b.peak = vp;
        b.jp = (double)p;
        b.v90 = b.base + 0.1 * (b.peak - b.base);
        b.flags &= ~BEAT_FOUND;
    } else if (!(b.flags & BEAT_FOUND) && (double)p > b.jp && vpm > b.v90 && b.v90 >= vp) {
        b.trep = tpm + (b.v90 - vpm) / (vp - vpm) * (tp - tpm);
        b.flags |= BEAT_FOUND;
    }
}

CG_DEV void beat_finish(Beat& b, const BatchArgs& a, int64_t cell, int& nb)
{
    if ((b.flags & BEAT_OPEN) && (b.flags & BEAT_HAS)) {
        if (nb < a.max_beats) {
            double* r = a.beats + ((size_t)cell * a.max_beats + nb) * BEAT_REC;
            r[0] = b.tu;
            r[1] = b.base;
            r[2] = b.peak;
            r[3] = b.v90;
            r[4] = (b.flags & BEAT_FOUND) ? b.trep : __builtin_nan("");
        }
        ++nb;
    }
    b.flags = 0;
}

template <class T> struct Detector {
    const BatchArgs& a;
    int64_t lane, cell;
    T* ring;
    double v1, v2, v3, dvp;
    Beat cur, old;
    int nb;

    CG_DEV double t_of(int64_t j) const { return (double)(j + 1) * a.dt; }

    CG_DEV void load()
    {
        const int64_t n = a.n;
        v1 = a.det[DS_V1 * n + lane];
        v2 = a.det[DS_V2 * n + lane];
        v3 = a.det[DS_V3 * n + lane];
        dvp = a.det[DS_DVP * n + lane];
        beat_load(cur, a.det, n, lane, DS_CUR);
        beat_load(old, a.det, n, lane, DS_OLD);
        nb = a.nbeats[cell];
    }
    CG_DEV void store()
    {
        const int64_t n = a.n;
        a.det[DS_V1 * n + lane] = v1;
        a.det[DS_V2 * n + lane] = v2;
        a.det[DS_V3 * n + lane] = v3;
        a.det[DS_DVP * n + lane] = dvp;
        beat_store(cur, a.det, n, lane, DS_CUR);
        beat_store(old, a.det, n, lane, DS_OLD);
        a.nbeats[cell] = nb;
    }

    // min of the ring samples with t in [tu - baseline, tu); samples up to index i held
    CG_DEV bool baseline(int64_t i, double tu, double& vmin) const
    {
        bool any = false;
        vmin = __builtin_inf();
        const int64_t lo = i - a.ring_len + 1 > 0 ? i - a.ring_len + 1 : 0;
        for (int64_t j = i; j >= lo; --j) {
            const double tj = t_of(j);
            if (tj >= tu) continue;
            if (tj < tu - a.baseline) break;
            const double v = (double)ring[(j % a.ring_len) * a.n + lane];
            if (v < vmin) vmin = v;
            any = true;
        }
        return any;
    }

    CG_DEV void sample_both(int64_t p, double vp, double vpm)
    {
        const double tp = t_of(p), tpm = t_of(p - 1);
        beat_sample(old, p, tp, vp, tpm, vpm);
        if ((old.flags & BEAT_OPEN) && tp >= old.tend) beat_finish(old, a, cell, nb);
        beat_sample(cur, p, tp, vp, tpm, vpm);
        if ((cur.flags & BEAT_OPEN) && tp >= cur.tend) beat_finish(cur, a, cell, nb);
    }

    // sample i (value vi) after step i + 1
    CG_DEV void push(int64_t i, double vi)
    {
        ring[(i % a.ring_len) * a.n + lane] = T(vi);
        if (i >= 2) {
            const double dv = (vi - v2) / (t_of(i) - t_of(i - 2));
            bool cand = false;
            double tu = 0.0;
            if (i >= 3 && dvp < a.thr && a.thr <= dv) {
                const double frac = (a.thr - dvp) / (dv - dvp);
                if (frac >= 0.0 && frac <= 1.0) {
                    cand = true;
                    tu = t_of(i - 2) + frac * (t_of(i - 1) - t_of(i - 2));
                }
            }
            dvp = dv;
            const int64_t p = i - 2;
            if (cand && (cur.flags & BEAT_OPEN) && tu < cur.tend) cur.tend = tu;
            sample_both(p, v2, v3);
            if (cand) {
                if (old.flags & BEAT_OPEN) beat_finish(old, a, cell, nb);
                old = cur;
                Beat b;
                b.tu = tu;
                b.tend = tu + a.peak_window;
                b.peak = -__builtin_inf();
                b.jp = -1.0;
                b.v90 = 0.0;
                b.trep = 0.0;
                b.flags = baseline(i, tu, b.base) ? BEAT_OPEN : 0u;
                cur = b;
                // the sample at exactly t_u opens the new window as well
                const double tp = t_of(p);
                beat_sample(cur, p, tp, v2, t_of(p - 1), v3);
                if ((cur.flags & BEAT_OPEN) && tp >= cur.tend) beat_finish(cur, a, cell, nb);
            }
        }
        v3 = v2;
        v2 = v1;
        v1 = vi;
    }

    // end of the trace at sample count m: the last two samples, then close every beat
    CG_DEV void flush(int64_t m)
    {
        if (m >= 2) sample_both(m - 2, v2, v3);
        if (m >= 1) sample_both(m - 1, v1, v2);
        if (old.flags & BEAT_OPEN) beat_finish(old, a, cell, nb);
        if (cur.flags & BEAT_OPEN) beat_finish(cur, a, cell, nb);
    }
};

// One adaptive step of a paced cell over model times [x0, x1] (RKF45 for INTEG 1,
// OpenFOAM's Euler for 2), the stimulus evaluated at every stage time.  FP64 integrates
// the full values y; FP32 integrates the step's increments from 0 against the step-start
// values (rates and error norm on base + d) and adds them to the stored words once, so
// increments below half an ulp of a state still count.  Returns false when the solver
// failed (step-size underflow or more than max_steps sub-steps).
template <class M, class T, int INTEG>
CG_DEV bool protocol_adaptive_step(T (&y)[M::NS], T (&sw)[M::NS], double x0, double x1,
                                   double& step, const BatchArgs& a, const Stimulus& stim,
                                   const T* p, int tissue, const T* gs)
{
    constexpr bool f32 = sizeof(T) == 4;
    T d[M::NS];
    // FP32: y holds the step-start values, unchanged until the solve ends
    auto f = [&](double t, const T* ss, T* r) __attribute__((always_inline)) {
        T yy[M::NS];
#pragma unroll
        for (int q = 0; q < M::NS; ++q) yy[q] = f32 ? y[q] + ss[q] : ss[q];
        Ops<M>::deriv(yy, r, p, tissue, gs, true, T(stimulus_current(t, stim)),
                      Ops<M>::vt(yy[0], p));
    };
    bool ok;
    if constexpr (f32) {
#pragma unroll
        for (int q = 0; q < M::NS; ++q) d[q] = T(0);
        ok = ode_solve<INTEG, M::NS>(f, d, x0, x1, step, T(a.abs_tol), T(a.rel_tol),
                                     a.max_steps, y);
#pragma unroll
        for (int q = 0; q < M::NS; ++q) {
            sw[q] = sw[q] + d[q];
            y[q] = offsets_stored<M, T>() ? T(StateRef<M>::at(q)) + sw[q] : sw[q];
        }
    } else {
        ok = ode_solve<INTEG, M::NS>(f, y, x0, x1, step, T(a.abs_tol), T(a.rel_tol),
                                     a.max_steps);
#pragma unroll
        for (int q = 0; q < M::NS; ++q) sw[q] = y[q];
    }
    return ok;
}

// One launch: lanes advance by up to a.chunk steps (stopping at their run length).
// BLOCK threads per block; WPE = minimum waves per SIMD the register budget must allow
// (0: the compiler's choice); INTEG 0 Rush-Larsen, 1 RKF45, 2 Euler.
template <class M, class T, int BLOCK, int WPE, int INTEG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_cells_protocol(T* __restrict__ st, KParams<T> P, const T* __restrict__ gsc, int tissue,
                      BatchArgs a)
{
    const int64_t lane = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (lane >= a.n) return;
    const int64_t cell = a.order ? a.order[lane] : lane;
    const int64_t len = a.len ? a.len[cell] : a.s0 + a.chunk;
    if (a.s0 >= len) return;
    const int64_t s1 = a.s0 + a.chunk < len ? a.s0 + a.chunk : len;
    const int64_t n = a.n;
    const T* p = P.p;  // model constants read from the kernel arguments
    T gs[4] = {T(1), T(1), T(1), T(1)};
    if (gsc) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gs[k] = gsc[k * n + cell];
    }
    T y[M::NS], sw[M::NS];
    load_model_states<M>(y, sw, st, n, cell, 0);
    const Stimulus stim = a.prot ? a.prot[cell] : a.stim;
    Detector<T> det{a, lane, cell, (T*)a.ring};
    if (a.det) det.load();
    if constexpr (INTEG == 0) {
        for (int64_t s = a.s0; s < s1; ++s) {
            cell_step_stored<M>(y, sw, T(a.dt), T(stimulus_current((double)s * a.dt, stim)),
                                p, tissue, gs);
            if (a.det) det.push(s, (double)y[0] * a.vscale - a.voff);
        }
    } else {
        double step = a.step[cell];
        bool ok = true;
        for (int64_t s = a.s0; s < s1 && ok; ++s) {
            if (!(step < a.dt)) step = a.dt;
            ok = protocol_adaptive_step<M, T, INTEG>(y, sw, (double)s * a.dt,
                                                     (double)(s + 1) * a.dt, step, a, stim, p,
                                                     tissue, gs);
            if (a.det) det.push(s, (double)y[0] * a.vscale - a.voff);
        }
        a.step[cell] = step;
        if (!ok) *a.fail = 1;
    }
    store_states(sw, st, n, cell, 0);
    if (a.det) {
        if (s1 == len) det.flush(len);
        det.store();
    }
}

template <class T> struct BatchLaunch {
    template <class M, int BLOCK, int WPE, int INTEG = 0>
    static hipError_t go(int tissue, void* states, const ModelParams& P, const void* gscale,
                         const BatchArgs& a, hipStream_t s)
    {
        const unsigned grid = (unsigned)((a.n + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL((k_cells_protocol<M, T, BLOCK, WPE, INTEG>), dim3(grid), dim3(BLOCK),
                           0, s, (T*)states, kparams<T>(P), (const T*)gscale, tissue, a);
        return hipGetLastError();
    }
    // the adaptive integrators run one shape, one wave per block (the lanes of a wave take
    // different numbers of sub-steps)
    template <class M>
    static hipError_t variant(int v, int tissue, void* states, const ModelParams& P,
                              const void* gscale, const BatchArgs& a, hipStream_t s)
    {
        if (a.method == 1) return go<M, 64, 0, 1>(tissue, states, P, gscale, a, s);
        if (a.method == 2) return go<M, 64, 0, 2>(tissue, states, P, gscale, a, s);
        switch (v) {
        case 0: return go<M, 256, 0>(tissue, states, P, gscale, a, s);
        case 1: return go<M, 64, 0>(tissue, states, P, gscale, a, s);
        case 2: return go<M, 256, 2>(tissue, states, P, gscale, a, s);
        case 3: return go<M, 64, 2>(tissue, states, P, gscale, a, s);
        case 4: return go<M, 256, 3>(tissue, states, P, gscale, a, s);
        case 5: return go<M, 128, 0>(tissue, states, P, gscale, a, s);
        default: return hipErrorInvalidValue;
        }
    }
};

#define CG_DEFINE_BATCH_ENTRY(T, SFX)                                                      \
    hipError_t launch_cells_protocol_##SFX(int model, int tissue, int variant, void* states, \
                                           const ModelParams& P, const void* gscale,       \
                                           const BatchArgs& a, hipStream_t s)              \
    {                                                                                      \
        switch (model) {                                                                   \
        case MODEL_BO: return BatchLaunch<T>::variant<BuenoOrovio>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_TNNP: return BatchLaunch<T>::variant<TNNP>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_ORD: return BatchLaunch<T>::variant<ORd>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_AP: return BatchLaunch<T>::variant<AlievPanfilov>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_CRN: return BatchLaunch<T>::variant<Courtemanche>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_PYS: return BatchLaunch<T>::variant<PerisYague>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_STW: return BatchLaunch<T>::variant<Stewart>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_GRD: return BatchLaunch<T>::variant<Grandi>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_TOR: return BatchLaunch<T>::variant<ToRORd>(variant, tissue, states, P, gscale, a, s); \
        case MODEL_GAUR: return BatchLaunch<T>::variant<Gaur>(variant, tissue, states, P, gscale, a, s); \
        default: return hipErrorInvalidValue;                                              \
        }                                                                                  \
    }

} // namespace cg
