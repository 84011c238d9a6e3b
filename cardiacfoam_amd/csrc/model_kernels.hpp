// Kernels that run an ionic model, templated on the model M and the arithmetic type T.
// They are instantiated in two translation units: model_fp64.hip (IEEE FP64, the plain
// path) and model_fp32.hip (FP32 with approximate reciprocal and exponential, the mixed
// path).  Launch entry points of both are declared in cg_internal.hpp.
//
// Memory layout: ionic states structure-of-arrays, state q of cell c at q*n + c, so each
// state load/store of a wavefront is one contiguous segment; one cell per lane with all
// states of the cell in registers for the whole step.
#pragma once
#include "cg_internal.hpp"
#include "cg_hexops.hpp"
#include "ode_rkf45.hpp"

namespace cg {

constexpr int MBS = 256;  // threads per block of the model kernels

// model constants already in the kernel's arithmetic type (no per-lane conversions)
template <class T> struct KParams {
    T p[MAX_PARAMS];
};
template <class T> inline KParams<T> kparams(const ModelParams& P)
{
    KParams<T> k;
    for (int i = 0; i < MAX_PARAMS; ++i) k.p[i] = T(P.p[i]);
    return k;
}

// states of one cell: state q of cell c at st[q*n + c], walked with a pointer step of n
template <int NS, class T>
CG_DEV void load_states(T (&y)[NS], const T* __restrict__ st, int64_t n, int64_t c, int first)
{
    const T* sp = st + c + first * n;
#pragma unroll
    for (int q = first; q < NS; ++q, sp += n) y[q] = *sp;
}
template <int NS, class T>
CG_DEV void store_states(const T (&y)[NS], T* __restrict__ st, int64_t n, int64_t c, int first)
{
    T* sp = st + c + first * n;
#pragma unroll
    for (int q = first; q < NS; ++q, sp += n) *sp = y[q];
}

// States of model M in the handle's storage (offsets from StateRef on FP32 handles):
// offsets_stored says whether this type stores offsets; load_model_states gives the full
// values y and the stored words s (the kernels keep both current and store s).
template <class M, class T> constexpr bool offsets_stored()
{
    return is_fp32<T>() && StateRef<M>::any;
}
template <class M, class T>
CG_DEV void load_model_states(T (&y)[M::NS], T (&s)[M::NS], const T* __restrict__ st, int64_t n,
                              int64_t c, int first)
{
    load_states(s, st, n, c, first);
#pragma unroll
    for (int q = first; q < M::NS; ++q)
        y[q] = offsets_stored<M, T>() ? T(StateRef<M>::at(q)) + s[q] : s[q];
}

// Ionic step of cell c in tissue mode with V held at v over the step, on full values in
// place; returns the total ionic current at the step end.  INTEG 0: Rush-Larsen; 1: RKF45,
// 2: OpenFOAM's adaptive Euler, over [o.t0, o.t1] from the cell's kept trial step (BuenoOrovio.C:124-132 and its
// TNNP/ORd/AP twins), the V terms computed once for all stages.
template <class M, class T, int INTEG>
CG_DEV T tissue_advance_full(T v, T (&y)[M::NS], T dt, const T* p, int tissue, const T* gs,
                             const OdeArgs& o, int64_t c)
{
    if constexpr (INTEG == 0) {
        return Ops<M>::tissue_step(v, y, dt, p, tissue, gs);
    } else {
        y[0] = v;
        const auto q = Ops<M>::vt(v, p);
        auto f = [&](double, const T* yy, T* d) __attribute__((always_inline)) {
            Ops<M>::deriv(yy, d, p, tissue, gs, false, T(0), q);
        };
        double step = o.step[c];
        if (!(step < o.dt_clamp)) step = o.dt_clamp;
        const bool ok = ode_solve<INTEG, M::NS>(f, y, o.t0, o.t1, step, T(o.abs_tol), T(o.rel_tol),
                                           o.max_steps);
        o.step[c] = step;
        if (!ok) *o.fail = 1;
        return Ops<M>::iion(y, p, tissue, gs);
    }
}

// The tissue step of one cell on the handle's storage: y enters with the full values and
// s with the stored words; on return s holds the words to store (the offsets on FP32
// handles of models with StateRef, else the full values).
template <class M, class T, int INTEG>
CG_DEV T tissue_advance(T v, T (&y)[M::NS], T (&s)[M::NS], T dt, const T* p, int tissue,
                        const T* gs, const OdeArgs& o, int64_t c)
{
    if constexpr (INTEG == 0 && offsets_stored<M, T>()) {
        return Ops<M>::tissue_step_off(v, y, s, dt, p, tissue, gs);
    } else if constexpr (INTEG != 0 && sizeof(T) == 4) {
        // FP32 adaptive solvers integrate this step's increments d from 0 against the
        // step-start values (rates at base + d, the norm on those values), then add d to
        // each stored word once: a sub-step's increment below half an FP32 ulp of the state
        // would otherwise be rounded away (many short Euler sub-steps leave slow states
        // frozen), while d resolves it.  y itself is the step-start base (V set to v): it
        // is not written until the solve ends, so no second copy is kept live.
        T d[M::NS];
#pragma unroll
        for (int q = 0; q < M::NS; ++q) d[q] = T(0);
        y[0] = v;
        const auto qv = Ops<M>::vt(v, p);
        auto f = [&](double, const T* dd, T* r) __attribute__((always_inline)) {
            T yy[M::NS];
#pragma unroll
            for (int q = 0; q < M::NS; ++q) yy[q] = y[q] + dd[q];
            Ops<M>::deriv(yy, r, p, tissue, gs, false, T(0), qv);
        };
        double step = o.step[c];
        if (!(step < o.dt_clamp)) step = o.dt_clamp;
        const bool ok = ode_solve<INTEG, M::NS>(f, d, o.t0, o.t1, step, T(o.abs_tol),
                                                T(o.rel_tol), o.max_steps, y);
        o.step[c] = step;
        if (!ok) *o.fail = 1;
        s[0] = v;
#pragma unroll
        for (int q = 1; q < M::NS; ++q) {
            s[q] = s[q] + d[q];
            y[q] = offsets_stored<M, T>() ? T(StateRef<M>::at(q)) + s[q] : s[q];
        }
        return Ops<M>::iion(y, p, tissue, gs);
    } else {
        const T I = tissue_advance_full<M, T, INTEG>(v, y, dt, p, tissue, gs, o, c);
#pragma unroll
        for (int q = 1; q < M::NS; ++q)
            s[q] = offsets_stored<M, T>() ? y[q] - T(StateRef<M>::at(q)) : y[q];
        return I;
    }
}

// One single-cell step (V inside the ODE) on the handle's storage: y the full values,
// s the stored words, both kept current (the offsets on FP32 handles of models with
// StateRef, where the increments land; else s mirrors y).
template <class M, class T>
CG_DEV void cell_step_stored(T (&y)[M::NS], T (&s)[M::NS], T dt, T istim, const T* p,
                             int tissue, const T* gs)
{
    if constexpr (offsets_stored<M, T>()) {
        Ops<M>::cell_step_off(y, s, dt, istim, p, tissue, gs);
    } else {
        Ops<M>::cell_step(y, dt, istim, p, tissue, gs);
#pragma unroll
        for (int q = 0; q < M::NS; ++q) s[q] = y[q];
    }
}

// ---------------------------------------------------------------- ionic only
template <class M, class T, int INTEG>
__global__ __launch_bounds__(MBS) void k_ionic_tissue(T* __restrict__ st, int64_t n,
                                                      const double* __restrict__ Vm,
                                                      double* __restrict__ Im, T dt,
                                                      KParams<T> P,
                                                      const T* __restrict__ gsc, int tissue,
                                                      OdeArgs ode)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T* p = P.p;  // model constants read from the kernel arguments
    T gs[4] = {T(1), T(1), T(1), T(1)};
    if (gsc) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gs[k] = gsc[k * n + i];
    }
    T y[M::NS], sw[M::NS];
    load_model_states<M>(y, sw, st, n, i, 1);
    const T v = T((Vm[i] * M::VMUL + M::VOFF) / M::VSCALE);
    const T I = tissue_advance<M, T, INTEG>(v, y, sw, dt, p, tissue, gs, ode, i);
    store_states(sw, st, n, i, 1);
    Im[i] = double(I) * M::ISCALE;
}

// n single cells with V integrated inside the ODE (singleCellSolver.C:179-229), n_steps
// steps of dt from model time t0; RKF45 evaluates the stimulus at every stage time.
template <class M, class T, int INTEG>
__global__ __launch_bounds__(MBS) void k_ionic_cells(T* __restrict__ st, int64_t n, double t0,
                                                     T dt, int n_steps, KParams<T> P,
                                                     const T* __restrict__ gsc, int tissue,
                                                     Stimulus stim, OdeArgs ode,
                                                     double* __restrict__ trace, int64_t tcell)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the traced cell stores its states and rates after every step
    const bool rec = trace != nullptr && i == tcell;
    const T* p = P.p;  // model constants read from the kernel arguments
    T gs[4] = {T(1), T(1), T(1), T(1)};
    if (gsc) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gs[k] = gsc[k * n + i];
    }
    T y[M::NS], sw[M::NS];
    load_model_states<M>(y, sw, st, n, i, 0);
    // trace row of the traced cell after a step ending at model time te: the states,
    // then their rates at te (the STATES and RATES columns of the single-cell trace)
    auto record = [&](int s, double te) __attribute__((always_inline)) {
        T d[M::NS];
        Ops<M>::deriv(y, d, p, tissue, gs, true, T(stimulus_current(te, stim)),
                      Ops<M>::vt(y[0], p));
        double* row = trace + (int64_t)s * 2 * M::NS;
#pragma unroll
        for (int k = 0; k < M::NS; ++k) {
            row[k] = double(y[k]);
            row[M::NS + k] = double(d[k]);
        }
    };
    if constexpr (INTEG == 0) {
        for (int s = 0; s < n_steps; ++s) {
            const double t = t0 + s * double(dt);
            cell_step_stored<M>(y, sw, dt, T(stimulus_current(t, stim)), p, tissue, gs);
            if (rec) record(s, t0 + (s + 1) * double(dt));
        }
        store_states(sw, st, n, i, 0);
        return;
    } else {
        // FP64: the solver integrates the full values y (the stored words are y).  FP32:
        // each step integrates its increments d from 0 against the step-start values
        // (rates at base + d, norm on those values) and adds d to the stored words once,
        // so sub-step increments below half an ulp of a state still count.
        // (FP32: y holds the step-start values, unchanged until that step's solve ends)
        constexpr bool f32 = sizeof(T) == 4;
        T d[M::NS];
        auto f = [&](double t, const T* ss, T* r) __attribute__((always_inline)) {
            T yy[M::NS];
#pragma unroll
            for (int q = 0; q < M::NS; ++q) yy[q] = f32 ? y[q] + ss[q] : ss[q];
            const auto q = Ops<M>::vt(yy[0], p);
            Ops<M>::deriv(yy, r, p, tissue, gs, true, T(stimulus_current(t, stim)), q);
        };
        double step = ode.step[i];
        bool ok = true;
        double tacc = ode.t_s;
        for (int s = 0; s < n_steps && ok; ++s) {
            const double tn = tacc + ode.dt_s;
            const double ts = tn - ode.dt_s;
            tacc = tn;
            // model time: ms (s for the manufactured model, tEnd = tStart + deltaT)
            constexpr bool secs = M::TSCALE == 1.0;
            const double x0 = secs ? ts : ode.ap ? ts * 1000.0 / 12.9 : ts * 1000.0;
            const double x1 = secs      ? ts + ode.dt_s
                              : ode.ap ? (ts + ode.dt_s) * 1000.0 / 12.9
                                       : (ts + ode.dt_s) * 1000.0;
            if (!(step < ode.dt_clamp)) step = ode.dt_clamp;
            if constexpr (f32) {
#pragma unroll
                for (int q = 0; q < M::NS; ++q) d[q] = T(0);
                ok = ode_solve<INTEG, M::NS>(f, d, x0, x1, step, T(ode.abs_tol),
                                             T(ode.rel_tol), ode.max_steps, y);
#pragma unroll
                for (int q = 0; q < M::NS; ++q) {
                    sw[q] = sw[q] + d[q];
                    y[q] = offsets_stored<M, T>() ? T(StateRef<M>::at(q)) + sw[q] : sw[q];
                }
            } else {
                ok = ode_solve<INTEG, M::NS>(f, y, x0, x1, step, T(ode.abs_tol),
                                             T(ode.rel_tol), ode.max_steps);
#pragma unroll
                for (int q = 0; q < M::NS; ++q) sw[q] = y[q];
            }
            if (rec) record(s, x1);
        }
        ode.step[i] = step;
        if (!ok) *ode.fail = 1;
    }
    store_states(sw, st, n, i, 0);
}

// ------------------------------------------------------- fused explicit step
// Box stimulus of myocardiumDomain.C:307-342: summed intensities, in box order, of the
// boxes whose closed extent contains the cell centre.  The membership mask (when the
// mesh's centres were given) is read only while some box is active.
CG_DEV double box_source(int i, int j, int k, int64_t c, const HexGeom& g, const ActiveBoxes& b)
{
    double src = 0.0;
    if (b.mask) {
        if (b.n == 0) return 0.0;
        const unsigned m = b.mask[c];
        for (int q = 0; q < b.n; ++q)
            if ((m >> b.bit[q]) & 1u) src += b.intensity[q];
        return src;
    }
    const double p0 = (i + 0.5) * g.hx, p1 = (j + 0.5) * g.hy, p2 = (k + 0.5) * g.hz;
    for (int q = 0; q < b.n; ++q) {
        if (p0 >= b.lo[q][0] && p0 <= b.hi[q][0] && p1 >= b.lo[q][1] && p1 <= b.hi[q][1] &&
            p2 >= b.lo[q][2] && p2 <= b.hi[q][2])
            src += b.intensity[q];
    }
    return src;
}

// Fused explicit monodomain step on the hex block (myocardiumDomain.C:487-529,
// monodomainSolver.C:103-118): box stimulus -> ionic step at V^n -> diagonal solve
//   V^{n+1} = (c0 V^n - c00 V^{n-1} + dt/chiCm (lap - chiCm Iion + src)) / c
// -> activation time where V first exceeds 1e-15 V (myocardiumDomain.C:345-374).
// Bytes per cell-step (fp64, BuenoOrovio): V 8 + gates 24 r/w + V^{n+1} 8 + Iion 8 +
// flag 1 (+ V^{n-1} 8 for backward, + off-diagonal sigma 24 with a tensor correction).
template <class M, class T, bool UNIFORM, int INTEG>
__global__ __launch_bounds__(MBS) void k_explicit_hex(
    HexGeom g, HexCoefs a, T* __restrict__ st, const double* __restrict__ V,
    const double* __restrict__ Vold, double* __restrict__ Vnew, double* __restrict__ Iion,
    uint8_t* __restrict__ flag, double* __restrict__ tact, ExplicitStep s, KParams<T> P,
    int tissue)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const IJK ix = ijk_of((uint32_t)c, g);
    const T* p = P.p;  // model constants read from the kernel arguments
    const T one[4] = {T(1), T(1), T(1), T(1)};
    T y[M::NS], sw[M::NS];
    load_model_states<M>(y, sw, st, n, c, 1);
    const double vn = V[c];
    const T v = T((vn * M::VMUL + M::VOFF) / M::VSCALE);
    const T I = tissue_advance<M, T, INTEG>(v, y, sw, T(s.dt_model), p, tissue, one, s.ode, c);
    store_states(sw, st, n, c, 1);
    const double iion = double(I) * M::ISCALE;
    if (Iion) Iion[c] = iion;

    double lap = hex_laplacian<UNIFORM>(V, c, ix, g, a);
    if (s.od) lap += hex_tensor_correction(V, s.od, n, c, ix.i, ix.j, ix.k, g, s.od_mask);
    const double src = box_source(ix.i, ix.j, ix.k, c, g, s.boxes);
    double rhs = s.c0 * vn + s.dt / s.chiCm * (lap - s.chiCm * iion + src);
    if (s.c00 != 0.0) rhs -= s.c00 * Vold[c];
    const double vnew = rhs / s.c;
    Vnew[c] = vnew;
    if (flag[c] && vnew > s.act_v + 1e-15) {
        flag[c] = 0;
        tact[c] = s.t_old + (s.act_v - vn) / (vnew - vn) * s.dt;
    }
}

// ------------------------------------------ ionic step + implicit right-hand side
// rhs = chiCm/dt (c0 V^n - c00 V^{n-1}) - chiCm Iion + source (the explicit tensor
// correction is added by the solver's first residual); x0 (when not null) receives the
// solver's extrapolated initial guess 2 V^n - V^{n-1}, riding on this kernel's pass over
// V^n; with s.act_flag it also completes the previous step's activation times.  With a
// gate (s.gate) the launch may be queued before the previous step's solve is known to have
// converged: it then does nothing unless the solver recorded the expected generation.
// waves per SIMD the register allocation of k_ionic_rhs aims at (1: the compiler's
// choice).  FP64 ORd: 2, at 256 VGPRs with 296 B/lane of spills instead of 256 + 84 AGPRs
// at one wave; measured 4.6 % faster per implicit 2.1M-cell FP64 step (0.938 vs 0.983 ms).
template <class M, class T> constexpr int rhs_waves()
{
    if constexpr (std::is_same_v<M, ORd> && sizeof(T) == 8) return 2;
    return 1;
}

template <class M, class T, int INTEG>
__global__ __launch_bounds__(MBS) __attribute__((amdgpu_waves_per_eu(rhs_waves<M, T>())))
void k_ionic_rhs(HexGeom g, T* __restrict__ st,
                                                   const double* __restrict__ Vn,
                                                   const double* __restrict__ Vm1,
                                                   double* __restrict__ x0,
                                                   double* __restrict__ Iion,
                                                   double* __restrict__ b, ExplicitStep s,
                                                   KParams<T> P, int tissue)
{
    if (s.gate && *s.gate != s.gate_val) return;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * MBS + threadIdx.x;
    if (c >= n) return;
    const IJK cc = ijk_of((uint32_t)c, g);
    const T* p = P.p;  // model constants read from the kernel arguments
    const T one[4] = {T(1), T(1), T(1), T(1)};
    T y[M::NS], sw[M::NS];
    load_model_states<M>(y, sw, st, n, c, 1);
    const double vn = Vn[c];
    // loads of the epilogue issued before the model's arithmetic, so their latency hides
    // behind it
    const double vo = (x0 || s.c00 != 0.0 || s.act_flag) ? Vm1[c] : 0.0;
    const uint8_t fl = s.act_flag ? s.act_flag[c] : 0;
    const T v = T((vn * M::VMUL + M::VOFF) / M::VSCALE);
    const T I = tissue_advance<M, T, INTEG>(v, y, sw, T(s.dt_model), p, tissue, one, s.ode, c);
    store_states(sw, st, n, c, 1);
    const double iion = double(I) * M::ISCALE;
    Iion[c] = iion;
    const double src = box_source(cc.i, cc.j, cc.k, c, g, s.boxes);
    double rhs = s.chiCm / s.dt * (s.c0 * vn - s.c00 * vo) + src - s.chiCm * iion;
    b[c] = rhs;
    if (x0) x0[c] = 2.0 * vn - vo;
    // the previous step's activation: V^n is its new potential, V^{n-1} its old one
    if (fl && vn > s.act_v + 1e-15) {
        s.act_flag[c] = 0;
        s.act_tact[c] = s.act_t_old + (s.act_v - vo) / (vn - vo) * s.act_dt;
    }
}

// RATES (kind 0) or ALGEBRAIC (kind 1) variables of model M at n given rows: states
// [NS][n] (V as state 0, model units), t [n] model time of each row (the single-cell
// stimulus when solve_v), conductance scales [4][n] or null; out [NV][n].  FP64.
template <class M>
__global__ __launch_bounds__(MBS) void k_ionic_eval(int tissue, int64_t n,
                                                    const double* __restrict__ st,
                                                    const double* __restrict__ t, int kind,
                                                    int solve_v, Stimulus stim,
                                                    KParams<double> P,
                                                    const double* __restrict__ gsc,
                                                    double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* p = P.p;
    double gs[4] = {1.0, 1.0, 1.0, 1.0};
    if (gsc)
        for (int k = 0; k < 4; ++k) gs[k] = gsc[k * n + i];
    double y[M::NS];
    for (int q = 0; q < M::NS; ++q) y[q] = st[q * n + i];
    const double istim = solve_v ? stimulus_current(t[i], stim) : 0.0;
    if (kind == 0) {
        double d[M::NS];
        Ops<M>::deriv(y, d, p, tissue, gs, solve_v != 0, istim, Ops<M>::vt(y[0], p));
        for (int q = 0; q < M::NS; ++q) out[q * n + i] = d[q];
    } else if constexpr (Algebraic<M>::NA > 0) {
        double a[Algebraic<M>::NA];
        Algebraic<M>::eval(y, p, istim, a);
        for (int q = 0; q < Algebraic<M>::NA; ++q) out[q * n + i] = a[q];
    }
}

// ---------------------------------------------------------------- launchers
inline unsigned grid_8(int64_t n, int bs)
{
    int64_t nb = (n + bs - 1) / bs;
    return (unsigned)((nb + 7) / 8 * 8);  // multiple of 8 for the XCD remap
}

template <class T> struct ModelLaunch {
    template <class M>
    static hipError_t tissue(int tissue, void* states, int64_t n, const double* Vm, double* Im,
                             double dt_model, const ModelParams& P, const void* gscale,
                             const OdeArgs& ode, hipStream_t s)
    {
        const unsigned grid = (unsigned)((n + MBS - 1) / MBS);
#define CG_IONIC_TISSUE(I)                                                                 \
    hipLaunchKernelGGL((k_ionic_tissue<M, T, I>), dim3(grid), dim3(MBS), 0, s, (T*)states, \
                       n, Vm, Im, T(dt_model), kparams<T>(P), (const T*)gscale, tissue, ode)
        if (ode.method == 2) CG_IONIC_TISSUE(2);
        else if (ode.method == 1) CG_IONIC_TISSUE(1);
        else CG_IONIC_TISSUE(0);
#undef CG_IONIC_TISSUE
        return hipGetLastError();
    }
    template <class M>
    static hipError_t cells(int tissue, void* states, int64_t n, double t0, double dt,
                            int n_steps, const ModelParams& P, const void* gscale,
                            const Stimulus& stim, const OdeArgs& ode, double* trace,
                            int64_t tcell, hipStream_t s)
    {
        const unsigned grid = (unsigned)((n + MBS - 1) / MBS);
#define CG_IONIC_CELLS(I)                                                                  \
    hipLaunchKernelGGL((k_ionic_cells<M, T, I>), dim3(grid), dim3(MBS), 0, s, (T*)states,  \
                       n, t0, T(dt), n_steps, kparams<T>(P), (const T*)gscale, tissue, stim, \
                       ode, trace, tcell)
        if (ode.method == 2) CG_IONIC_CELLS(2);
        else if (ode.method == 1) CG_IONIC_CELLS(1);
        else CG_IONIC_CELLS(0);
#undef CG_IONIC_CELLS
        return hipGetLastError();
    }
    template <class M>
    static hipError_t explicit_hex(int tissue, const HexGeom& g, const HexCoefs& a,
                                   void* states, const double* V, const double* Vold,
                                   double* Vnew, double* Iion, uint8_t* flag, double* tact,
                                   const ExplicitStep& st, const ModelParams& P, hipStream_t s)
    {
        const int64_t n = (int64_t)g.nx * g.ny * g.nz;
        const unsigned grid = (unsigned)((n + MBS - 1) / MBS);
#define CG_EXPLICIT_HEX(U, I)                                                              \
    hipLaunchKernelGGL((k_explicit_hex<M, T, U, I>), dim3(grid), dim3(MBS), 0, s, g, a,    \
                       (T*)states, V, Vold, Vnew, Iion, flag, tact, st, kparams<T>(P), tissue)
        if (st.ode.method == 2) {
            if (a.fx) CG_EXPLICIT_HEX(false, 2);
            else CG_EXPLICIT_HEX(true, 2);
        } else if (st.ode.method == 1) {
            if (a.fx) CG_EXPLICIT_HEX(false, 1);
            else CG_EXPLICIT_HEX(true, 1);
        } else {
            if (a.fx) CG_EXPLICIT_HEX(false, 0);
            else CG_EXPLICIT_HEX(true, 0);
        }
#undef CG_EXPLICIT_HEX
        return hipGetLastError();
    }
    template <class M>
    static hipError_t ionic_rhs(int tissue, const HexGeom& g, void* states, const double* Vn,
                                const double* Vm1, double* x0, double* Iion,
                                double* b, const ExplicitStep& st, const ModelParams& P,
                                hipStream_t s)
    {
        const int64_t n = (int64_t)g.nx * g.ny * g.nz;
#define CG_IONIC_RHS(I)                                                                    \
    hipLaunchKernelGGL((k_ionic_rhs<M, T, I>), dim3(grid_8(n, MBS)), dim3(MBS), 0, s, g,    \
                       (T*)states, Vn, Vm1, x0, Iion, b, st, kparams<T>(P), tissue)
        if (st.ode.method == 2) CG_IONIC_RHS(2);
        else if (st.ode.method == 1) CG_IONIC_RHS(1);
        else CG_IONIC_RHS(0);
#undef CG_IONIC_RHS
        return hipGetLastError();
    }
};

// switch over the models for one arithmetic type
#define CG_MODEL_SWITCH(T, FN, ...)                                                        \
    switch (model) {                                                                       \
    case MODEL_BO: return ModelLaunch<T>::template FN<BuenoOrovio>(__VA_ARGS__);           \
    case MODEL_TNNP: return ModelLaunch<T>::template FN<TNNP>(__VA_ARGS__);                \
    case MODEL_ORD: return ModelLaunch<T>::template FN<ORd>(__VA_ARGS__);                  \
    case MODEL_AP: return ModelLaunch<T>::template FN<AlievPanfilov>(__VA_ARGS__);         \
    case MODEL_FDA: return ModelLaunch<T>::template FN<FdaManufactured>(__VA_ARGS__);      \
    case MODEL_CRN: return ModelLaunch<T>::template FN<Courtemanche>(__VA_ARGS__);         \
    case MODEL_PYS: return ModelLaunch<T>::template FN<PerisYague>(__VA_ARGS__);           \
    case MODEL_STW: return ModelLaunch<T>::template FN<Stewart>(__VA_ARGS__);              \
    case MODEL_GRD: return ModelLaunch<T>::template FN<Grandi>(__VA_ARGS__);               \
    case MODEL_TOR: return ModelLaunch<T>::template FN<ToRORd>(__VA_ARGS__);               \
    case MODEL_GAUR: return ModelLaunch<T>::template FN<Gaur>(__VA_ARGS__);                \
    default: return hipErrorInvalidValue;                                                  \
    }

// defines the four per-precision entry points of cg_internal.hpp for type T, suffix SFX
#define CG_DEFINE_MODEL_ENTRIES(T, SFX)                                                    \
    hipError_t launch_ionic_tissue_##SFX(int model, int tissue, void* states, int64_t n,   \
                                         const double* Vm, double* Im, double dt_model,    \
                                         const ModelParams& P, const void* gscale,         \
                                         const OdeArgs& ode, hipStream_t s)                \
    {                                                                                      \
        CG_MODEL_SWITCH(T, tissue, tissue, states, n, Vm, Im, dt_model, P, gscale, ode, s) \
    }                                                                                      \
    hipError_t launch_ionic_cells_##SFX(int model, int tissue, void* states, int64_t n,    \
                                        double t0, double dt, int n_steps,                 \
                                        const ModelParams& P, const void* gscale,          \
                                        const Stimulus& stim, const OdeArgs& ode,          \
                                        double* trace, int64_t tcell, hipStream_t s)       \
    {                                                                                      \
        CG_MODEL_SWITCH(T, cells, tissue, states, n, t0, dt, n_steps, P, gscale, stim,     \
                        ode, trace, tcell, s)                                              \
    }                                                                                      \
    hipError_t launch_explicit_hex_##SFX(int model, int tissue, const HexGeom& g,          \
                                         const HexCoefs& a, void* states, const double* V, \
                                         const double* Vold, double* Vnew, double* Iion,   \
                                         uint8_t* flag, double* tact,                      \
                                         const ExplicitStep& st, const ModelParams& P,     \
                                         hipStream_t s)                                    \
    {                                                                                      \
        CG_MODEL_SWITCH(T, explicit_hex, tissue, g, a, states, V, Vold, Vnew, Iion, flag,  \
                        tact, st, P, s)                                                    \
    }                                                                                      \
    hipError_t launch_ionic_rhs_##SFX(int model, int tissue, const HexGeom& g,             \
                                      void* states, const double* Vn, const double* Vm1,   \
                                      double* x0, double* Iion, double* b,                 \
                                      const ExplicitStep& st, const ModelParams& P,        \
                                      hipStream_t s)                                       \
    {                                                                                      \
        CG_MODEL_SWITCH(T, ionic_rhs, tissue, g, states, Vn, Vm1, x0, Iion, b, st, P, s)   \
    }

} // namespace cg
