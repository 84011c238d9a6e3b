// Ionic step kernels: one cell per lane, states structure-of-arrays (var*n + cell) so
// that each state load/store of a wavefront is one contiguous 256 B (fp32) / 512 B
// (fp64) segment; all gates of a cell stay in registers for the whole step.
#include "cg_internal.hpp"

namespace cg {

template <class M, class T>
__global__ __launch_bounds__(256) void k_ionic_tissue(T* __restrict__ st, int64_t n,
                                                      const double* __restrict__ Vm,
                                                      double* __restrict__ Im, T dt,
                                                      ModelParams P,
                                                      const T* __restrict__ gsc, int tissue)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T p[M::NP];
#pragma unroll
    for (int k = 0; k < M::NP; ++k) p[k] = T(P.p[k]);
    T gs[4] = {T(1), T(1), T(1), T(1)};
    if (gsc) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gs[k] = gsc[k * n + i];
    }
    T y[M::NS];
#pragma unroll
    for (int k = 1; k < M::NS; ++k) y[k] = st[k * n + i];
    const T v = T((Vm[i] * 1000.0 + M::VOFF) / M::VSCALE);
    const T I = Ops<M>::tissue_step(v, y, dt, p, tissue, gs);
#pragma unroll
    for (int k = 1; k < M::NS; ++k) st[k * n + i] = y[k];
    Im[i] = double(I) * M::ISCALE;
}

template <class M, class T>
__global__ __launch_bounds__(256) void k_ionic_cells(T* __restrict__ st, int64_t n, double t0,
                                                     T dt, int n_steps, ModelParams P,
                                                     const T* __restrict__ gsc, int tissue,
                                                     Stimulus stim)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T p[M::NP];
#pragma unroll
    for (int k = 0; k < M::NP; ++k) p[k] = T(P.p[k]);
    T gs[4] = {T(1), T(1), T(1), T(1)};
    if (gsc) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gs[k] = gsc[k * n + i];
    }
    T y[M::NS];
#pragma unroll
    for (int k = 0; k < M::NS; ++k) y[k] = st[k * n + i];
    for (int s = 0; s < n_steps; ++s) {
        const double t = t0 + s * double(dt);
        Ops<M>::cell_step(y, dt, T(stimulus_current(t, stim)), p, tissue, gs);
    }
#pragma unroll
    for (int k = 0; k < M::NS; ++k) st[k * n + i] = y[k];
}

template <class M, class T>
static hipError_t ionic_tissue_t(int tissue, void* states, int64_t n, const double* Vm,
                                 double* Im, double dt_model, const ModelParams& P,
                                 const void* gscale, hipStream_t s)
{
    const int bs = 256;
    const unsigned grid = (unsigned)((n + bs - 1) / bs);
    hipLaunchKernelGGL((k_ionic_tissue<M, T>), dim3(grid), dim3(bs), 0, s, (T*)states, n, Vm,
                       Im, T(dt_model), P, (const T*)gscale, tissue);
    return hipGetLastError();
}

template <class M, class T>
static hipError_t ionic_cells_t(int tissue, void* states, int64_t n, double t0, double dt,
                                int n_steps, const ModelParams& P, const void* gscale,
                                const Stimulus& stim, hipStream_t s)
{
    const int bs = 256;
    const unsigned grid = (unsigned)((n + bs - 1) / bs);
    hipLaunchKernelGGL((k_ionic_cells<M, T>), dim3(grid), dim3(bs), 0, s, (T*)states, n, t0,
                       T(dt), n_steps, P, (const T*)gscale, tissue, stim);
    return hipGetLastError();
}

#define CG_DISPATCH(model, prec, FN, ...)                                                  \
    switch ((model) * 2 + (prec)) {                                                        \
    case 0: return FN<BuenoOrovio, double>(__VA_ARGS__);                                   \
    case 1: return FN<BuenoOrovio, float>(__VA_ARGS__);                                    \
    case 2: return FN<TNNP, double>(__VA_ARGS__);                                          \
    case 3: return FN<TNNP, float>(__VA_ARGS__);                                           \
    case 4: return FN<ORd, double>(__VA_ARGS__);                                           \
    case 5: return FN<ORd, float>(__VA_ARGS__);                                            \
    case 6: return FN<AlievPanfilov, double>(__VA_ARGS__);                                 \
    case 7: return FN<AlievPanfilov, float>(__VA_ARGS__);                                  \
    default: return hipErrorInvalidValue;                                                  \
    }

hipError_t launch_ionic_tissue(int model, int prec, int tissue, void* states, int64_t n,
                               const double* Vm, double* Im, double dt_model,
                               const ModelParams& P, const void* gscale, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    CG_DISPATCH(model, prec, ionic_tissue_t, tissue, states, n, Vm, Im, dt_model, P, gscale, s)
}

hipError_t launch_ionic_cells(int model, int prec, int tissue, void* states, int64_t n,
                              double t0_model, double dt_model, int n_steps,
                              const ModelParams& P, const void* gscale, const Stimulus& stim,
                              hipStream_t s)
{
    if (n <= 0 || n_steps <= 0) return hipSuccess;
    CG_DISPATCH(model, prec, ionic_cells_t, tissue, states, n, t0_model, dt_model, n_steps, P,
                gscale, stim, s)
}

} // namespace cg
