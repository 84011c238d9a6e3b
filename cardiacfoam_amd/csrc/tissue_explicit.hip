// Fused explicit monodomain step on a uniform hex block (blockMesh cell order,
// c = i + nx*(j + ny*k)):  box stimulus -> ionic step -> explicit diffusion ->
// activation time, one cell per lane, one launch per time step.
//
// Reference operations (in order, myocardiumDomain.C:487-529, monodomainSolver.C:103-118):
//   source  = sum of active box intensities            (myocardiumDomain.C:307-342)
//   Iion    = ionic solveODE at V^n                     (ionicModel.H:250-256)
//   chiCm ddt(V) = laplacian(sigma, V^n) - chiCm Iion + source, solved on the
//   diagonal:  V^{n+1} = (c0 V^n - c00 V^{n-1} + dt/chiCm (lap - chiCm Iion + src)) / c
//   activation time where V first exceeds 1e-15 V      (myocardiumDomain.C:345-374)
//
// Bytes per cell-step (fp64, BuenoOrovio): read V 8 + gates 24 + write gates 24 +
// write V 8 + flag 1 (+ V^{n-1} 8 for backward) ; neighbour V values come from L2.
#include "cg_internal.hpp"

namespace cg {

template <bool UNIFORM>
CG_DEV double hex_laplacian(const double* __restrict__ V, int64_t c, int i, int j, int k,
                            const HexGeom& g, const HexCoefs& a)
{
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const double vc = V[c];
    double acc = 0.0;
    // x faces
    if (i < g.nx - 1) acc += (UNIFORM ? a.ax : a.fx[c]) * (V[c + 1] - vc);
    if (i > 0) acc += (UNIFORM ? a.ax : a.fx[c - 1]) * (V[c - 1] - vc);
    if (j < g.ny - 1) acc += (UNIFORM ? a.ay : a.fy[c]) * (V[c + sy] - vc);
    if (j > 0) acc += (UNIFORM ? a.ay : a.fy[c - sy]) * (V[c - sy] - vc);
    if (k < g.nz - 1) acc += (UNIFORM ? a.az : a.fz[c]) * (V[c + sz] - vc);
    if (k > 0) acc += (UNIFORM ? a.az : a.fz[c - sz]) * (V[c - sz] - vc);
    return acc;
}

CG_DEV double box_source(int i, int j, int k, const HexGeom& g, const ActiveBoxes& b)
{
    double src = 0.0;
    const double p0 = (i + 0.5) * g.hx, p1 = (j + 0.5) * g.hy, p2 = (k + 0.5) * g.hz;
    for (int q = 0; q < b.n; ++q) {
        if (p0 >= b.lo[q][0] && p0 <= b.hi[q][0] && p1 >= b.lo[q][1] && p1 <= b.hi[q][1] &&
            p2 >= b.lo[q][2] && p2 <= b.hi[q][2])
            src += b.intensity[q];
    }
    return src;
}

template <class M, class T, bool UNIFORM>
__global__ __launch_bounds__(256) void k_explicit_hex(
    HexGeom g, HexCoefs a, T* __restrict__ st, const double* __restrict__ V,
    const double* __restrict__ Vold, double* __restrict__ Vnew, double* __restrict__ Iion,
    uint8_t* __restrict__ flag, double* __restrict__ tact, ExplicitStep s, ModelParams P,
    int tissue)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % g.nx);
    const int j = (int)((c / g.nx) % g.ny);
    const int k = (int)(c / ((int64_t)g.nx * g.ny));

    // ionic step with V^n
    T p[M::NP];
#pragma unroll
    for (int q = 0; q < M::NP; ++q) p[q] = T(P.p[q]);
    const T one[4] = {T(1), T(1), T(1), T(1)};
    T y[M::NS];
#pragma unroll
    for (int q = 1; q < M::NS; ++q) y[q] = st[q * n + c];
    const double vn = V[c];
    const T v = T((vn * 1000.0 + M::VOFF) / M::VSCALE);
    const T I = Ops<M>::tissue_step(v, y, T(s.dt_model), p, tissue, one);
#pragma unroll
    for (int q = 1; q < M::NS; ++q) st[q * n + c] = y[q];
    const double iion = double(I) * M::ISCALE;
    if (Iion) Iion[c] = iion;

    // explicit diffusion
    double lap = hex_laplacian<UNIFORM>(V, c, i, j, k, g, a);
    if (s.corr) lap += s.corr[c];
    const double src = box_source(i, j, k, g, s.boxes);
    double rhs = s.c0 * vn + s.dt / s.chiCm * (lap - s.chiCm * iion + src);
    if (s.c00 != 0.0) rhs -= s.c00 * Vold[c];
    const double vnew = rhs / s.c;
    Vnew[c] = vnew;

    // activation time, linear interpolation of the 0 V crossing inside the step
    if (flag[c] && vnew > 1e-15) {
        flag[c] = 0;
        tact[c] = s.t_old + (0.0 - vn) / (vnew - vn) * s.dt;
    }
}

template <class M, class T>
static hipError_t explicit_t(int tissue, const HexGeom& g, const HexCoefs& a, void* states,
                             const double* V, const double* Vold, double* Vnew, double* Iion,
                             uint8_t* flag, double* tact, const ExplicitStep& s,
                             const ModelParams& P, hipStream_t st)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int bs = 256;
    const unsigned grid = (unsigned)((n + bs - 1) / bs);
    if (a.fx)
        hipLaunchKernelGGL((k_explicit_hex<M, T, false>), dim3(grid), dim3(bs), 0, st, g, a,
                           (T*)states, V, Vold, Vnew, Iion, flag, tact, s, P, tissue);
    else
        hipLaunchKernelGGL((k_explicit_hex<M, T, true>), dim3(grid), dim3(bs), 0, st, g, a,
                           (T*)states, V, Vold, Vnew, Iion, flag, tact, s, P, tissue);
    return hipGetLastError();
}

hipError_t launch_explicit_hex(int model, int prec, int tissue, const HexGeom& g,
                               const HexCoefs& k, void* states, const double* V,
                               const double* Vold, double* Vnew, double* Iion, uint8_t* flag,
                               double* tact, const ExplicitStep& st, const ModelParams& P,
                               hipStream_t s)
{
    switch (model * 2 + prec) {
    case 0: return explicit_t<BuenoOrovio, double>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 1: return explicit_t<BuenoOrovio, float>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 2: return explicit_t<TNNP, double>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 3: return explicit_t<TNNP, float>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 4: return explicit_t<ORd, double>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 5: return explicit_t<ORd, float>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 6: return explicit_t<AlievPanfilov, double>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    case 7: return explicit_t<AlievPanfilov, float>(tissue, g, k, states, V, Vold, Vnew, Iion, flag, tact, st, P, s);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------- tensor correction
// grad V (Gauss linear, zeroGradient walls: wall face value = cell value)
__global__ __launch_bounds__(256) void k_hex_grad(HexGeom g, const double* __restrict__ V,
                                                  double* __restrict__ grad)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((int64_t)g.nx * g.ny));
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const double v = V[c];
    const double xp = i < g.nx - 1 ? 0.5 * (v + V[c + 1]) : v;
    const double xm = i > 0 ? 0.5 * (v + V[c - 1]) : v;
    const double yp = j < g.ny - 1 ? 0.5 * (v + V[c + sy]) : v;
    const double ym = j > 0 ? 0.5 * (v + V[c - sy]) : v;
    const double zp = k < g.nz - 1 ? 0.5 * (v + V[c + sz]) : v;
    const double zm = k > 0 ? 0.5 * (v + V[c - sz]) : v;
    grad[3 * c + 0] = (xp - xm) / g.hx;
    grad[3 * c + 1] = (yp - ym) / g.hy;
    grad[3 * c + 2] = (zp - zm) / g.hz;
}

// div of SfGammaCorr . interp(grad V) per unit volume; wall faces use the cell tensor
// and the cell gradient with its normal component set to snGrad = 0.
__global__ __launch_bounds__(256) void k_hex_corr(HexGeom g, const double* __restrict__ sig,
                                                  const double* __restrict__ grad,
                                                  double* __restrict__ out)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int idx[3] = {(int)(c % g.nx), (int)((c / g.nx) % g.ny), (int)(c / ((int64_t)g.nx * g.ny))};
    const int nn[3] = {g.nx, g.ny, g.nz};
    const int64_t st[3] = {1, g.nx, (int64_t)g.nx * g.ny};
    const double h[3] = {g.hx, g.hy, g.hz};
    double acc = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double A_over_V = 1.0 / h[d];   // face area / cell volume
        for (int side = 0; side < 2; ++side) {
            const bool interior = side == 0 ? idx[d] < nn[d] - 1 : idx[d] > 0;
            const double sgn = side == 0 ? 1.0 : -1.0;
            double q = 0.0;
            if (interior) {
                const int64_t nb = side == 0 ? c + st[d] : c - st[d];
                for (int e = 0; e < 3; ++e) {
                    if (e == d) continue;
                    const double sf = 0.5 * (sig[9 * c + 3 * d + e] + sig[9 * nb + 3 * d + e]);
                    const double gf = 0.5 * (grad[3 * c + e] + grad[3 * nb + e]);
                    q += sf * gf;
                }
            } else {
                for (int e = 0; e < 3; ++e) {
                    if (e == d) continue;
                    q += sig[9 * c + 3 * d + e] * grad[3 * c + e];
                }
            }
            acc += sgn * A_over_V * q;
        }
    }
    out[c] = acc;
}

hipError_t launch_hex_correction(const HexGeom& g, const double* sigma9, const double* V,
                                 double* grad3, double* out, hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int bs = 256;
    const unsigned grid = (unsigned)((n + bs - 1) / bs);
    hipLaunchKernelGGL(k_hex_grad, dim3(grid), dim3(bs), 0, s, g, V, grad3);
    hipLaunchKernelGGL(k_hex_corr, dim3(grid), dim3(bs), 0, s, g, sigma9, grad3, out);
    return hipGetLastError();
}

} // namespace cg
