// Pseudo-ECG of the monodomain tissue on the hex block (PseudoECGSolver::compute,
// src/electroModels/ecgModels/pseudoECGSolver/pseudoECGSolver.C:36-83, the Gima-Rudy
// dipole sum):
//   phi(P_e) = sum_c ((sigma_c . grad V_c) vol_c) . r / |r|^3,   r = C_c - P_e,
// skipping cells with |r| <= VSMALL.  grad V is fvc::grad with Gauss linear on the
// uniform block: (V_{i+1} - V_{i-1}) / 2h, the missing neighbour of a wall cell replaced by
// the cell itself (zeroGradient walls give the face the cell value).
//
// One pass reads V (8 B per cell, neighbours through the caches) and, for a per-cell
// conductivity, its 6 (symmetric) or 9 components (48 or 72 B); the arithmetic is about 15
// FP64 instructions per cell and electrode (a reciprocal square root among them).
// Each lane accumulates ECG_EG electrodes of a grid-stride range of cells; per-block sums
// (wave shuffles, then the waves in order) go to partial[e][block] and a second kernel
// adds those in a fixed order, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cg_internal.hpp"
#include "cg_ecg.hpp"

// Scratch of the allocation-free entry: each GPU holds its own copy until the process
// ends.  Host code never names it in a launch or copy: it uses the address that
// hipGetSymbolAddress returns for the current device (the name in host code is only a
// host-side stand-in).
static __device__ __align__(256) double scratch[cg::ECG_SCRATCH_DOUBLES];

namespace cg {

hipError_t ecg_scratch(double** out)
{
    constexpr int MAX_DEV = 64;
    static std::mutex mu;
    static double* ptr[MAX_DEV] = {};
    int dev = -1;
    if (hipError_t e = hipGetDevice(&dev)) return e;
    if (dev < 0 || dev >= MAX_DEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) return e;
        ptr[dev] = static_cast<double*>(p);
    }
    *out = ptr[dev];
    return hipSuccess;
}

namespace {

constexpr int ECG_BS = 256;   // threads per block
constexpr int ECG_EG = 16;    // electrodes per pass (accumulators per lane)

struct EcgArgs {
    HexGeom g;
    const double* V;
    const double* sig;      // per cell [6][n] (xx xy xz yy yz zz) or [9][n], or null
    int sym;                // sig holds 6 symmetric components
    double s9[9];
    double ihx2, ihy2, ihz2;  // 1 / (2 h) per direction
    const double* pos;      // [E][3]
    int e0, ne;             // electrodes e0 .. e0+ne-1 in this pass (ne <= ECG_EG)
    double* partial;        // [E][gridDim.x]
};

__global__ __launch_bounds__(ECG_BS) void k_pseudo_ecg(EcgArgs a)
{
    const HexGeom g = a.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const double vol = g.hx * g.hy * g.hz;
    double px[ECG_EG], py[ECG_EG], pz[ECG_EG], acc[ECG_EG];
#pragma unroll
    for (int q = 0; q < ECG_EG; ++q) {
        const int e = a.e0 + (q < a.ne ? q : 0);
        px[q] = a.pos[3 * e];
        py[q] = a.pos[3 * e + 1];
        pz[q] = a.pos[3 * e + 2];
        acc[q] = 0.0;
    }
    // grid-stride walk over the cells with (i, j, k) advanced by the stride's own
    // decomposition (carries instead of a division per cell)
    const int64_t stride = (int64_t)gridDim.x * ECG_BS;
    const int64_t c0 = (int64_t)blockIdx.x * ECG_BS + threadIdx.x;
    const int di = (int)(stride % g.nx), dj = (int)((stride / g.nx) % g.ny);
    const int64_t dk = stride / sz;
    int i = (int)(c0 % g.nx), j = (int)((c0 / g.nx) % g.ny);
    int64_t k = c0 / sz;
    for (int64_t c = c0; c < n; c += stride) {
        const double vc = a.V[c];
        const double vxp = i < g.nx - 1 ? a.V[c + 1] : vc, vxm = i > 0 ? a.V[c - 1] : vc;
        const double vyp = j < g.ny - 1 ? a.V[c + sy] : vc, vym = j > 0 ? a.V[c - sy] : vc;
        const double vzp = k < g.nz - 1 ? a.V[c + sz] : vc, vzm = k > 0 ? a.V[c - sz] : vc;
        const double gx = (vxp - vxm) * a.ihx2;
        const double gy = (vyp - vym) * a.ihy2;
        const double gz = (vzp - vzm) * a.ihz2;
        double s[9];
        if (a.sig && a.sym) {
            const double* p = a.sig + c;
            s[0] = p[0]; s[1] = p[n]; s[2] = p[2 * n];
            s[3] = s[1]; s[4] = p[3 * n]; s[5] = p[4 * n];
            s[6] = s[2]; s[7] = s[5]; s[8] = p[5 * n];
        } else if (a.sig) {
#pragma unroll
            for (int q = 0; q < 9; ++q) s[q] = a.sig[q * n + c];
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) s[q] = a.s9[q];
        }
        const double dx = (s[0] * gx + s[1] * gy + s[2] * gz) * vol;
        const double dy = (s[3] * gx + s[4] * gy + s[5] * gz) * vol;
        const double dz = (s[6] * gx + s[7] * gy + s[8] * gz) * vol;
        const double cx = (i + 0.5) * g.hx, cy = (j + 0.5) * g.hy, cz = ((double)k + 0.5) * g.hz;
#pragma unroll
        for (int q = 0; q < ECG_EG; ++q) {
            if (q < a.ne) {
                const double rx = cx - px[q], ry = cy - py[q], rz = cz - pz[q];
                const double r2 = rx * rx + ry * ry + rz * rz;
                // |r| > VSMALL (1e-300): r2 underflows to 0 below 1e-162, so for any
                // distance a mesh can hold the test is r2 > 0.  1/|r| from the hardware
                // reciprocal square root and one Newton step (relative error ~1e-16),
                // in place of a square root and a division
                if (r2 > 0.0) {
                    double y = __builtin_amdgcn_rsq(r2);
                    y = y * (1.5 - 0.5 * r2 * y * y);
                    acc[q] += (dx * rx + dy * ry + dz * rz) * (y * y * y);
                }
            }
        }
        i += di;
        int cj = 0;
        if (i >= g.nx) { i -= g.nx; cj = 1; }
        j += dj + cj;
        int64_t ck = 0;
        if (j >= g.ny) { j -= g.ny; ck = 1; }
        k += dk + ck;
    }
    // per-wave sums by butterfly shuffles (the same order on every lane), then the block's
    // four wave sums added in wave order
    constexpr int NW = ECG_BS / 64;
    __shared__ double red[NW][ECG_EG];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < ECG_EG; ++q) {
        if (q < a.ne) {
            double v = acc[q];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[wave][q] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < a.ne) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) v += red[w][threadIdx.x];
        a.partial[(int64_t)(a.e0 + threadIdx.x) * gridDim.x + blockIdx.x] = v;
    }
}

// out[e] = sum over blocks of partial[e][block], one block per electrode, fixed order
__global__ __launch_bounds__(ECG_BS) void k_ecg_sum(const double* __restrict__ partial,
                                                    int nblocks, double* __restrict__ out)
{
    __shared__ double red[ECG_BS];
    const double* p = partial + (int64_t)blockIdx.x * nblocks;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += ECG_BS) s += p[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = ECG_BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

} // namespace

int ecg_partial_blocks(int64_t n)
{
    const int64_t want = (n + ECG_BS - 1) / ECG_BS;
    return (int)(want < ECG_BLOCKS ? want : ECG_BLOCKS);
}

hipError_t launch_pseudo_ecg(const HexGeom& g, const double* V, const double* sig, int sym,
                             const double* s9, int n_el, const double* pos_dev,
                             double* partial, double* out_dev, hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = ecg_partial_blocks(n);
    EcgArgs a{};
    a.g = g;
    a.V = V;
    a.sig = sig;
    a.sym = sym;
    for (int q = 0; q < 9; ++q) a.s9[q] = s9 ? s9[q] : 0.0;
    a.ihx2 = 0.5 / g.hx;
    a.ihy2 = 0.5 / g.hy;
    a.ihz2 = 0.5 / g.hz;
    a.pos = pos_dev;
    a.partial = partial;
    for (int e0 = 0; e0 < n_el; e0 += ECG_EG) {
        a.e0 = e0;
        a.ne = n_el - e0 < ECG_EG ? n_el - e0 : ECG_EG;
        hipLaunchKernelGGL(k_pseudo_ecg, dim3(nb), dim3(ECG_BS), 0, s, a);
        if (hipError_t e = hipGetLastError()) return e;
    }
    hipLaunchKernelGGL(k_ecg_sum, dim3(n_el), dim3(ECG_BS), 0, s, partial, nb, out_dev);
    return hipGetLastError();
}

} // namespace cg
