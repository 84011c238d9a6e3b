// Monodomain diffusion on an arbitrary polyhedral mesh (monodomainSolver.C:103-150 with
// OpenFOAM's Gauss linear corrected laplacian of a tensor conductivity).
//
// Geometry, computed once on the host from what the mesh hands over:
//   w_f    = |Sf.(C_N - Cf)| / (|Sf.(Cf - C_P)| + |Sf.(C_N - Cf)|)   (linear weights)
//   n_f    = Sf/|Sf|,  d_f = C_N - C_P
//   D_f    = 1 / max(n_f.d_f, 0.05 |d_f|)                             (nonOrthDeltaCoeffs)
//   k_f    = n_f - D_f d_f                                            (nonOrthCorrectionVectors)
// and from the conductivity (sigma_f = w sigma_P + (1 - w) sigma_N):
//   SfGamma = Sf . sigma_f,  SfGammaSn = SfGamma . n_f,  SfGammaCorr = SfGamma - SfGammaSn n_f
//   a_f = SfGammaSn D_f,  b_f = SfGammaSn k_f + SfGammaCorr
// so that the face flux of fvc::laplacian (and of fvm::laplacian's matrix plus its explicit
// correction) is
//   flux_f = a_f (V_N - V_P) + b_f . (w grad V_P + (1 - w) grad V_N),
// with grad V the Gauss-linear gradient.  A zeroGradient boundary face carries only
// SfGammaCorr_b . (grad V_P - n (n . grad V_P)) (the gradient's boundary value with its
// normal part replaced by snGrad = 0; sigma_b = sigma_P; no correction vector on walls).
//
// Device layout: cell-major ELL over the faces of each cell, K slots, column-major
// ([K][n]), slot = face index (owner side) or ~face (neighbour side), INT32_MIN empty; the
// boundary faces in a second ELL of KB slots.  Every cell evaluates each of its faces'
// value or flux with the same expression and operand order as its neighbour does, so the
// two sides see bit-identical numbers and the divergence conserves exactly without a face
// pass or atomics.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <vector>
#include <hip/hip_runtime.h>

#include "cg_general.hpp"

namespace cg {

namespace {

constexpr int GBS = 256;
constexpr int32_t EMPTY = INT32_MIN;  // unused slot (~f of a face f >= 0 is at least -2^31+1)

struct EcgSigma {
    double v[9];
};

struct GView {
    int64_t n;
    int K, KB;
    const int32_t* slot;   // [K][n]
    const int32_t* nbc;    // [K][n] the other cell of the slot's face
    const int32_t* bslot;  // [KB][n] boundary face or -1
    const double* w;       // [F]
    const double* Sf;      // [3][F]
    const double* a;       // [F]
    const double* b;       // [3][F]
    const double* Sfb;     // [3][B]
    const double* bb;      // [3][B]  SfGammaCorr of the boundary face
    const double* nb;      // [3][B]  unit normal
    const double* vol;     // [n]
    const double* C;       // [3][n]
    int64_t F, B;
};

__global__ __launch_bounds__(GBS) void k_ggrad(GView g, const double* __restrict__ V,
                                               double* __restrict__ grad)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c >= g.n) return;
    const double vp = V[c];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < g.K; ++k) {
        const int32_t s = g.slot[(int64_t)k * g.n + c];
        if (s == EMPTY) continue;
        const bool own = s >= 0;
        const int64_t f = own ? s : ~s;
        const double vn = V[g.nbc[(int64_t)k * g.n + c]];
        const double wf = g.w[f];
        // face value w V_owner + (1 - w) V_neighbour, the same on both sides
        const double vf = own ? wf * vp + (1.0 - wf) * vn : wf * vn + (1.0 - wf) * vp;
        const double sg = own ? vf : -vf;
        sx += g.Sf[f] * sg;
        sy += g.Sf[g.F + f] * sg;
        sz += g.Sf[2 * g.F + f] * sg;
    }
    for (int k = 0; k < g.KB; ++k) {
        const int32_t bf = g.bslot[(int64_t)k * g.n + c];
        if (bf < 0) continue;
        sx += g.Sfb[bf] * vp;
        sy += g.Sfb[g.B + bf] * vp;
        sz += g.Sfb[2 * g.B + bf] * vp;
    }
    const double iv = 1.0 / g.vol[c];
    grad[c] = sx * iv;
    grad[g.n + c] = sy * iv;
    grad[2 * g.n + c] = sz * iv;
}

// sum over the cell's faces of the outward flux (volume-integrated)
__device__ __forceinline__ double flux_sum(const GView& g, const double* __restrict__ V,
                                           const double* __restrict__ grad, int64_t c,
                                           double vp, bool with_orth)
{
    const double gpx = grad[c], gpy = grad[g.n + c], gpz = grad[2 * g.n + c];
    double sum = 0.0;
    for (int k = 0; k < g.K; ++k) {
        const int32_t s = g.slot[(int64_t)k * g.n + c];
        if (s == EMPTY) continue;
        const bool own = s >= 0;
        const int64_t f = own ? s : ~s;
        const int64_t o = g.nbc[(int64_t)k * g.n + c];
        const double vo_ = V[o];
        const double gox = grad[o], goy = grad[g.n + o], goz = grad[2 * g.n + o];
        const double wf = g.w[f];
        // owner (P_) and neighbour (N_) of the face
        const double vP = own ? vp : vo_, vN = own ? vo_ : vp;
        const double gPx = own ? gpx : gox, gPy = own ? gpy : goy, gPz = own ? gpz : goz;
        const double gNx = own ? gox : gpx, gNy = own ? goy : gpy, gNz = own ? goz : gpz;
        const double fx = wf * gPx + (1.0 - wf) * gNx;
        const double fy = wf * gPy + (1.0 - wf) * gNy;
        const double fz = wf * gPz + (1.0 - wf) * gNz;
        double flux = g.b[f] * fx + g.b[g.F + f] * fy + g.b[2 * g.F + f] * fz;
        if (with_orth) flux += g.a[f] * (vN - vP);
        sum += own ? flux : -flux;
    }
    for (int k = 0; k < g.KB; ++k) {
        const int32_t bf = g.bslot[(int64_t)k * g.n + c];
        if (bf < 0) continue;
        const double nx = g.nb[bf], ny = g.nb[g.B + bf], nz = g.nb[2 * g.B + bf];
        const double gn = nx * gpx + ny * gpy + nz * gpz;
        sum += g.bb[bf] * (gpx - gn * nx) + g.bb[g.B + bf] * (gpy - gn * ny) +
               g.bb[2 * g.B + bf] * (gpz - gn * nz);
    }
    return sum;
}

// box stimulus on the mesh's centres; a hex block described as a general mesh uses the
// membership mask made from the caller's centres when it has one
__device__ __forceinline__ double general_box(const GView& g, int64_t c, const ActiveBoxes& b)
{
    double src = 0.0;
    if (b.mask) {
        if (b.n == 0) return 0.0;
        const unsigned m = b.mask[c];
        for (int q = 0; q < b.n; ++q)
            if ((m >> b.bit[q]) & 1u) src += b.intensity[q];
        return src;
    }
    const double p0 = g.C[c], p1 = g.C[g.n + c], p2 = g.C[2 * g.n + c];
    for (int q = 0; q < b.n; ++q) {
        if (p0 >= b.lo[q][0] && p0 <= b.hi[q][0] && p1 >= b.lo[q][1] && p1 <= b.hi[q][1] &&
            p2 >= b.lo[q][2] && p2 <= b.hi[q][2])
            src += b.intensity[q];
    }
    return src;
}

__global__ __launch_bounds__(GBS) void k_gexplicit(GView g, const double* __restrict__ V,
                                                   const double* __restrict__ grad,
                                                   const double* __restrict__ Vold,
                                                   double* __restrict__ Vnew,
                                                   const double* __restrict__ Iion,
                                                   uint8_t* __restrict__ flag,
                                                   double* __restrict__ tact, GeneralStep st)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c >= g.n) return;
    const double vn = V[c];
    const double lap = flux_sum(g, V, grad, c, vn, true) / g.vol[c];
    const double src = st.src ? st.src[c] : general_box(g, c, st.boxes);
    double rhs = st.c0 * vn + st.dt / st.chiCm * (lap - st.chiCm * Iion[c] + src);
    if (st.c00 != 0.0) rhs -= st.c00 * Vold[c];
    const double vnew = rhs / st.c;
    Vnew[c] = vnew;
    if (flag[c] && vnew > st.act_v + 1e-15) {
        flag[c] = 0;
        tact[c] = st.t_old + (st.act_v - vn) / (vnew - vn) * st.dt;
    }
}

__global__ __launch_bounds__(GBS) void k_glap(GView g, const double* __restrict__ V,
                                              const double* __restrict__ grad,
                                              double* __restrict__ lap)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c >= g.n) return;
    lap[c] = flux_sum(g, V, grad, c, V[c], true) / g.vol[c];
}

__global__ __launch_bounds__(GBS) void k_grhs0(GView g, const double* __restrict__ Vn,
                                               const double* __restrict__ Vm1,
                                               const double* __restrict__ Iion,
                                               double* __restrict__ b0, GeneralStep st)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c >= g.n) return;
    const double vo = st.c00 != 0.0 ? Vm1[c] : 0.0;
    const double src = st.src ? st.src[c] : general_box(g, c, st.boxes);
    b0[c] = g.vol[c] * (st.chiCm / st.dt * (st.c0 * Vn[c] - st.c00 * vo) -
                        st.chiCm * Iion[c] + src);
}

__global__ __launch_bounds__(GBS) void k_grhs_corr(GView g, const double* __restrict__ psi,
                                                   const double* __restrict__ grad,
                                                   const double* __restrict__ b0,
                                                   double* __restrict__ b)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c >= g.n) return;
    b[c] = b0[c] + flux_sum(g, psi, grad, c, psi[c], false);
}

__global__ __launch_bounds__(GBS) void k_gguess(int64_t n, const double* __restrict__ Vn,
                                                const double* __restrict__ Vm1,
                                                double* __restrict__ psi)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c < n) psi[c] = 2.0 * Vn[c] - Vm1[c];
}

__global__ __launch_bounds__(GBS) void k_gactivation(int64_t n, const double* __restrict__ Vn,
                                                     const double* __restrict__ Vnew,
                                                     uint8_t* __restrict__ flag,
                                                     double* __restrict__ tact, double t_old,
                                                     double dt, double vthr)
{
    const int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x;
    if (c >= n) return;
    const double vnew = Vnew[c];
    if (flag[c] && vnew > vthr + 1e-15) {
        const double vn = Vn[c];
        flag[c] = 0;
        tact[c] = t_old + (vthr - vn) / (vnew - vn) * dt;
    }
}

// pseudo-ECG partial sums on a general mesh (the dipole of pseudoECGSolver.C:36-83 from the
// Gauss gradient, the cell centres and volumes); one pass covers up to 16 electrodes
constexpr int GE_EG = 16;
__global__ __launch_bounds__(GBS) void k_gecg(GView g, const double* __restrict__ grad,
                                              const double* __restrict__ sig, int sym,
                                              EcgSigma s9, const double* __restrict__ pos,
                                              int e0, int ne, double* __restrict__ partial)
{
    double px[GE_EG], py[GE_EG], pz[GE_EG], acc[GE_EG];
#pragma unroll
    for (int q = 0; q < GE_EG; ++q) {
        const int e = e0 + (q < ne ? q : 0);
        px[q] = pos[3 * e];
        py[q] = pos[3 * e + 1];
        pz[q] = pos[3 * e + 2];
        acc[q] = 0.0;
    }
    const int64_t stride = (int64_t)gridDim.x * GBS;
    for (int64_t c = (int64_t)blockIdx.x * GBS + threadIdx.x; c < g.n; c += stride) {
        const double gx = grad[c], gy = grad[g.n + c], gz = grad[2 * g.n + c];
        double s[9];
        if (sig && sym) {
            const double* p = sig + c;
            s[0] = p[0]; s[1] = p[g.n]; s[2] = p[2 * g.n];
            s[3] = s[1]; s[4] = p[3 * g.n]; s[5] = p[4 * g.n];
            s[6] = s[2]; s[7] = s[5]; s[8] = p[5 * g.n];
        } else if (sig) {
#pragma unroll
            for (int q = 0; q < 9; ++q) s[q] = sig[q * g.n + c];
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) s[q] = s9.v[q];
        }
        const double vol = g.vol[c];
        const double dx = (s[0] * gx + s[1] * gy + s[2] * gz) * vol;
        const double dy = (s[3] * gx + s[4] * gy + s[5] * gz) * vol;
        const double dz = (s[6] * gx + s[7] * gy + s[8] * gz) * vol;
        const double cx = g.C[c], cy = g.C[g.n + c], cz = g.C[2 * g.n + c];
#pragma unroll
        for (int q = 0; q < GE_EG; ++q) {
            if (q < ne) {
                const double rx = cx - px[q], ry = cy - py[q], rz = cz - pz[q];
                const double r2 = rx * rx + ry * ry + rz * rz;
                // |r| > VSMALL as r2 > 0; 1/|r| from the reciprocal square root and one
                // Newton step
                if (r2 > 0.0) {
                    double y = __builtin_amdgcn_rsq(r2);
                    y = y * (1.5 - 0.5 * r2 * y * y);
                    acc[q] += (dx * rx + dy * ry + dz * rz) * (y * y * y);
                }
            }
        }
    }
    constexpr int NW = GBS / 64;
    __shared__ double red[NW][GE_EG];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < GE_EG; ++q) {
        if (q < ne) {
            double v = acc[q];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[wave][q] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < ne) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) v += red[w][threadIdx.x];
        partial[(int64_t)(e0 + threadIdx.x) * gridDim.x + blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(GBS) void k_gecg_sum(const double* __restrict__ partial,
                                                  int nblocks, double* __restrict__ out)
{
    __shared__ double red[GBS];
    const double* p = partial + (int64_t)blockIdx.x * nblocks;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += GBS) v += p[b];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = GBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

} // namespace

struct GeneralMesh {
    int64_t n = 0, F = 0, B = 0;
    int K = 0, KB = 0;
    // host geometry
    std::vector<int32_t> owner, nbr, bown;
    std::vector<double> Sf, Cf, C, vol, bSf;       // [F][3], [F][3], [n][3], [n], [B][3]
    std::vector<double> w, D, kv, dmag;           // [F], [F], [F][3], [F]
    std::vector<double> a;                        // [F] (after sigma)
    // device
    int32_t *slot = nullptr, *nbc = nullptr, *bslot = nullptr;
    double *dw = nullptr, *dSf = nullptr, *da = nullptr, *db = nullptr;
    double *dSfb = nullptr, *dbb = nullptr, *dnb = nullptr, *dvol = nullptr, *dC = nullptr;
    double* grad = nullptr;  // [3][n]
    bool sigma = false;
    GView view() const
    {
        return {n, K, KB, slot, nbc, bslot, dw, dSf, da, db, dSfb, dbb, dnb, dvol, dC, F, B};
    }
};

static bool dmalloc(void** p, size_t bytes)
{
    return hipMalloc(p, bytes ? bytes : 8) == hipSuccess;
}

int general_create(GeneralMesh** out, int64_t n, int64_t nf, const int32_t* owner,
                   const int32_t* nbr, const double* Sf, const double* Cf, const double* C,
                   const double* vol, int64_t nb, const int32_t* b_owner, const double* b_Sf,
                   std::string* err)
{
    *out = nullptr;
    auto bad = [&](const char* m) {
        if (err) *err = m;
        return 1;
    };
    if (n < 1 || nf < 0 || nb < 0 || n >= ((int64_t)1 << 31) || nf >= ((int64_t)1 << 31))
        return bad("bad mesh sizes");
    std::vector<int> deg(n, 0), bdeg(n, 0);
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t o = owner[f], q = nbr[f];
        if (o < 0 || o >= n || q < 0 || q >= n || o == q) return bad("face cell out of range");
        ++deg[o];
        ++deg[q];
    }
    for (int64_t f = 0; f < nb; ++f) {
        if (b_owner[f] < 0 || b_owner[f] >= n) return bad("boundary face cell out of range");
        ++bdeg[b_owner[f]];
    }
    for (int64_t c = 0; c < n; ++c)
        if (!(vol[c] > 0.0) || !std::isfinite(vol[c])) return bad("cell volume not positive");
    GeneralMesh* G = new GeneralMesh();
    G->n = n; G->F = nf; G->B = nb;
    G->K = *std::max_element(deg.begin(), deg.end());
    G->KB = *std::max_element(bdeg.begin(), bdeg.end());
    G->owner.assign(owner, owner + nf);
    G->nbr.assign(nbr, nbr + nf);
    G->bown.assign(b_owner, b_owner + nb);
    G->Sf.assign(Sf, Sf + 3 * nf);
    G->Cf.assign(Cf, Cf + 3 * nf);
    G->C.assign(C, C + 3 * n);
    G->vol.assign(vol, vol + n);
    G->bSf.assign(b_Sf, b_Sf + 3 * nb);
    G->w.resize(nf); G->D.resize(nf); G->kv.resize(3 * nf); G->dmag.resize(nf);
    for (int64_t f = 0; f < nf; ++f) {
        const double* s = &Sf[3 * f];
        const double* cf = &Cf[3 * f];
        const double* cp = &C[3 * owner[f]];
        const double* cn = &C[3 * nbr[f]];
        const double mag = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        if (!(mag > 0.0)) {
            delete G;
            return bad("face with zero area");
        }
        // surfaceInterpolation::makeWeights
        const double sown = std::fabs(s[0] * (cf[0] - cp[0]) + s[1] * (cf[1] - cp[1]) +
                                      s[2] * (cf[2] - cp[2]));
        const double snei = std::fabs(s[0] * (cn[0] - cf[0]) + s[1] * (cn[1] - cf[1]) +
                                      s[2] * (cn[2] - cf[2]));
        G->w[f] = snei / (sown + snei);
        const double d[3] = {cn[0] - cp[0], cn[1] - cp[1], cn[2] - cp[2]};
        const double nv[3] = {s[0] / mag, s[1] / mag, s[2] / mag};
        const double dm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        // makeNonOrthDeltaCoeffs and makeNonOrthCorrectionVectors
        const double D = 1.0 / std::max(nv[0] * d[0] + nv[1] * d[1] + nv[2] * d[2], 0.05 * dm);
        G->D[f] = D;
        G->dmag[f] = dm;
        for (int q = 0; q < 3; ++q) G->kv[3 * f + q] = nv[q] - D * d[q];
    }
    // cell-major slots
    std::vector<int32_t> slot((size_t)G->K * n, EMPTY), nbc((size_t)G->K * n, 0),
        bslot((size_t)std::max(G->KB, 1) * n, -1);
    std::vector<int> fill(n, 0), bfill(n, 0);
    for (int64_t f = 0; f < nf; ++f) {
        const int32_t o = owner[f], q = nbr[f];
        slot[(size_t)fill[o] * n + o] = (int32_t)f;
        nbc[(size_t)fill[o] * n + o] = q;
        ++fill[o];
        slot[(size_t)fill[q] * n + q] = ~(int32_t)f;
        nbc[(size_t)fill[q] * n + q] = o;
        ++fill[q];
    }
    for (int64_t c = 0; c < n; ++c)
        for (int k = fill[c]; k < G->K; ++k) nbc[(size_t)k * n + c] = (int32_t)c;
    for (int64_t f = 0; f < nb; ++f) {
        const int32_t o = b_owner[f];
        bslot[(size_t)bfill[o] * n + o] = (int32_t)f;
        ++bfill[o];
    }
    std::vector<double> soaSf(3 * nf), soaC(3 * n), soaSfb(3 * nb), soan(3 * nb);
    for (int64_t f = 0; f < nf; ++f)
        for (int q = 0; q < 3; ++q) soaSf[q * nf + f] = Sf[3 * f + q];
    for (int64_t c = 0; c < n; ++c)
        for (int q = 0; q < 3; ++q) soaC[q * n + c] = C[3 * c + q];
    for (int64_t f = 0; f < nb; ++f) {
        const double* s = &b_Sf[3 * f];
        const double mag = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (int q = 0; q < 3; ++q) {
            soaSfb[q * nb + f] = s[q];
            soan[q * nb + f] = mag > 0.0 ? s[q] / mag : 0.0;
        }
    }
    const size_t K = G->K, KB = std::max(G->KB, 1);
    bool ok = dmalloc((void**)&G->slot, 4 * K * n) && dmalloc((void**)&G->nbc, 4 * K * n) &&
              dmalloc((void**)&G->bslot, 4 * KB * n) && dmalloc((void**)&G->dw, 8 * nf) &&
              dmalloc((void**)&G->dSf, 24 * nf) && dmalloc((void**)&G->da, 8 * nf) &&
              dmalloc((void**)&G->db, 24 * nf) && dmalloc((void**)&G->dSfb, 24 * nb) &&
              dmalloc((void**)&G->dbb, 24 * nb) && dmalloc((void**)&G->dnb, 24 * nb) &&
              dmalloc((void**)&G->dvol, 8 * n) && dmalloc((void**)&G->dC, 24 * n) &&
              dmalloc((void**)&G->grad, 24 * n);
    if (!ok) {
        general_destroy(G);
        if (err) *err = "general mesh allocation failed";
        return 5;
    }
    auto up = [](void* d, const void* h, size_t bytes) {
        return bytes == 0 || hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    ok = up(G->slot, slot.data(), 4 * K * n) && up(G->nbc, nbc.data(), 4 * K * n) &&
         up(G->bslot, bslot.data(), 4 * KB * n) && up(G->dw, G->w.data(), 8 * nf) &&
         up(G->dSf, soaSf.data(), 24 * nf) && up(G->dSfb, soaSfb.data(), 24 * nb) &&
         up(G->dnb, soan.data(), 24 * nb) && up(G->dvol, vol, 8 * n) &&
         up(G->dC, soaC.data(), 24 * n);
    if (!ok) {
        general_destroy(G);
        if (err) *err = "general mesh upload failed";
        return 5;
    }
    *out = G;
    return 0;
}

int general_from_hex(GeneralMesh** out, const HexGeom& g, std::string* err)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t n = (int64_t)nx * ny * nz;
    const double h[3] = {g.hx, g.hy, g.hz};
    const double area[3] = {g.hy * g.hz, g.hx * g.hz, g.hx * g.hy};
    std::vector<int32_t> own, nbr, bown;
    std::vector<double> Sf, Cf, C(3 * n), V(n, g.hx * g.hy * g.hz), bSf;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int32_t c = i + nx * (j + ny * k);
                const double cc[3] = {(i + 0.5) * g.hx, (j + 0.5) * g.hy, (k + 0.5) * g.hz};
                for (int q = 0; q < 3; ++q) C[3 * c + q] = cc[q];
                const int idx[3] = {i, j, k}, nn[3] = {nx, ny, nz};
                const int32_t st[3] = {1, nx, nx * ny};
                for (int d = 0; d < 3; ++d) {
                    if (idx[d] < nn[d] - 1) {   // internal face towards +d
                        own.push_back(c);
                        nbr.push_back(c + st[d]);
                        for (int q = 0; q < 3; ++q) {
                            Sf.push_back(q == d ? area[d] : 0.0);
                            Cf.push_back(q == d ? cc[q] + 0.5 * h[d] : cc[q]);
                        }
                    }
                    for (int side = 0; side < 2; ++side) {  // walls
                        const bool wall = side == 0 ? idx[d] == 0 : idx[d] == nn[d] - 1;
                        if (!wall) continue;
                        bown.push_back(c);
                        for (int q = 0; q < 3; ++q)
                            bSf.push_back(q == d ? (side == 0 ? -area[d] : area[d]) : 0.0);
                    }
                }
            }
    return general_create(out, n, (int64_t)own.size(), own.data(), nbr.data(), Sf.data(),
                          Cf.data(), C.data(), V.data(), (int64_t)bown.size(), bown.data(),
                          bSf.data(), err);
}

void general_destroy(GeneralMesh* G)
{
    if (!G) return;
    void* p[] = {G->slot, G->nbc, G->bslot, G->dw, G->dSf, G->da, G->db,
                 G->dSfb, G->dbb, G->dnb, G->dvol, G->dC, G->grad};
    for (void* q : p)
        if (q) (void)hipFree(q);
    delete G;
}

int general_set_sigma(GeneralMesh* G, const double* s9, std::string* err)
{
    const int64_t nf = G->F, nb = G->B, n = G->n;
    std::vector<double> a(nf), b(3 * nf), bb(3 * nb);
    for (int64_t f = 0; f < nf; ++f) {
        const double* s = &G->Sf[3 * f];
        const double mag = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        const double nv[3] = {s[0] / mag, s[1] / mag, s[2] / mag};
        const double w = G->w[f];
        const double* sp = &s9[9 * G->owner[f]];
        const double* sn = &s9[9 * G->nbr[f]];
        double g[3];  // Sf . sigma_f, (v & T)_j = v_i T_ij
        for (int j = 0; j < 3; ++j) {
            g[j] = 0.0;
            for (int i = 0; i < 3; ++i)
                g[j] += s[i] * (w * sp[3 * i + j] + (1.0 - w) * sn[3 * i + j]);
        }
        const double gsn = g[0] * nv[0] + g[1] * nv[1] + g[2] * nv[2];
        a[f] = gsn * G->D[f];
        for (int q = 0; q < 3; ++q)
            b[q * nf + f] = gsn * G->kv[3 * f + q] + (g[q] - gsn * nv[q]);
    }
    for (int64_t f = 0; f < nb; ++f) {
        const double* s = &G->bSf[3 * f];
        const double mag = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        const double nv[3] = {s[0] / mag, s[1] / mag, s[2] / mag};
        const double* sp = &s9[9 * G->bown[f]];
        double g[3];
        for (int j = 0; j < 3; ++j) g[j] = s[0] * sp[j] + s[1] * sp[3 + j] + s[2] * sp[6 + j];
        const double gsn = g[0] * nv[0] + g[1] * nv[1] + g[2] * nv[2];
        for (int q = 0; q < 3; ++q) bb[q * nb + f] = g[q] - gsn * nv[q];
    }
    (void)n;
    if ((nf && (hipMemcpy(G->da, a.data(), 8 * nf, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(G->db, b.data(), 24 * nf, hipMemcpyHostToDevice) != hipSuccess)) ||
        (nb && hipMemcpy(G->dbb, bb.data(), 24 * nb, hipMemcpyHostToDevice) != hipSuccess)) {
        if (err) *err = "face coefficient upload failed";
        return 4;
    }
    G->a = a;
    G->sigma = true;
    return 0;
}

double general_explicit_dt(const GeneralMesh* G, double chiCm, double max_co)
{
    // Df = n . sigma_f . n / chiCm, dx = 1/deltaCoeffs = |d|, dt = maxCo min dx^2/Df
    // (SfGammaSn = |Sf| n . sigma_f . n, so n . sigma_f . n = a_f / (|Sf| D_f))
    double best = INFINITY;
    for (int64_t f = 0; f < G->F; ++f) {
        const double* s = &G->Sf[3 * f];
        const double mag = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        const double snn = G->a[f] / (mag * G->D[f]);
        const double q = G->dmag[f] * G->dmag[f] / (snn / chiCm);
        if (q < best) best = q;
    }
    return max_co * best;
}

const int32_t* general_owner(const GeneralMesh* G) { return G->owner.data(); }
const int32_t* general_neighbour(const GeneralMesh* G) { return G->nbr.data(); }
int64_t general_faces(const GeneralMesh* G) { return G->F; }

static unsigned grid_of(int64_t n) { return (unsigned)((n + GBS - 1) / GBS); }

hipError_t general_laplacian(GeneralMesh* G, const double* V, double* lap, hipStream_t s)
{
    const GView g = G->view();
    hipLaunchKernelGGL(k_ggrad, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, V, G->grad);
    hipLaunchKernelGGL(k_glap, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, V,
                       (const double*)G->grad, lap);
    return hipGetLastError();
}

hipError_t general_explicit(GeneralMesh* G, const double* Vn, const double* Vm1,
                            double* Vnew, const double* Iion, uint8_t* flag, double* tact,
                            const GeneralStep& st, hipStream_t s)
{
    const GView g = G->view();
    hipLaunchKernelGGL(k_ggrad, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, Vn, G->grad);
    hipLaunchKernelGGL(k_gexplicit, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, Vn,
                       (const double*)G->grad, Vm1, Vnew, Iion, flag, tact, st);
    return hipGetLastError();
}

void general_matrix(const GeneralMesh* G, double chiCm, double c, double dt, double* diag,
                    double* upper)
{
    for (int64_t i = 0; i < G->n; ++i) diag[i] = chiCm * c * G->vol[i] / dt;
    for (int64_t f = 0; f < G->F; ++f) {
        upper[f] = -G->a[f];
        diag[G->owner[f]] += G->a[f];
        diag[G->nbr[f]] += G->a[f];
    }
}

hipError_t general_rhs0(GeneralMesh* G, const double* Vn, const double* Vm1,
                        const double* Iion, double* b0, const GeneralStep& st, hipStream_t s)
{
    hipLaunchKernelGGL(k_grhs0, dim3(grid_of(G->n)), dim3(GBS), 0, s, G->view(), Vn, Vm1,
                       Iion, b0, st);
    return hipGetLastError();
}

hipError_t general_rhs_corr(GeneralMesh* G, const double* psi, const double* b0, double* b,
                            hipStream_t s)
{
    const GView g = G->view();
    hipLaunchKernelGGL(k_ggrad, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, psi, G->grad);
    hipLaunchKernelGGL(k_grhs_corr, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, psi,
                       (const double*)G->grad, b0, b);
    return hipGetLastError();
}

hipError_t general_pseudo_ecg(GeneralMesh* G, const double* V, const double* sig, int sym,
                              const double* s9, int n_el, const double* pos, int nblocks,
                              double* partial, double* out, hipStream_t s)
{
    const GView g = G->view();
    hipLaunchKernelGGL(k_ggrad, dim3(grid_of(G->n)), dim3(GBS), 0, s, g, V, G->grad);
    EcgSigma u{};
    for (int q = 0; q < 9; ++q) u.v[q] = s9 ? s9[q] : 0.0;
    const int64_t want = (G->n + GBS - 1) / GBS;
    const int nb = (int)(want < nblocks ? want : nblocks);
    for (int e0 = 0; e0 < n_el; e0 += GE_EG) {
        const int ne = n_el - e0 < GE_EG ? n_el - e0 : GE_EG;
        hipLaunchKernelGGL(k_gecg, dim3(nb), dim3(GBS), 0, s, g, (const double*)G->grad, sig,
                           sym, u, pos, e0, ne, partial);
    }
    hipLaunchKernelGGL(k_gecg_sum, dim3(n_el), dim3(GBS), 0, s, (const double*)partial, nb,
                       out);
    return hipGetLastError();
}

hipError_t general_guess(GeneralMesh* G, const double* Vn, const double* Vm1, double* psi,
                         bool extrapolate, hipStream_t s)
{
    if (!extrapolate)
        return hipMemcpyAsync(psi, Vn, sizeof(double) * G->n, hipMemcpyDeviceToDevice, s);
    hipLaunchKernelGGL(k_gguess, dim3(grid_of(G->n)), dim3(GBS), 0, s, G->n, Vn, Vm1, psi);
    return hipGetLastError();
}

hipError_t general_activation(GeneralMesh* G, const double* Vn, const double* Vnew,
                              uint8_t* flag, double* tact, double t_old, double dt,
                              double vthr, hipStream_t s)
{
    hipLaunchKernelGGL(k_gactivation, dim3(grid_of(G->n)), dim3(GBS), 0, s, G->n, Vn, Vnew,
                       flag, tact, t_old, dt, vthr);
    return hipGetLastError();
}

} // namespace cg
