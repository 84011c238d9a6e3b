// Reverse-mode gradients of the explicit monodomain step (BuenoOrovio, forward-Euler ddt,
// FP64) with respect to a per-cell conductivity scale kappa and the model constants
// tau_fi and tau_si, for losses built on Vm at probe cells.
//
// Forward map of one step n -> n+1, cell c (monodomainSolver.C:103-118 with ddt Euler,
// BuenoOrovio.C:107-152 scalings, RL gates with u frozen):
//   u = (1000 V + 84) / 85.7
//   (v', w', s') = RL(v, w, s; u),  J = Jion(u, v', w', s'),  I = 85.7 J
//   V' = V + dt/chiCm (L(kappa) V - chiCm I + src)
//   (L V)_c = sum_f a_f (V_nb - V_c),  a_f = 1/2 (kappa_c + kappa_nb) sigma0_dd / h_d^2
// Adjoint of step n, with mu = dLoss/dV^{n+1} and (lv', lw', ls') = dLoss/d(v', w', s'):
//   lJ = -dt 85.7 mu
//   (lv, lw, ls, lu, dtau_fi, dtau_si) = vector-Jacobian product of the cell step
//   dLoss/dV^n = mu + dt/chiCm (L mu) + (1000/85.7) lu + probe seed of V^n
//   dLoss/dkappa_c += dt/chiCm sum_{f of c} 1/2 sigma0_dd/h^2 (V_nb - V_c)(mu_c - mu_nb)
// The derivatives of the gates and currents are written out by hand from the model
// equations (BuenoOrovio_2008.H:166-225); Heaviside switches contribute zero derivative.
//
// Checkpoints: the forward pass stores (V, v, w, s) every K steps; the backward pass
// re-runs each segment from its checkpoint into a (K+1)-slot store, then sweeps it in
// reverse.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "cg_adjoint.hpp"
#include "cg_hexops.hpp"
#include "model_kernels.hpp"

namespace cg {

namespace {

constexpr int ABS = 256;
constexpr int AG = 1024;  // blocks of the backward kernel (parameter partials)

using BO = BuenoOrovio;

struct AdjStep {
    HexGeom g;
    const double* fx;  // face coefficients a_f (per unit volume), + faces, 0 on walls
    const double* fy;
    const double* fz;
    double dt, dt_model, chiCm;
    ActiveBoxes boxes;
};

// ------------------------------------------------------------------ forward
// One step from slot (Vin, yin[3][n]) into (Vout, yout); Vm at probes into probe_row.
__global__ __launch_bounds__(ABS) void k_adj_forward(AdjStep s, KParams<double> P,
                                                     const double* __restrict__ Vin,
                                                     const double* __restrict__ yin,
                                                     double* __restrict__ Vout,
                                                     double* __restrict__ yout,
                                                     const int* __restrict__ probe_slot,
                                                     double* __restrict__ probe_row)
{
    const HexGeom& g = s.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c >= n) return;
    const IJK q = ijk_of((uint32_t)c, g);
    double p[BO::NP];
#pragma unroll
    for (int k = 0; k < BO::NP; ++k) p[k] = P.p[k];
    const double one[4] = {1, 1, 1, 1};
    double y[BO::NS];
    y[1] = yin[c];
    y[2] = yin[n + c];
    y[3] = yin[2 * n + c];
    const double vn = Vin[c];
    const double u = (vn * 1000.0 + BO::VOFF) / BO::VSCALE;
    const double iion = Ops<BO>::tissue_step(u, y, s.dt_model, p, 1, one) * BO::ISCALE;
    yout[c] = y[1];
    yout[n + c] = y[2];
    yout[2 * n + c] = y[3];
    const HexCoefs a{0, 0, 0, s.fx, s.fy, s.fz};
    const double lap = hex_laplacian<false>(Vin, c, q, g, a);
    const double src = box_source(q.i, q.j, q.k, g, s.boxes);
    const double vnew = (1.0 * vn + s.dt / s.chiCm * (lap - s.chiCm * iion + src)) / 1.0;
    Vout[c] = vnew;
    if (probe_row) {
        const int slot = probe_slot[c];
        if (slot >= 0) probe_row[slot] = vnew;
    }
}

// face coefficients from kappa: a_x(c) = 1/2 (kappa_c + kappa_{c+x}) gx, 0 on the + wall
__global__ __launch_bounds__(ABS) void k_adj_faces(HexGeom g, const double* __restrict__ kappa,
                                                   double gx, double gy, double gz,
                                                   double* __restrict__ fx,
                                                   double* __restrict__ fy,
                                                   double* __restrict__ fz)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c >= n) return;
    const IJK q = ijk_of((uint32_t)c, g);
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const double kc = kappa[c];
    fx[c] = q.i < g.nx - 1 ? 0.5 * (kc + kappa[c + 1]) * gx : 0.0;
    fy[c] = q.j < g.ny - 1 ? 0.5 * (kc + kappa[c + sy]) * gy : 0.0;
    fz[c] = q.k < g.nz - 1 ? 0.5 * (kc + kappa[c + sz]) * gz : 0.0;
}

// ----------------------------------------------------------------- cell VJP
struct CellAdj {
    double lv, lw, ls, lu, dtfi, dtsi;
};

// Vector-Jacobian product of the BuenoOrovio tissue step (u frozen) at (u, v, w, s):
// cotangents (lv1, lw1, ls1) of the updated gates and lJ of Jion at the step end.
CG_DEV CellAdj bo_vjp(double u, double v, double w, double s, double dt, const double* p,
                      double lv1, double lw1, double ls1, double lJ)
{
    const bool Hv = pos(u - p[BO::thv]), Hw = pos(u - p[BO::thw]), Ho = pos(u - p[BO::tho]),
               Hvm = pos(u - p[BO::thvm]);
    // gates, as BuenoOrovio::gates
    const double vinf = Hv ? 0.0 : (pos(p[BO::thvm] - u) ? 1.0 : 0.0);
    const double tv = Hv ? p[BO::tvp] : (Hvm ? p[BO::tv2m] : p[BO::tv1m]);
    const double thw_ = tanh(p[BO::kwm] * (u - p[BO::uwm]));
    const double twm = p[BO::tw1m] + (p[BO::tw2m] - p[BO::tw1m]) * (1.0 + thw_) / 2.0;
    const double dtwm = (p[BO::tw2m] - p[BO::tw1m]) * p[BO::kwm] * (1.0 - thw_ * thw_) / 2.0;
    const double winf = Hw ? 0.0 : (Ho ? p[BO::winfstar] : 1.0 - u / p[BO::twinf]);
    const double dwinf = (!Hw && !Ho) ? -1.0 / p[BO::twinf] : 0.0;
    const double tw = Hw ? p[BO::twp] : twm;
    const double dtw = Hw ? 0.0 : dtwm;
    const double ths = tanh(p[BO::ks] * (u - p[BO::us]));
    const double sinf = (1.0 + ths) / 2.0;
    const double dsinf = p[BO::ks] * (1.0 - ths * ths) / 2.0;
    const double ts = Hw ? p[BO::ts2] : p[BO::ts1];
    const double Ev = exp(-dt / tv), Ew = exp(-dt / tw), Es = exp(-dt / ts);
    const double v1 = vinf + (v - vinf) * Ev;
    const double w1 = winf + (w - winf) * Ew;
    const double s1 = sinf + (s - sinf) * Es;
    const double dw1du = dwinf * (1.0 - Ew) + (w - winf) * Ew * dt / (tw * tw) * dtw;
    const double ds1du = dsinf * (1.0 - Es);
    // currents at (u, v1, w1, s1), as BuenoOrovio::jion
    const double tfi = p[BO::tfi], tsi = p[BO::tsi];
    const double a1 = u - p[BO::thv], a2 = p[BO::uu] - u;
    const double thso = tanh(p[BO::kso] * (u - p[BO::uso]));
    const double tso = p[BO::tso1] + (p[BO::tso2] - p[BO::tso1]) * (1.0 + thso) / 2.0;
    const double dtso = (p[BO::tso2] - p[BO::tso1]) * p[BO::kso] * (1.0 - thso * thso) / 2.0;
    const double to = Ho ? p[BO::to2] : p[BO::to1];
    const double dJ_dv1 = Hv ? -a1 * a2 / tfi : 0.0;
    const double dJ_dw1 = Hw ? -s1 / tsi : 0.0;
    const double dJ_ds1 = Hw ? -w1 / tsi : 0.0;
    const double dJ_du = (Hv ? -v1 * (a2 - a1) / tfi : 0.0) + (Hw ? -dtso / (tso * tso) : 1.0 / to);
    const double cv = lv1 + lJ * dJ_dv1, cw = lw1 + lJ * dJ_dw1, cs = ls1 + lJ * dJ_ds1;
    CellAdj r;
    r.lv = cv * Ev;
    r.lw = cw * Ew;
    r.ls = cs * Es;
    r.lu = lJ * dJ_du + cw * dw1du + cs * ds1du;  // v1 does not depend on u (a.e.)
    r.dtfi = Hv ? lJ * v1 * a1 * a2 / (tfi * tfi) : 0.0;
    r.dtsi = Hw ? lJ * w1 * s1 / (tsi * tsi) : 0.0;
    return r;
}

template <int NV> CG_DEV void block_sum_a(double (&v)[NV])
{
    constexpr int NW = ABS / 64;
    __shared__ double sm[NV][NW];
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sm[k][w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = 0.0;
        for (int q = 0; q < NW; ++q) t += sm[k][q];
        v[k] = t;
    }
}

// ----------------------------------------------------------------- backward
// Adjoint of step n.  mu = dLoss/dV^{n+1} (includes the probe seed of V^{n+1}),
// ly1 = dLoss/dy^{n+1}; writes lV = dLoss/dV^n (plus the seed of V^n from seed_row when
// not null), ly = dLoss/dy^n, adds into gk and into the per-block parameter partials.
// Grid-stride over cells with AG blocks (fixed block -> cell assignment, so the partial
// of block b is always accumulated by block b).
__global__ __launch_bounds__(ABS) void k_adj_backward(AdjStep s, KParams<double> P, double gxh,
                                                      double gyh, double gzh,
                                                      const double* __restrict__ Vn,
                                                      const double* __restrict__ yn,
                                                      const double* __restrict__ mu,
                                                      const double* __restrict__ ly1,
                                                      double* __restrict__ lV,
                                                      double* __restrict__ ly,
                                                      const int* __restrict__ probe_slot,
                                                      const double* __restrict__ seed_row,
                                                      double* __restrict__ gk,
                                                      double* __restrict__ gtheta)
{
    const HexGeom& g = s.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    double p[BO::NP];
#pragma unroll
    for (int k = 0; k < BO::NP; ++k) p[k] = P.p[k];
    const double a = s.dt / s.chiCm;
    const double lJ_per_mu = -s.dt * BO::ISCALE;
    const double du_dV = 1000.0 / BO::VSCALE;
    double acc[2] = {0.0, 0.0};
    for (int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x; c < n; c += (int64_t)gridDim.x * ABS) {
        const IJK q = ijk_of((uint32_t)c, g);
        const Nbr b = nbr_of(c, q, g);
        const double muc = mu[c];
        const double u = (Vn[c] * 1000.0 + BO::VOFF) / BO::VSCALE;
        const CellAdj r = bo_vjp(u, yn[c], yn[n + c], yn[2 * n + c], s.dt_model, p, ly1[c],
                                 ly1[n + c], ly1[2 * n + c], lJ_per_mu * muc);
        ly[c] = r.lv;
        ly[n + c] = r.lw;
        ly[2 * n + c] = r.ls;
        acc[0] += r.dtfi;
        acc[1] += r.dtsi;
        // diffusion adjoint (L symmetric) and the kappa gradient, one pass over the faces
        const double fxp = s.fx[c], fxm = s.fx[b.xm], fyp = s.fy[c], fym = s.fy[b.ym];
        const double fzp = s.fz[c], fzm = s.fz[b.zm];
        const double mxp = mu[b.xp], mxm = mu[b.xm], myp = mu[b.yp], mym = mu[b.ym];
        const double mzp = mu[b.zp], mzm = mu[b.zm];
        const double vc = Vn[c];
        const double Lmu = fxp * (mxp - muc) + fxm * (mxm - muc) + fyp * (myp - muc) +
                           fym * (mym - muc) + fzp * (mzp - muc) + fzm * (mzm - muc);
        double lv = muc + a * Lmu + du_dV * r.lu;
        if (seed_row) {
            const int slot = probe_slot[c];
            if (slot >= 0) lv += seed_row[slot];
        }
        lV[c] = lv;
        const double dk = gxh * ((Vn[b.xp] - vc) * (muc - mxp) + (Vn[b.xm] - vc) * (muc - mxm)) +
                          gyh * ((Vn[b.yp] - vc) * (muc - myp) + (Vn[b.ym] - vc) * (muc - mym)) +
                          gzh * ((Vn[b.zp] - vc) * (muc - mzp) + (Vn[b.zm] - vc) * (muc - mzm));
        gk[c] += a * dk;
    }
    block_sum_a<2>(acc);
    if (threadIdx.x == 0) {
        gtheta[blockIdx.x * 2 + 0] += acc[0];
        gtheta[blockIdx.x * 2 + 1] += acc[1];
    }
}

// lV = seed of V^N (the last probe row) at probes, 0 elsewhere; ly = 0
__global__ __launch_bounds__(ABS) void k_adj_seed(int64_t n, const int* __restrict__ probe_slot,
                                                  const double* __restrict__ seed_row,
                                                  double* __restrict__ lV,
                                                  double* __restrict__ ly)
{
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c >= n) return;
    const int slot = probe_slot[c];
    lV[c] = slot >= 0 ? seed_row[slot] : 0.0;
    ly[c] = 0.0;
    ly[n + c] = 0.0;
    ly[2 * n + c] = 0.0;
}

__global__ __launch_bounds__(ABS) void k_fill(int64_t n, double* __restrict__ x, double v)
{
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c < n) x[c] = v;
}

unsigned blocks_of(int64_t n) { return (unsigned)((n + ABS - 1) / ABS); }

} // namespace

// ------------------------------------------------------------------ host side
struct Adjoint {
    HexGeom g;
    int64_t n;
    double gx, gy, gz;          // sigma0_dd / h_d^2 (kappa = 1 conductivity)
    double dt, dt_model, chiCm;
    ModelParams P;
    int K;                      // checkpoint interval
    int n_probes;
    std::vector<int64_t> probes;
    int* d_slot = nullptr;      // [n] probe slot or -1
    double* d_V0 = nullptr;     // initial state
    double* d_y0 = nullptr;     // [3][n]
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double* ck = nullptr;       // checkpoints [n_ck][4][n]
    int n_ck = 0;
    double* seg = nullptr;      // recompute store [K+1][4][n]
    double *lV[2] = {nullptr, nullptr}, *ly[2] = {nullptr, nullptr};
    double* gtheta = nullptr;   // [AG][2]
    int last_steps = 0;
    std::vector<ActiveBoxes> boxes;  // per step of the last forward
    hipStream_t stream;
};

void adjoint_destroy(Adjoint* A)
{
    if (!A) return;
    void* p[] = {A->d_slot, A->d_V0, A->d_y0, A->fx,   A->fy,   A->fz,   A->ck,
                 A->seg,    A->lV[0], A->lV[1], A->ly[0], A->ly[1], A->gtheta};
    for (void* q : p)
        if (q) hipFree(q);
    delete A;
}

int adjoint_create(Adjoint** out, const HexGeom& g, const double* sigma0_diag, double dt,
                   double chiCm, const ModelParams& P, const double* V0, const void* y0_states,
                   int K, int n_probes, const int64_t* probes, hipStream_t stream,
                   std::string* err)
{
    Adjoint* A = new Adjoint();
    A->g = g;
    A->n = (int64_t)g.nx * g.ny * g.nz;
    A->gx = sigma0_diag[0] / (g.hx * g.hx);
    A->gy = sigma0_diag[1] / (g.hy * g.hy);
    A->gz = sigma0_diag[2] / (g.hz * g.hz);
    A->dt = dt;
    A->dt_model = dt * 1000.0;
    A->chiCm = chiCm;
    A->P = P;
    A->K = K;
    A->n_probes = n_probes;
    A->probes.assign(probes, probes + n_probes);
    A->stream = stream;
    const int64_t n = A->n;
    bool ok = true;
    auto M = [&](void** p, size_t b) { ok &= hipMalloc(p, b) == hipSuccess; };
    M((void**)&A->d_slot, sizeof(int) * n);
    M((void**)&A->d_V0, 8 * n);
    M((void**)&A->d_y0, 8 * 3 * n);
    M((void**)&A->fx, 8 * n); M((void**)&A->fy, 8 * n); M((void**)&A->fz, 8 * n);
    M((void**)&A->seg, 8 * 4 * n * (size_t)(K + 1));
    for (int i = 0; i < 2; ++i) { M((void**)&A->lV[i], 8 * n); M((void**)&A->ly[i], 8 * 3 * n); }
    M((void**)&A->gtheta, 8 * 2 * AG);
    if (!ok) {
        adjoint_destroy(A);
        *err = "adjoint allocation failed";
        return 5;
    }
    std::vector<int> slot((size_t)n, -1);
    for (int k = 0; k < n_probes; ++k) {
        if (probes[k] < 0 || probes[k] >= n) {
            adjoint_destroy(A);
            *err = "probe cell out of range";
            return 1;
        }
        if (slot[probes[k]] >= 0) {
            adjoint_destroy(A);
            *err = "probe cell listed twice";
            return 1;
        }
        slot[probes[k]] = k;
    }
    hipMemcpy(A->d_slot, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    hipMemcpy(A->d_V0, V0, 8 * n, hipMemcpyDeviceToDevice);
    // gates v, w, s: state planes 1..3 of the ionic handle
    hipMemcpy(A->d_y0, (const double*)y0_states + n, 8 * 3 * n, hipMemcpyDeviceToDevice);
    *out = A;
    return 0;
}

static AdjStep step_of(const Adjoint* A, const ActiveBoxes& bx)
{
    AdjStep s;
    s.g = A->g;
    s.fx = A->fx;
    s.fy = A->fy;
    s.fz = A->fz;
    s.dt = A->dt;
    s.dt_model = A->dt_model;
    s.chiCm = A->chiCm;
    s.boxes = bx;
    return s;
}

int adjoint_forward(Adjoint* A, const double* kappa, const double* theta, const int* theta_index,
                    int n_theta, int n_steps, const std::vector<ActiveBoxes>& boxes,
                    double* probes, std::string* err)
{
    const int64_t n = A->n;
    const unsigned nb = blocks_of(n);
    for (int k = 0; k < n_theta; ++k) A->P.p[theta_index[k]] = theta[k];
    const int n_ck = (n_steps + A->K - 1) / A->K;
    if (n_ck > A->n_ck) {
        if (A->ck) hipFree(A->ck);
        A->ck = nullptr;
        if (hipMalloc((void**)&A->ck, 8 * 4 * n * (size_t)n_ck) != hipSuccess) {
            A->n_ck = 0;
            *err = "checkpoint allocation failed";
            return 5;
        }
        A->n_ck = n_ck;
    }
    hipLaunchKernelGGL(k_adj_faces, dim3(nb), dim3(ABS), 0, A->stream, A->g, kappa, A->gx,
                       A->gy, A->gz, A->fx, A->fy, A->fz);
    // ping-pong between two slots of the segment store
    double* slot[2] = {A->seg, A->seg + 4 * n};
    hipMemcpyAsync(slot[0], A->d_V0, 8 * n, hipMemcpyDeviceToDevice, A->stream);
    hipMemcpyAsync(slot[0] + n, A->d_y0, 8 * 3 * n, hipMemcpyDeviceToDevice, A->stream);
    const KParams<double> KP = kparams<double>(A->P);
    for (int st = 0; st < n_steps; ++st) {
        double* in = slot[st & 1];
        double* outp = slot[(st + 1) & 1];
        if (st % A->K == 0)
            hipMemcpyAsync(A->ck + (size_t)(st / A->K) * 4 * n, in, 8 * 4 * n,
                           hipMemcpyDeviceToDevice, A->stream);
        hipLaunchKernelGGL(k_adj_forward, dim3(nb), dim3(ABS), 0, A->stream, step_of(A, boxes[st]),
                           KP, in, in + n, outp, outp + n, A->d_slot,
                           probes + (size_t)st * A->n_probes);
    }
    A->last_steps = n_steps;
    A->boxes = boxes;
    const hipError_t e = hipStreamSynchronize(A->stream);
    if (e != hipSuccess || hipGetLastError() != hipSuccess) {
        *err = std::string("adjoint forward: ") + hipGetErrorString(e);
        return 4;
    }
    return 0;
}

int adjoint_backward(Adjoint* A, const double* dprobes, double* dkappa, double* dtheta,
                     std::string* err)
{
    const int64_t n = A->n;
    const int N = A->last_steps;
    if (N <= 0) {
        *err = "backward before forward";
        return 9;
    }
    const unsigned nb = blocks_of(n);
    const KParams<double> KP = kparams<double>(A->P);
    const double gxh = 0.5 * A->gx, gyh = 0.5 * A->gy, gzh = 0.5 * A->gz;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(ABS), 0, A->stream, n, dkappa, 0.0);
    hipLaunchKernelGGL(k_fill, dim3((2 * AG + ABS - 1) / ABS), dim3(ABS), 0, A->stream,
                       (int64_t)2 * AG, A->gtheta, 0.0);
    // lV = dLoss/dV^N from the last probe row
    int cur = 0;
    hipLaunchKernelGGL(k_adj_seed, dim3(nb), dim3(ABS), 0, A->stream, n, A->d_slot,
                       dprobes + (size_t)(N - 1) * A->n_probes, A->lV[cur], A->ly[cur]);
    const int n_seg = (N + A->K - 1) / A->K;
    for (int sgi = n_seg - 1; sgi >= 0; --sgi) {
        const int s0 = sgi * A->K, s1 = s0 + A->K < N ? s0 + A->K : N;
        // recompute the segment: slot j holds the state at step s0 + j
        hipMemcpyAsync(A->seg, A->ck + (size_t)sgi * 4 * n, 8 * 4 * n, hipMemcpyDeviceToDevice,
                       A->stream);
        for (int st = s0; st < s1 - 1; ++st) {
            double* in = A->seg + (size_t)(st - s0) * 4 * n;
            double* outp = in + 4 * n;
            hipLaunchKernelGGL(k_adj_forward, dim3(nb), dim3(ABS), 0, A->stream,
                               step_of(A, A->boxes[st]), KP, in, in + n, outp, outp + n,
                               A->d_slot, (double*)nullptr);
        }
        // reverse sweep over the segment's steps
        for (int st = s1 - 1; st >= s0; --st) {
            const double* Vn = A->seg + (size_t)(st - s0) * 4 * n;
            const int nxt = cur ^ 1;
            const double* seed = st >= 1 ? dprobes + (size_t)(st - 1) * A->n_probes : nullptr;
            hipLaunchKernelGGL(k_adj_backward, dim3(AG), dim3(ABS), 0, A->stream,
                               step_of(A, A->boxes[st]), KP, gxh, gyh, gzh, Vn, Vn + n,
                               (const double*)A->lV[cur], (const double*)A->ly[cur], A->lV[nxt],
                               A->ly[nxt], A->d_slot, seed, dkappa, A->gtheta);
            cur = nxt;
        }
    }
    std::vector<double> gt(2 * AG);
    hipError_t e = hipMemcpyAsync(gt.data(), A->gtheta, 8 * 2 * AG, hipMemcpyDeviceToHost,
                                  A->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(A->stream);
    if (e != hipSuccess || hipGetLastError() != hipSuccess) {
        *err = std::string("adjoint backward: ") + hipGetErrorString(e);
        return 4;
    }
    double s0 = 0.0, s1 = 0.0;
    for (int b = 0; b < AG; ++b) {
        s0 += gt[2 * b];
        s1 += gt[2 * b + 1];
    }
    dtheta[0] = s0;
    dtheta[1] = s1;
    return 0;
}

} // namespace cg
