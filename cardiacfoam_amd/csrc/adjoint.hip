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
// Implicit (backward-Euler) diffusion, monodomainSolver.C:121-136 with ddt Euler and
// fvm::laplacian: M V' = V - dt I + a src, M = 1 - a L, a = dt/chiCm (M symmetric
// positive definite).  Adjoint: lambda = M^-1 mu, then the cell VJP with lJ = -dt 85.7
// lambda and dLoss/dV^n = lambda + (1000/85.7) lu; the kappa gradient is the same face sum
// with V^{n+1} and lambda in place of V^n and mu.  Both solves are Jacobi PCG to a relative
// residual of 1e-14 (tight, so that the forward map the gradient differentiates is the
// solved one).
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

// slot layout of the state store: V, v, w, s, then the step's transcendental factors
// exp(-dt/tau_v), exp(-dt/tau_w), exp(-dt/tau_s), tanh of tau_w-, s_inf and tau_so
enum { SL_V, SL_v, SL_w, SL_s, SL_Ev, SL_Ew, SL_Es, SL_thw, SL_ths, SL_thso, SL_N };

// stimulus boxes of the tissue (myocardiumDomain.C:307-342), active when the step start
// t0 lies in [t_start, t_start + duration]
struct BoxTable {
    int n;
    double lo[MAX_BOXES][3], hi[MAX_BOXES][3], t0[MAX_BOXES], dur[MAX_BOXES], I[MAX_BOXES];
    const uint8_t* mask;  // bit b of mask[c]: box b contains cell c's centre (the mesh's
                          // own centres); null: the (i+1/2)h lattice
};

struct AdjGeom {
    HexGeom g;
    const double* fx;   // face coefficients a_f (per unit volume), + faces, 0 on walls
    const double* fy;
    const double* fz;
    const double* prm;  // BuenoOrovio constants (device)
    const BoxTable* boxes;
    double dt, dt_model, chiCm;
};

CG_DEV double box_src(int i, int j, int k, const HexGeom& g, const BoxTable& b, double t0)
{
    double src = 0.0;
    const double p0 = (i + 0.5) * g.hx, p1 = (j + 0.5) * g.hy, p2 = (k + 0.5) * g.hz;
    for (int q = 0; q < b.n; ++q) {
        if (t0 < b.t0[q] || t0 > b.t0[q] + b.dur[q]) continue;
        if (b.mask) {
            if ((b.mask[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)] >> q) & 1u) src += b.I[q];
            continue;
        }
        if (p0 >= b.lo[q][0] && p0 <= b.hi[q][0] && p1 >= b.lo[q][1] && p1 <= b.hi[q][1] &&
            p2 >= b.lo[q][2] && p2 <= b.hi[q][2])
            src += b.I[q];
    }
    return src;
}

// The BuenoOrovio tissue step (u frozen) written out with its transcendental factors,
// operation for operation as BuenoOrovio::gates and ::jion; returns Jion at the step end.
struct BoFactors {
    double Ev, Ew, Es, thw, ths, thso;
};
CG_DEV double bo_step(double u, double& v, double& w, double& s, double dt, const double* p,
                      BoFactors& f)
{
    const bool Hv = pos(u - p[BO::thv]), Hw = pos(u - p[BO::thw]), Ho = pos(u - p[BO::tho]),
               Hvm = pos(u - p[BO::thvm]);
    const double vinf = Hv ? 0.0 : (pos(p[BO::thvm] - u) ? 1.0 : 0.0);
    const double tv = Hv ? p[BO::tvp] : (Hvm ? p[BO::tv2m] : p[BO::tv1m]);
    f.thw = tanh(p[BO::kwm] * (u - p[BO::uwm]));
    const double twm = p[BO::tw1m] + (p[BO::tw2m] - p[BO::tw1m]) * (1.0 + f.thw) / 2.0;
    const double winf = Hw ? 0.0 : (Ho ? p[BO::winfstar] : 1.0 - u / p[BO::twinf]);
    const double tw = Hw ? p[BO::twp] : twm;
    f.ths = tanh(p[BO::ks] * (u - p[BO::us]));
    const double sinf = (1.0 + f.ths) / 2.0;
    const double ts = Hw ? p[BO::ts2] : p[BO::ts1];
    f.Ev = exp(-dt / tv);
    f.Ew = exp(-dt / tw);
    f.Es = exp(-dt / ts);
    v = vinf + (v - vinf) * f.Ev;
    w = winf + (w - winf) * f.Ew;
    s = sinf + (s - sinf) * f.Es;
    f.thso = tanh(p[BO::kso] * (u - p[BO::uso]));
    const double tso = p[BO::tso1] + (p[BO::tso2] - p[BO::tso1]) * (1.0 + f.thso) / 2.0;
    const double to = Ho ? p[BO::to2] : p[BO::to1];
    const double jfi = Hv ? -v * (u - p[BO::thv]) * (p[BO::uu] - u) / p[BO::tfi] : 0.0;
    const double jso = Hw ? 1.0 / tso : (u - p[BO::uo]) / to;
    const double jsi = Hw ? -w * s / p[BO::tsi] : 0.0;
    return jfi + jso + jsi;
}

// ------------------------------------------------------------------ forward
// One step from slot `in` into slot `out` (planes SL_V..SL_s, [plane][n]); when aux is
// not null (planes SL_Ev.. of the input slot) the step's factors are stored there; Vm at
// probes into probe_row.
// IMPL: the diffusion is implicit; instead of V^{n+1} the kernel writes the right-hand
// side V - dt I + a src of M V^{n+1} into bvec and V^n (the solve's initial guess) into the
// output V plane, and records no probes (the solve's result is V^{n+1}).
template <bool IMPL>
__global__ __launch_bounds__(ABS) void k_adj_forward(AdjGeom G, double t0,
                                                     const double* __restrict__ in,
                                                     double* __restrict__ out,
                                                     double* __restrict__ aux,
                                                     const int* __restrict__ probe_slot,
                                                     double* __restrict__ probe_row,
                                                     double* __restrict__ bvec)
{
    const HexGeom& g = G.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c >= n) return;
    const IJK q = ijk_of((uint32_t)c, g);
    double v = in[SL_v * n + c], w = in[SL_w * n + c], s = in[SL_s * n + c];
    const double vn = in[SL_V * n + c];
    const double u = (vn * 1000.0 + BO::VOFF) / BO::VSCALE;
    BoFactors f;
    const double iion = bo_step(u, v, w, s, G.dt_model, G.prm, f) * BO::ISCALE;
    out[SL_v * n + c] = v;
    out[SL_w * n + c] = w;
    out[SL_s * n + c] = s;
    if (aux) {  // planes SL_Ev.. of the input slot
        aux[(SL_Ev - SL_Ev) * n + c] = f.Ev;
        aux[(SL_Ew - SL_Ev) * n + c] = f.Ew;
        aux[(SL_Es - SL_Ev) * n + c] = f.Es;
        aux[(SL_thw - SL_Ev) * n + c] = f.thw;
        aux[(SL_ths - SL_Ev) * n + c] = f.ths;
        aux[(SL_thso - SL_Ev) * n + c] = f.thso;
    }
    const double src = box_src(q.i, q.j, q.k, g, *G.boxes, t0);
    if constexpr (IMPL) {
        bvec[c] = vn + G.dt / G.chiCm * (src - G.chiCm * iion);
        out[SL_V * n + c] = vn;
        return;
    }
    const HexCoefs hc{0, 0, 0, G.fx, G.fy, G.fz};
    const double lap = hex_laplacian<false>(in + SL_V * n, c, q, g, hc);
    const double vnew = (1.0 * vn + G.dt / G.chiCm * (lap - G.chiCm * iion + src)) / 1.0;
    out[SL_V * n + c] = vnew;
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

// Vector-Jacobian product of bo_step at (u, v, w, s) with its stored factors f: cotangents
// (lv1, lw1, ls1) of the updated gates and lJ of Jion.  Derivatives by hand from the
// model equations; the Heaviside switches are piecewise constant (zero derivative).
CG_DEV CellAdj bo_vjp(double u, double v, double w, double s, const BoFactors& f, double dt,
                      const double* p, double lv1, double lw1, double ls1, double lJ)
{
    const bool Hv = pos(u - p[BO::thv]), Hw = pos(u - p[BO::thw]), Ho = pos(u - p[BO::tho]);
    const double vinf = Hv ? 0.0 : (pos(p[BO::thvm] - u) ? 1.0 : 0.0);
    const double twm = p[BO::tw1m] + (p[BO::tw2m] - p[BO::tw1m]) * (1.0 + f.thw) / 2.0;
    const double dtwm = (p[BO::tw2m] - p[BO::tw1m]) * p[BO::kwm] * (1.0 - f.thw * f.thw) / 2.0;
    const double winf = Hw ? 0.0 : (Ho ? p[BO::winfstar] : 1.0 - u / p[BO::twinf]);
    const double dwinf = (!Hw && !Ho) ? -1.0 / p[BO::twinf] : 0.0;
    const double tw = Hw ? p[BO::twp] : twm;
    const double dtw = Hw ? 0.0 : dtwm;
    const double sinf = (1.0 + f.ths) / 2.0;
    const double dsinf = p[BO::ks] * (1.0 - f.ths * f.ths) / 2.0;
    const double v1 = vinf + (v - vinf) * f.Ev;
    const double w1 = winf + (w - winf) * f.Ew;
    const double s1 = sinf + (s - sinf) * f.Es;
    const double dw1du = dwinf * (1.0 - f.Ew) + (w - winf) * f.Ew * dt / (tw * tw) * dtw;
    const double ds1du = dsinf * (1.0 - f.Es);
    const double tfi = p[BO::tfi], tsi = p[BO::tsi];
    const double a1 = u - p[BO::thv], a2 = p[BO::uu] - u;
    const double tso = p[BO::tso1] + (p[BO::tso2] - p[BO::tso1]) * (1.0 + f.thso) / 2.0;
    const double dtso = (p[BO::tso2] - p[BO::tso1]) * p[BO::kso] * (1.0 - f.thso * f.thso) / 2.0;
    const double to = Ho ? p[BO::to2] : p[BO::to1];
    const double dJ_dv1 = Hv ? -a1 * a2 / tfi : 0.0;
    const double dJ_dw1 = Hw ? -s1 / tsi : 0.0;
    const double dJ_ds1 = Hw ? -w1 / tsi : 0.0;
    const double dJ_du = (Hv ? -v1 * (a2 - a1) / tfi : 0.0) + (Hw ? -dtso / (tso * tso) : 1.0 / to);
    const double cv = lv1 + lJ * dJ_dv1, cw = lw1 + lJ * dJ_dw1, cs = ls1 + lJ * dJ_ds1;
    CellAdj r;
    r.lv = cv * f.Ev;
    r.lw = cw * f.Ew;
    r.ls = cs * f.Es;
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
// Adjoint of step n from its slot (state and factors of step n).  mu = dLoss/dV^{n+1}
// (with the probe seed of V^{n+1}), ly1 = dLoss/dy^{n+1}; writes lV = dLoss/dV^n (plus the
// seed of V^n from seed_row when not null) and ly = dLoss/dy^n, adds into gk and into the
// per-block parameter partials.  Grid-stride over cells with AG blocks (block b always
// accumulates partial b).
// IMPL: mu is lambda = M^-1 dLoss/dV^{n+1}; the diffusion term is inside the solve, and
// the kappa gradient takes V^{n+1} (the next slot's V plane, vnext).
template <bool IMPL>
__global__ __launch_bounds__(ABS) void k_adj_backward(AdjGeom G, double gxh, double gyh,
                                                      double gzh, const double* __restrict__ sl,
                                                      const double* __restrict__ vnext,
                                                      const double* __restrict__ mu,
                                                      const double* __restrict__ ly1,
                                                      double* __restrict__ lV,
                                                      double* __restrict__ ly,
                                                      const int* __restrict__ probe_slot,
                                                      const double* __restrict__ seed_row,
                                                      double* __restrict__ gk,
                                                      double* __restrict__ gtheta)
{
    const HexGeom& g = G.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const double* __restrict__ p = G.prm;
    const double a = G.dt / G.chiCm;
    const double lJ_per_mu = -G.dt * BO::ISCALE;
    const double du_dV = 1000.0 / BO::VSCALE;
    const double* __restrict__ Vn = sl + SL_V * n;
    double acc[2] = {0.0, 0.0};
    for (int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x; c < n; c += (int64_t)gridDim.x * ABS) {
        const IJK q = ijk_of((uint32_t)c, g);
        const Nbr b = nbr_of(c, q, g);
        const double muc = mu[c];
        const double vc = Vn[c];
        const double u = (vc * 1000.0 + BO::VOFF) / BO::VSCALE;
        BoFactors f;
        f.Ev = sl[SL_Ev * n + c];
        f.Ew = sl[SL_Ew * n + c];
        f.Es = sl[SL_Es * n + c];
        f.thw = sl[SL_thw * n + c];
        f.ths = sl[SL_ths * n + c];
        f.thso = sl[SL_thso * n + c];
        const CellAdj r = bo_vjp(u, sl[SL_v * n + c], sl[SL_w * n + c], sl[SL_s * n + c], f,
                                 G.dt_model, p, ly1[c], ly1[n + c], ly1[2 * n + c],
                                 lJ_per_mu * muc);
        ly[c] = r.lv;
        ly[n + c] = r.lw;
        ly[2 * n + c] = r.ls;
        acc[0] += r.dtfi;
        acc[1] += r.dtsi;
        // diffusion adjoint (L symmetric) and the kappa gradient, one pass over the faces
        const double mxp = mu[b.xp], mxm = mu[b.xm], myp = mu[b.yp], mym = mu[b.ym];
        const double mzp = mu[b.zp], mzm = mu[b.zm];
        double lv = muc + du_dV * r.lu;
        if constexpr (!IMPL) {
            const double fxp = G.fx[c], fxm = G.fx[b.xm], fyp = G.fy[c], fym = G.fy[b.ym];
            const double fzp = G.fz[c], fzm = G.fz[b.zm];
            const double Lmu = fxp * (mxp - muc) + fxm * (mxm - muc) + fyp * (myp - muc) +
                               fym * (mym - muc) + fzp * (mzp - muc) + fzm * (mzm - muc);
            lv += a * Lmu;
        }
        if (seed_row) {
            const int slot = probe_slot[c];
            if (slot >= 0) lv += seed_row[slot];
        }
        lV[c] = lv;
        const double* __restrict__ Vx = IMPL ? vnext : Vn;  // the field L acted on
        const double xc = Vx[c];
        const double dk = gxh * ((Vx[b.xp] - xc) * (muc - mxp) + (Vx[b.xm] - xc) * (muc - mxm)) +
                          gyh * ((Vx[b.yp] - xc) * (muc - myp) + (Vx[b.ym] - xc) * (muc - mym)) +
                          gzh * ((Vx[b.zp] - xc) * (muc - mzp) + (Vx[b.zm] - xc) * (muc - mzm));
        gk[c] += a * dk;
    }
    block_sum_a<2>(acc);
    if (threadIdx.x == 0) {
        gtheta[blockIdx.x * 2 + 0] += acc[0];
        gtheta[blockIdx.x * 2 + 1] += acc[1];
    }
}

// ---------------------------------------------------------- implicit solves
// Jacobi PCG on M = 1 - a L with the adjoint's face coefficients (kappa-scaled, wall faces
// 0, wall neighbours clamped to the cell as in hex_laplacian).  p and q are updated as
// p = z + beta p, q = M z + beta q (q = M p without a second stencil pass), so an
// iteration is two kernels; the host sums the per-block partials in a fixed order.
CG_DEV double m_apply(const AdjGeom& G, double a, const double* __restrict__ x, int64_t c,
                      const IJK& q)
{
    const Nbr b = nbr_of(c, q, G.g);
    const double xc = x[c];
    const double L = G.fx[c] * (x[b.xp] - xc) + G.fx[b.xm] * (x[b.xm] - xc) +
                     G.fy[c] * (x[b.yp] - xc) + G.fy[b.ym] * (x[b.ym] - xc) +
                     G.fz[c] * (x[b.zp] - xc) + G.fz[b.zm] * (x[b.zm] - xc);
    return xc - a * L;
}

// 1/diag of M
__global__ __launch_bounds__(ABS) void k_adj_dinv(AdjGeom G, double a, double* __restrict__ dinv)
{
    const HexGeom& g = G.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c >= n) return;
    const IJK q = ijk_of((uint32_t)c, g);
    const Nbr b = nbr_of(c, q, g);
    const double s = G.fx[c] + (q.i > 0 ? G.fx[b.xm] : 0.0) + G.fy[c] +
                     (q.j > 0 ? G.fy[b.ym] : 0.0) + G.fz[c] + (q.k > 0 ? G.fz[b.zm] : 0.0);
    dinv[c] = 1.0 / (1.0 + a * s);
}

// r = b - M x, z = D^-1 r; partials (r, z), (r, r), (b, b)
__global__ __launch_bounds__(ABS) void k_adj_cg0(AdjGeom G, double a, const double* __restrict__ bv,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ dinv,
                                                 double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ part)
{
    const HexGeom& g = G.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x; c < n; c += (int64_t)gridDim.x * ABS) {
        const IJK q = ijk_of((uint32_t)c, g);
        const double rc = bv[c] - m_apply(G, a, x, c, q);
        const double zc = dinv[c] * rc;
        r[c] = rc;
        z[c] = zc;
        acc[0] += rc * zc;
        acc[1] += rc * rc;
        acc[2] += bv[c] * bv[c];
    }
    block_sum_a<3>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[k * AG + blockIdx.x] = acc[k];
}

// p = z + beta p, q = M z + beta q (first: p = z, q = M z); partial (p, q)
__global__ __launch_bounds__(ABS) void k_adj_cg1(AdjGeom G, double a, double beta, int first,
                                                 const double* __restrict__ z,
                                                 double* __restrict__ p, double* __restrict__ qv,
                                                 double* __restrict__ part)
{
    const HexGeom& g = G.g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    double acc[1] = {0.0};
    for (int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x; c < n; c += (int64_t)gridDim.x * ABS) {
        const IJK q = ijk_of((uint32_t)c, g);
        const double mz = m_apply(G, a, z, c, q);
        const double pc = first ? z[c] : z[c] + beta * p[c];
        const double qc = first ? mz : mz + beta * qv[c];
        p[c] = pc;
        qv[c] = qc;
        acc[0] += pc * qc;
    }
    block_sum_a<1>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// x += alpha p, r -= alpha q, z = D^-1 r; partials (r, z), (r, r)
__global__ __launch_bounds__(ABS) void k_adj_cg2(int64_t n, double alpha,
                                                 const double* __restrict__ p,
                                                 const double* __restrict__ qv,
                                                 const double* __restrict__ dinv,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ z, double* __restrict__ part)
{
    double acc[2] = {0.0, 0.0};
    for (int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x; c < n; c += (int64_t)gridDim.x * ABS) {
        x[c] += alpha * p[c];
        const double rc = r[c] - alpha * qv[c];
        const double zc = dinv[c] * rc;
        r[c] = rc;
        z[c] = zc;
        acc[0] += rc * zc;
        acc[1] += rc * rc;
    }
    block_sum_a<2>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < 2; ++k) part[k * AG + blockIdx.x] = acc[k];
}

// Vm at the probes from a solved V plane
__global__ __launch_bounds__(ABS) void k_adj_probe(int64_t n, const double* __restrict__ v,
                                                   const int* __restrict__ probe_slot,
                                                   double* __restrict__ probe_row)
{
    const int64_t c = (int64_t)blockIdx.x * ABS + threadIdx.x;
    if (c >= n) return;
    const int slot = probe_slot[c];
    if (slot >= 0) probe_row[slot] = v[c];
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
    double* d_prm = nullptr;    // model constants [NP]
    BoxTable* d_boxes = nullptr;
    uint8_t* d_mask = nullptr;  // copy of the tissue's box-membership mask, or null
    int* d_slot = nullptr;      // [n] probe slot or -1
    double* d_V0 = nullptr;     // initial state
    double* d_y0 = nullptr;     // [3][n]
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double* ck = nullptr;       // checkpoints [n_ck][4][n]
    int n_ck = 0;
    double* seg = nullptr;      // state store [slots][SL_N][n]
    int seg_slots = 0;
    bool store_all = false;     // last forward kept every step's slot (no recompute)
    double *lV[2] = {nullptr, nullptr}, *ly[2] = {nullptr, nullptr};
    double* gtheta = nullptr;   // [AG][2]
    bool implicit = false;      // backward-Euler diffusion (solves in both passes)
    double *cg_r = nullptr, *cg_z = nullptr, *cg_p = nullptr, *cg_q = nullptr;
    double *dinv = nullptr, *bvec = nullptr, *lam = nullptr, *part = nullptr;  // part [3][AG]
    std::vector<double> h_part;
    int last_steps = 0;
    std::vector<double> t0;     // start time of every step of the last forward
    hipStream_t stream;
};

void adjoint_destroy(Adjoint* A)
{
    if (!A) return;
    void* p[] = {A->d_prm, A->d_boxes, A->d_slot, A->d_V0,  A->d_y0,  A->fx,   A->fy,
                 A->fz,    A->ck,      A->seg,    A->lV[0], A->lV[1], A->ly[0], A->ly[1],
                 A->gtheta, A->d_mask, A->cg_r,   A->cg_z,  A->cg_p,  A->cg_q,  A->dinv,
                 A->bvec,  A->lam,     A->part};
    for (void* q : p)
        if (q) hipFree(q);
    delete A;
}

int adjoint_create(Adjoint** out, const HexGeom& g, const double* sigma0_diag, double dt,
                   double chiCm, const ModelParams& P, const double* V0, const void* y0_states,
                   const ActiveBoxesAll& boxes, int K, int n_probes, const int64_t* probes,
                   bool implicit, hipStream_t stream, std::string* err)
{
    Adjoint* A = new Adjoint();
    A->implicit = implicit;
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
    A->stream = stream;
    const int64_t n = A->n;
    std::vector<int> slot((size_t)n, -1);
    for (int k = 0; k < n_probes; ++k) {
        if (probes[k] < 0 || probes[k] >= n) {
            delete A;
            *err = "probe cell out of range";
            return 1;
        }
        if (slot[probes[k]] >= 0) {
            delete A;
            *err = "probe cell listed twice";
            return 1;
        }
        slot[probes[k]] = k;
    }
    bool ok = true;
    auto M = [&](void** p, size_t b) { ok &= hipMalloc(p, b) == hipSuccess; };
    M((void**)&A->d_prm, sizeof(double) * MAX_PARAMS);
    M((void**)&A->d_boxes, sizeof(BoxTable));
    if (boxes.mask) M((void**)&A->d_mask, n);
    M((void**)&A->d_slot, sizeof(int) * n);
    M((void**)&A->d_V0, 8 * n);
    M((void**)&A->d_y0, 8 * 3 * n);
    M((void**)&A->fx, 8 * n); M((void**)&A->fy, 8 * n); M((void**)&A->fz, 8 * n);
    for (int i = 0; i < 2; ++i) { M((void**)&A->lV[i], 8 * n); M((void**)&A->ly[i], 8 * 3 * n); }
    M((void**)&A->gtheta, 8 * 2 * AG);
    if (implicit) {
        M((void**)&A->cg_r, 8 * n); M((void**)&A->cg_z, 8 * n); M((void**)&A->cg_p, 8 * n);
        M((void**)&A->cg_q, 8 * n); M((void**)&A->dinv, 8 * n); M((void**)&A->bvec, 8 * n);
        M((void**)&A->lam, 8 * n); M((void**)&A->part, 8 * 3 * AG);
        A->h_part.resize(3 * AG);
    }
    if (!ok) {
        adjoint_destroy(A);
        *err = "adjoint allocation failed";
        return 5;
    }
    BoxTable bt{};
    bt.n = boxes.n;
    for (int b = 0; b < boxes.n; ++b) {
        for (int d = 0; d < 3; ++d) { bt.lo[b][d] = boxes.lo[b][d]; bt.hi[b][d] = boxes.hi[b][d]; }
        bt.t0[b] = boxes.t0[b];
        bt.dur[b] = boxes.dur[b];
        bt.I[b] = boxes.I[b];
    }
    if (boxes.mask) hipMemcpy(A->d_mask, boxes.mask, n, hipMemcpyDeviceToDevice);
    bt.mask = A->d_mask;
    hipMemcpy(A->d_boxes, &bt, sizeof bt, hipMemcpyHostToDevice);
    hipMemcpy(A->d_slot, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    hipMemcpy(A->d_V0, V0, 8 * n, hipMemcpyDeviceToDevice);
    // gates v, w, s: state planes 1..3 of the ionic handle
    hipMemcpy(A->d_y0, (const double*)y0_states + n, 8 * 3 * n, hipMemcpyDeviceToDevice);
    *out = A;
    return 0;
}

static AdjGeom geom_of(const Adjoint* A)
{
    AdjGeom G;
    G.g = A->g;
    G.fx = A->fx;
    G.fy = A->fy;
    G.fz = A->fz;
    G.prm = A->d_prm;
    G.boxes = A->d_boxes;
    G.dt = A->dt;
    G.dt_model = A->dt_model;
    G.chiCm = A->chiCm;
    return G;
}

// the first k partial sets [k][AG] summed in a fixed order (host)
static bool read_partials(Adjoint* A, int k, double* out)
{
    if (hipMemcpyAsync(A->h_part.data(), A->part, 8 * (size_t)k * AG, hipMemcpyDeviceToHost,
                       A->stream) != hipSuccess ||
        hipStreamSynchronize(A->stream) != hipSuccess)
        return false;
    for (int j = 0; j < k; ++j) {
        double t = 0.0;
        for (int b = 0; b < AG; ++b) t += A->h_part[(size_t)j * AG + b];
        out[j] = t;
    }
    return true;
}

// x = M^-1 b by Jacobi PCG from the x given, to ||r|| <= 1e-14 ||b||
static int adj_solve(Adjoint* A, const AdjGeom& G, const double* b, double* x,
                     std::string* err)
{
    constexpr double TOL = 1e-14;
    constexpr int MAX_IT = 1000;
    const double a = A->dt / A->chiCm;
    double s3[3];
    hipLaunchKernelGGL(k_adj_cg0, dim3(AG), dim3(ABS), 0, A->stream, G, a, b,
                       (const double*)x, (const double*)A->dinv, A->cg_r, A->cg_z, A->part);
    if (!read_partials(A, 3, s3)) {
        *err = "adjoint solve: partial read-back failed";
        return 4;
    }
    double rz = s3[0], rr = s3[1], beta = 0.0;
    const double lim = TOL * TOL * s3[2];
    for (int it = 0; rr > lim; ++it) {
        if (it >= MAX_IT) {
            *err = "adjoint solve did not converge";
            return 4;
        }
        double pq, s2[2];
        hipLaunchKernelGGL(k_adj_cg1, dim3(AG), dim3(ABS), 0, A->stream, G, a, beta,
                           it == 0 ? 1 : 0, (const double*)A->cg_z, A->cg_p, A->cg_q, A->part);
        if (!read_partials(A, 1, &pq)) {
            *err = "adjoint solve: partial read-back failed";
            return 4;
        }
        const double alpha = rz / pq;
        hipLaunchKernelGGL(k_adj_cg2, dim3(AG), dim3(ABS), 0, A->stream, A->n, alpha,
                           (const double*)A->cg_p, (const double*)A->cg_q,
                           (const double*)A->dinv, x, A->cg_r, A->cg_z, A->part);
        if (!read_partials(A, 2, s2)) {
            *err = "adjoint solve: partial read-back failed";
            return 4;
        }
        beta = s2[0] / rz;
        rz = s2[0];
        rr = s2[1];
    }
    return 0;
}

// one forward step from slot `in` into slot `outp` (factors into aux, probes into
// probe_row when not null)
static int adj_step(Adjoint* A, const AdjGeom& G, double t0, const double* in, double* outp,
                    double* aux, double* probe_row, std::string* err)
{
    const unsigned nb = blocks_of(A->n);
    if (!A->implicit) {
        hipLaunchKernelGGL(k_adj_forward<false>, dim3(nb), dim3(ABS), 0, A->stream, G, t0, in,
                           outp, aux, (const int*)A->d_slot, probe_row, (double*)nullptr);
        return 0;
    }
    hipLaunchKernelGGL(k_adj_forward<true>, dim3(nb), dim3(ABS), 0, A->stream, G, t0, in,
                       outp, aux, (const int*)A->d_slot, (double*)nullptr, A->bvec);
    if (int rc = adj_solve(A, G, A->bvec, outp + SL_V * A->n, err)) return rc;
    if (probe_row)
        hipLaunchKernelGGL(k_adj_probe, dim3(nb), dim3(ABS), 0, A->stream, A->n,
                           (const double*)(outp + SL_V * A->n), (const int*)A->d_slot,
                           probe_row);
    return 0;
}

static bool ensure(double** buf, int* have, int want, size_t per)
{
    if (want <= *have) return true;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    if (hipMalloc((void**)buf, per * (size_t)want) != hipSuccess) return false;
    *have = want;
    return true;
}

int adjoint_forward(Adjoint* A, const double* kappa, const double* theta, const int* theta_index,
                    int n_theta, int n_steps, const std::vector<double>& t0, double* probes,
                    size_t store_budget, std::string* err)
{
    const int64_t n = A->n;
    const unsigned nb = blocks_of(n);
    const size_t slot_bytes = 8 * (size_t)SL_N * n;
    for (int k = 0; k < n_theta; ++k) A->P.p[theta_index[k]] = theta[k];
    // every step's slot kept when the interval covers the run and the store fits the budget
    A->store_all = A->K >= n_steps && slot_bytes * (size_t)(n_steps + 1) <= store_budget;
    const int slots = A->store_all ? n_steps + 1 : A->K + 1;
    if (!ensure(&A->seg, &A->seg_slots, slots, slot_bytes)) {
        *err = "state store allocation failed";
        return 5;
    }
    const int n_ck = A->store_all ? 0 : (n_steps + A->K - 1) / A->K;
    if (!ensure(&A->ck, &A->n_ck, n_ck, 8 * 4 * (size_t)n)) {
        *err = "checkpoint allocation failed";
        return 5;
    }
    hipMemcpyAsync(A->d_prm, A->P.p, sizeof(double) * MAX_PARAMS, hipMemcpyHostToDevice,
                   A->stream);
    hipLaunchKernelGGL(k_adj_faces, dim3(nb), dim3(ABS), 0, A->stream, A->g, kappa, A->gx,
                       A->gy, A->gz, A->fx, A->fy, A->fz);
    if (A->implicit)
        hipLaunchKernelGGL(k_adj_dinv, dim3(nb), dim3(ABS), 0, A->stream, geom_of(A),
                           A->dt / A->chiCm, A->dinv);
    hipMemcpyAsync(A->seg + SL_V * n, A->d_V0, 8 * n, hipMemcpyDeviceToDevice, A->stream);
    hipMemcpyAsync(A->seg + SL_v * n, A->d_y0, 8 * 3 * n, hipMemcpyDeviceToDevice, A->stream);
    const AdjGeom G = geom_of(A);
    for (int st = 0; st < n_steps; ++st) {
        double* in;
        double* outp;
        double* aux = nullptr;
        if (A->store_all) {
            in = A->seg + (size_t)st * SL_N * n;
            outp = in + (size_t)SL_N * n;
            aux = in + SL_Ev * n;
        } else {  // two slots in turn, a checkpoint every K steps
            in = A->seg + (size_t)(st & 1) * SL_N * n;
            outp = A->seg + (size_t)((st + 1) & 1) * SL_N * n;
            if (st % A->K == 0)
                hipMemcpyAsync(A->ck + (size_t)(st / A->K) * 4 * n, in, 8 * 4 * n,
                               hipMemcpyDeviceToDevice, A->stream);
        }
        if (int rc = adj_step(A, G, t0[st], in, outp, aux, probes + (size_t)st * A->n_probes,
                              err))
            return rc;
    }
    A->last_steps = n_steps;
    A->t0 = t0;
    const hipError_t e = hipStreamSynchronize(A->stream);
    const hipError_t e2 = hipGetLastError();
    if (e != hipSuccess || e2 != hipSuccess) {
        *err = std::string("adjoint forward: ") + hipGetErrorString(e != hipSuccess ? e : e2);
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
    const AdjGeom G = geom_of(A);
    const double gxh = 0.5 * A->gx, gyh = 0.5 * A->gy, gzh = 0.5 * A->gz;
    const size_t SN = (size_t)SL_N * n;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(ABS), 0, A->stream, n, dkappa, 0.0);
    hipLaunchKernelGGL(k_fill, dim3((2 * AG + ABS - 1) / ABS), dim3(ABS), 0, A->stream,
                       (int64_t)2 * AG, A->gtheta, 0.0);
    // lV = dLoss/dV^N from the last probe row
    int cur = 0;
    hipLaunchKernelGGL(k_adj_seed, dim3(nb), dim3(ABS), 0, A->stream, n,
                       (const int*)A->d_slot, dprobes + (size_t)(N - 1) * A->n_probes,
                       A->lV[cur], A->ly[cur]);
    const int K = A->store_all ? N : A->K;
    const int n_seg = (N + K - 1) / K;
    for (int sgi = n_seg - 1; sgi >= 0; --sgi) {
        const int s0 = sgi * K, s1 = s0 + K < N ? s0 + K : N;
        const double* base = A->store_all ? A->seg + (size_t)s0 * SN : A->seg;
        if (!A->store_all) {
            // recompute the segment: slot j <- state at step s0 + j, with the factors
            hipMemcpyAsync(A->seg, A->ck + (size_t)sgi * 4 * n, 8 * 4 * n,
                           hipMemcpyDeviceToDevice, A->stream);
            for (int st = s0; st < s1; ++st) {
                double* in = A->seg + (size_t)(st - s0) * SN;
                if (int rc = adj_step(A, G, A->t0[st], in, in + SN, in + SL_Ev * n, nullptr,
                                      err))
                    return rc;
            }
        }
        for (int st = s1 - 1; st >= s0; --st) {
            const double* sl = base + (size_t)(st - s0) * SN;
            const int nxt = cur ^ 1;
            const double* seed = st >= 1 ? dprobes + (size_t)(st - 1) * A->n_probes : nullptr;
            if (A->implicit) {
                // lambda = M^-1 dLoss/dV^{n+1}, from lambda = 0
                hipLaunchKernelGGL(k_fill, dim3(nb), dim3(ABS), 0, A->stream, n, A->lam, 0.0);
                if (int rc = adj_solve(A, G, A->lV[cur], A->lam, err)) return rc;
                hipLaunchKernelGGL(k_adj_backward<true>, dim3(AG), dim3(ABS), 0, A->stream, G,
                                   gxh, gyh, gzh, sl, sl + SN, (const double*)A->lam,
                                   (const double*)A->ly[cur], A->lV[nxt], A->ly[nxt],
                                   (const int*)A->d_slot, seed, dkappa, A->gtheta);
            } else {
                hipLaunchKernelGGL(k_adj_backward<false>, dim3(AG), dim3(ABS), 0, A->stream, G,
                                   gxh, gyh, gzh, sl, (const double*)nullptr,
                                   (const double*)A->lV[cur], (const double*)A->ly[cur],
                                   A->lV[nxt], A->ly[nxt], (const int*)A->d_slot, seed, dkappa,
                                   A->gtheta);
            }
            cur = nxt;
        }
    }
    std::vector<double> gt(2 * AG);
    hipError_t e = hipMemcpyAsync(gt.data(), A->gtheta, 8 * 2 * AG, hipMemcpyDeviceToHost,
                                  A->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(A->stream);
    const hipError_t e2 = hipGetLastError();
    if (e != hipSuccess || e2 != hipSuccess) {
        *err = std::string("adjoint backward: ") + hipGetErrorString(e != hipSuccess ? e : e2);
        return 4;
    }
    double a0 = 0.0, a1 = 0.0;
    for (int b = 0; b < AG; ++b) {
        a0 += gt[2 * b];
        a1 += gt[2 * b + 1];
    }
    dtheta[0] = a0;
    dtheta[1] = a1;
    return 0;
}

} // namespace cg
