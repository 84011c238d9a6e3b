/* Tangent-linear (forward-mode) run of the BuenoOrovio monodomain step.  Test
 * infrastructure: the FP64 reference gradient that the device adjoint is checked against
 * over a whole horizon, where finite differences fail (the model's Heaviside switches make
 * the discrete loss jump wherever a perturbation moves a threshold crossing by a step).
 *
 * Configuration: the one the adjoint differentiates, from rest (V = -84 mV,
 * myocardiumDomain.C:207; the model's initial gates): explicit diffusion with Euler ddt
 * (orc_tissue_step with ddt 0, monodomainSolver.C:103-118), a diagonal conductivity
 * sigma_c = kappa_c sig0 per cell, BuenoOrovio's Rush-Larsen step with V fed by the PDE
 * (orc_ionic_tissue_step, orc_rl_step; BuenoOrovio_2008.H:116-225), box stimulus
 * (orc_box_stimulus).  Loss: the sum over steps and probes of (Vm - target)^2, Vm after
 * each step.
 *
 * Every quantity is a dual number: its value repeats the oracle's operations in the
 * oracle's order, so the value trajectory is orc_tissue_step's bit for bit (checked by the
 * tests), and its K tangent parts carry d/d eps along K directions of (kappa, tau_fi,
 * tau_si) at once.  Heaviside switches act on the value; their derivative is 0, the
 * convention of the adjoint (it differentiates the branch each step takes).  The tensor
 * correction of fvc::laplacian is exactly 0 for a diagonal conductivity and is omitted. */
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" {
#include "oracle.h"
}
#include "bo_consts.h"

namespace {

constexpr int KD = 8;   // tangent directions carried per pass

struct D {
    double v;
    double d[KD];
};

inline D cst(double x)
{
    D r;
    r.v = x;
    for (int k = 0; k < KD; ++k) r.d[k] = 0.0;
    return r;
}
inline D operator+(const D& a, const D& b)
{
    D r;
    r.v = a.v + b.v;
    for (int k = 0; k < KD; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline D operator-(const D& a, const D& b)
{
    D r;
    r.v = a.v - b.v;
    for (int k = 0; k < KD; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline D operator-(const D& a)
{
    D r;
    r.v = -a.v;
    for (int k = 0; k < KD; ++k) r.d[k] = -a.d[k];
    return r;
}
inline D operator*(const D& a, const D& b)
{
    D r;
    r.v = a.v * b.v;
    for (int k = 0; k < KD; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline D operator/(const D& a, const D& b)
{
    D r;
    r.v = a.v / b.v;
    for (int k = 0; k < KD; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
inline D operator+(const D& a, double b) { D r = a; r.v = a.v + b; return r; }
inline D operator+(double a, const D& b) { D r = b; r.v = a + b.v; return r; }
inline D operator-(const D& a, double b) { D r = a; r.v = a.v - b; return r; }
inline D operator-(double a, const D& b)
{
    D r;
    r.v = a - b.v;
    for (int k = 0; k < KD; ++k) r.d[k] = -b.d[k];
    return r;
}
inline D operator*(const D& a, double b)
{
    D r;
    r.v = a.v * b;
    for (int k = 0; k < KD; ++k) r.d[k] = a.d[k] * b;
    return r;
}
inline D operator*(double a, const D& b) { return b * a; }
inline D operator/(const D& a, double b)
{
    D r;
    r.v = a.v / b;
    for (int k = 0; k < KD; ++k) r.d[k] = a.d[k] / b;
    return r;
}
inline D operator/(double a, const D& b)
{
    D r;
    r.v = a / b.v;
    for (int k = 0; k < KD; ++k) r.d[k] = -r.v * b.d[k] / b.v;
    return r;
}
inline D dexp(const D& a)
{
    D r;
    r.v = std::exp(a.v);
    for (int k = 0; k < KD; ++k) r.d[k] = r.v * a.d[k];
    return r;
}
inline D dtanh(const D& a)
{
    D r;
    r.v = std::tanh(a.v);
    const double s = 1.0 - r.v * r.v;
    for (int k = 0; k < KD; ++k) r.d[k] = s * a.d[k];
    return r;
}
// heaviside.H:39-49 on the value
inline double H(double x) { return x > 0.0 ? 1.0 : 0.0; }

// BuenoOrovio constants: tau_fi and tau_si carry tangents, the rest are plain
struct Consts {
    double c[BO_NC];
    D tfi, tsi;
};

// orc_rl_step (tissue mode, V frozen) then the step-end Jion, in bo_eval's order
inline D bo_step(const Consts& K, D* y, double dt)
{
    const double* c = K.c;
    {
        const D u = y[0];
        const double Hv = H(u.v - c[BO_thv]), Hw = H(u.v - c[BO_thw]);
        const double Ho = H(u.v - c[BO_tho]), Hvm = H(u.v - c[BO_thvm]);
        const double tvm = (1.0 - Hvm) * c[BO_tv1m] + Hvm * c[BO_tv2m];
        const D twm = c[BO_tw1m] + (c[BO_tw2m] - c[BO_tw1m]) *
                                       (1.0 + dtanh(c[BO_kwm] * (u - c[BO_uwm]))) / 2.0;
        const double tS = (1.0 - Hw) * c[BO_ts1] + Hw * c[BO_ts2];
        const double vinf = H(c[BO_thvm] - u.v);
        const D winf = (1.0 - Ho) * (1.0 - u / c[BO_twinf]) + Ho * c[BO_winfstar];
        const D sinf = (1.0 + dtanh(c[BO_ks] * (u - c[BO_us]))) / 2.0;
        // xi + (y - xi) exp(-dt / tau) for v, w, s
        {
            const double xi = Hv > 0 ? 0.0 : vinf, ta = Hv > 0 ? c[BO_tvp] : tvm;
            y[1] = xi + (y[1] - xi) * std::exp(-dt / ta);
        }
        if (Hw > 0) {
            y[2] = 0.0 + (y[2] - 0.0) * std::exp(-dt / c[BO_twp]);
        } else {
            y[2] = winf + (y[2] - winf) * dexp(-dt / twm);
        }
        y[3] = sinf + (y[3] - sinf) * std::exp(-dt / tS);
    }
    // computeVariables at the step end
    const D u = y[0], v = y[1], w = y[2], s = y[3];
    const double Hv = H(u.v - c[BO_thv]), Hw = H(u.v - c[BO_thw]);
    const double Ho = H(u.v - c[BO_tho]);
    const D tso = c[BO_tso1] + (c[BO_tso2] - c[BO_tso1]) *
                                   (1.0 + dtanh(c[BO_kso] * (u - c[BO_uso]))) / 2.0;
    const double to = (1.0 - Ho) * c[BO_to1] + Ho * c[BO_to2];
    const D Jfi = -v * Hv * (u - c[BO_thv]) * (c[BO_uu] - u) / K.tfi;
    const D Jso = (u - c[BO_uo]) * (1.0 - Hw) / to + Hw / tso;
    const D Jsi = -Hw * w * s / K.tsi;
    return Jfi + Jso + Jsi;
}

} // namespace

extern "C" int orc_bo_tangent(int nx, int ny, int nz, double h, const double* sig0,
                              int tissue, double dt, double chi, double cm,
                              const orc_boxes_t* boxes, const double* kappa,
                              const double* theta, int n_dir, const double* dkappa,
                              const double* dtheta, long n_steps, int n_probes,
                              const long* probes, const double* target, int nthreads,
                              double* loss, double* dloss, double* v_final)
{
    if (n_dir < 0 || n_dir > KD || nx < 1 || ny < 1 || nz < 1) return -1;
    orc_mesh_t mesh;
    if (orc_mesh_build(&mesh, nx, ny, nz, h, h, h)) return -2;
    const long n = mesh.n_cells;
    for (int p = 0; p < n_probes; ++p)
        if (probes[p] < 0 || probes[p] >= n) {
            orc_mesh_free(&mesh);
            return -3;
        }
    Consts K;
    double call[ORC_MAX_CONSTS], y0[ORC_MAX_STATES];
    if (orc_model_init(ORC_BO, tissue, call, y0)) {
        orc_mesh_free(&mesh);
        return -4;
    }
    std::memcpy(K.c, call, sizeof K.c);
    K.c[BO_tfi] = theta[0];
    K.c[BO_tsi] = theta[1];
    K.tfi = cst(theta[0]);
    K.tsi = cst(theta[1]);
    for (int k = 0; k < n_dir; ++k) {
        K.tfi.d[k] = dtheta[2 * k];
        K.tsi.d[k] = dtheta[2 * k + 1];
    }
    // per-cell conductivity diagonal, sigma_c = kappa_c sig0 (as the tests build sigma9)
    std::vector<D> sg(3 * (size_t)n);
    for (long c = 0; c < n; ++c)
        for (int d = 0; d < 3; ++d) {
            D s = cst(kappa[c] * sig0[d]);
            for (int k = 0; k < n_dir; ++k) s.d[k] = dkappa[(size_t)k * n + c] * sig0[d];
            sg[3 * c + d] = s;
        }
    std::vector<D> V(n, cst(-0.084)), V0(n, cst(0.0)), Iion(n, cst(0.0)), lap(n, cst(0.0));
    std::vector<D> st(3 * (size_t)n);
    for (long c = 0; c < n; ++c)
        for (int q = 0; q < 3; ++q) st[3 * c + q] = cst(y0[q + 1]);
    std::vector<double> src(n);
    const double chiCm = chi * cm, dtm = dt * 1000.0, vol = h * h * h;
    const double A = h * h;
    double t = 0.0;
    D L = cst(0.0);
    const int nt = nthreads > 0 ? nthreads : 1;
    for (long step = 0; step < n_steps; ++step) {
        t = t + dt;
        const double t0 = t - dt;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long c = 0; c < n; ++c) V0[c] = V[c];
        orc_box_stimulus(&mesh, boxes, t0, src.data());
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long c = 0; c < n; ++c) {
            D y[4];
            y[0] = (V[c] * 1000.0 + 84.0) / 85.7;
            for (int q = 0; q < 3; ++q) y[q + 1] = st[3 * c + q];
            Iion[c] = bo_step(K, y, dtm) * 85.7;
            for (int q = 0; q < 3; ++q) st[3 * c + q] = y[q + 1];
        }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long c = 0; c < n; ++c) {
            const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
            const long stp[3] = {1, nx, (long)nx * ny};
            const int idx[3] = {i, j, k}, nn[3] = {nx, ny, nz};
            D acc = cst(0.0);
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    const int interior = side == 0 ? idx[d] < nn[d] - 1 : idx[d] > 0;
                    if (!interior) continue;
                    const long nb = side == 0 ? c + stp[d] : c - stp[d];
                    const D sf = 0.5 * (sg[3 * c + d] + sg[3 * nb + d]);
                    acc = acc + A * sf * (V0[nb] - V0[c]) / h;
                }
            lap[c] = acc / vol;
        }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long c = 0; c < n; ++c)
            V[c] = V0[c] + dt / chiCm * (lap[c] - chiCm * Iion[c] + src[c]);
        for (int p = 0; p < n_probes; ++p) {
            const D e = V[probes[p]] - target[step * n_probes + p];
            L = L + e * e;
        }
    }
    *loss = L.v;
    for (int k = 0; k < n_dir; ++k) dloss[k] = L.d[k];
    if (v_final)
        for (long c = 0; c < n; ++c) v_final[c] = V[c].v;
    orc_mesh_free(&mesh);
    return 0;
}
