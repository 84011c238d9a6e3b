// Implicit monodomain step on the hex block.
//
// Matrix (per unit volume, monodomainSolver.C:129-135 with fvm::laplacian):
//   (c chiCm/dt + sum_f a_f) V_P - sum_f a_f V_N = rhs,   a_f = sigma_nn,f / h^2
// stored as three face-coefficient planes fx, fy, fz (face to +x/+y/+z, 0 on the wall)
// plus 1/diag; the operator is applied as a 7-point stencil (no index arrays):
//   bytes per row and product: 3 coefficients + x + y  (fp64 40 B, fp32 20 B).
// Right-hand side (computed in the ionic kernel, so Iion never makes an extra pass):
//   rhs = chiCm/dt (c0 V^n - c00 V^{n-1}) + div(corr) - chiCm Iion + source
// Solver: Jacobi-preconditioned CG.  fp64 path: CG in fp64 to ||b-Ax||/||b|| <= tol.
// mixed path: iterative refinement -- residual and solution update in fp64, correction
// from an fp32 CG (fp32 operator and vectors, fp64 dot products).
// Reductions are deterministic: per-block partials, the last block to finish sums them
// in a fixed order and updates the scalars of the iteration (alpha, beta, convergence).
#include <cstring>
#include <vector>
#include "cg_implicit.hpp"

namespace cg {

constexpr int BS = 256;
constexpr int NWAVE = BS / 64;

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD owns a
// contiguous run of cells and the +-nx, +-nx*ny neighbour rows stay in its own L2.
CG_DEV int64_t xcd_block(int64_t b, int64_t nb)
{
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

enum { S_RZ, S_RZ_OLD, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TOL2, S_R0, S_NSCAL };
struct DevState {
    double s[S_NSCAL];
    int iters;
    int done;
    int max_iter;
    int pad;
    unsigned counter[4];
};

template <int NV>
CG_DEV bool reduce_last(double (&v)[NV], double* __restrict__ partials, unsigned* counter,
                        int nblocks, int64_t bid, double (&out)[NV])
{
    __shared__ double sm[NV][NWAVE];
    __shared__ int last;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sm[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int q = 0; q < NWAVE; ++q) s += sm[k][q];
            partials[k * nblocks + bid] = s;
        }
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = prev == (unsigned)(nblocks - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int q = threadIdx.x; q < nblocks; q += BS) s += partials[k * nblocks + q];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) sm[k][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int q = 0; q < NWAVE; ++q) s += sm[k][q];
        out[k] = s;
    }
    if (threadIdx.x == 0) atomicExch(counter, 0u);
    return true;
}

template <class T> struct HexOp {
    const T* fx;
    const T* fy;
    const T* fz;
    const T* invd;
    T mass;  // c chiCm / dt
};

template <class T, class TX>
CG_DEV T apply_row(const HexOp<T>& A, const TX* __restrict__ x, int64_t c, int i, int j, int k,
                   const HexGeom& g)
{
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const T xc = T(x[c]);
    T diag = A.mass, off = T(0);
    if (i < g.nx - 1) { const T a = A.fx[c]; diag += a; off += a * T(x[c + 1]); }
    if (i > 0) { const T a = A.fx[c - 1]; diag += a; off += a * T(x[c - 1]); }
    if (j < g.ny - 1) { const T a = A.fy[c]; diag += a; off += a * T(x[c + sy]); }
    if (j > 0) { const T a = A.fy[c - sy]; diag += a; off += a * T(x[c - sy]); }
    if (k < g.nz - 1) { const T a = A.fz[c]; diag += a; off += a * T(x[c + sz]); }
    if (k > 0) { const T a = A.fz[c - sz]; diag += a; off += a * T(x[c - sz]); }
    return diag * xc - off;
}

#define CELL_INDEX                                                                         \
    const int64_t nb_ = gridDim.x;                                                         \
    const int64_t bid = xcd_block(blockIdx.x, nb_);                                        \
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;                                         \
    const int64_t c = bid * BS + threadIdx.x;                                              \
    const bool act = c < n;                                                                \
    const int ci = act ? (int)(c % g.nx) : 0;                                              \
    const int cj = act ? (int)((c / g.nx) % g.ny) : 0;                                     \
    const int ck = act ? (int)(c / ((int64_t)g.nx * g.ny)) : 0;

// ---- 1/diag for a given mass coefficient
template <class T>
__global__ __launch_bounds__(BS) void k_invdiag(HexGeom g, HexOp<T> A, T* __restrict__ invd)
{
    CELL_INDEX
    if (!act) return;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    T d = A.mass;
    if (ci < g.nx - 1) d += A.fx[c];
    if (ci > 0) d += A.fx[c - 1];
    if (cj < g.ny - 1) d += A.fy[c];
    if (cj > 0) d += A.fy[c - sy];
    if (ck < g.nz - 1) d += A.fz[c];
    if (ck > 0) d += A.fz[c - sz];
    invd[c] = T(1) / d;
}

// ---- ionic step + right-hand side
template <class M, class T>
__global__ __launch_bounds__(BS) void k_ionic_rhs(HexGeom g, T* __restrict__ st,
                                                  const double* __restrict__ Vn,
                                                  const double* __restrict__ Vm1,
                                                  double* __restrict__ Iion,
                                                  double* __restrict__ b, ExplicitStep s,
                                                  ModelParams P, int tissue)
{
    CELL_INDEX
    if (!act) return;
    T p[M::NP];
#pragma unroll
    for (int q = 0; q < M::NP; ++q) p[q] = T(P.p[q]);
    const T one[4] = {T(1), T(1), T(1), T(1)};
    T y[M::NS];
#pragma unroll
    for (int q = 1; q < M::NS; ++q) y[q] = st[q * n + c];
    const double vn = Vn[c];
    const T v = T((vn * 1000.0 + M::VOFF) / M::VSCALE);
    const T I = Ops<M>::tissue_step(v, y, T(s.dt_model), p, tissue, one);
#pragma unroll
    for (int q = 1; q < M::NS; ++q) st[q * n + c] = y[q];
    const double iion = double(I) * M::ISCALE;
    Iion[c] = iion;
    double src = 0.0;
    {
        const double p0 = (ci + 0.5) * g.hx, p1 = (cj + 0.5) * g.hy, p2 = (ck + 0.5) * g.hz;
        for (int q = 0; q < s.boxes.n; ++q)
            if (p0 >= s.boxes.lo[q][0] && p0 <= s.boxes.hi[q][0] && p1 >= s.boxes.lo[q][1] &&
                p1 <= s.boxes.hi[q][1] && p2 >= s.boxes.lo[q][2] && p2 <= s.boxes.hi[q][2])
                src += s.boxes.intensity[q];
    }
    double rhs = s.chiCm / s.dt * (s.c0 * vn - s.c00 * (s.c00 != 0.0 ? Vm1[c] : 0.0)) +
                 src - s.chiCm * iion;
    if (s.corr) rhs += s.corr[c];
    b[c] = rhs;
}

// ---- fp64 residual r = b - A x, with ||r||^2 and ||b||^2; also starts a CG in TP:
//      rt = r (TP), p = z = invd rt, rz = r.z.  init_state: set tolerance/counters.
template <class TP>
__global__ __launch_bounds__(BS) void k_residual(HexGeom g, HexOp<double> A, HexOp<TP> Ap,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ b,
                                                 TP* __restrict__ rt, TP* __restrict__ p,
                                                 TP* __restrict__ xt, double* partials,
                                                 DevState* S, int which)
{
    CELL_INDEX
    double v[3] = {0.0, 0.0, 0.0};
    if (act) {
        const double r = b[c] - apply_row(A, x, c, ci, cj, ck, g);
        const TP rr = TP(r);
        const TP z = Ap.invd[c] * rr;
        rt[c] = rr;
        p[c] = z;
        if (xt) xt[c] = TP(0);
        v[0] = r * r;
        v[1] = b[c] * b[c];
        v[2] = double(rr) * double(z);
    }
    double out[3];
    if (reduce_last<3>(v, partials, &S->counter[which], (int)nb_, bid, out) && threadIdx.x == 0) {
        S->s[S_RR] = out[0];
        S->s[S_BB] = out[1];
        S->s[S_RZ] = out[2];
        S->s[S_R0] = out[0];
    }
}

// ---- q = A p ; alpha = rz / (p.q)
template <class T>
__global__ __launch_bounds__(BS) void k_spmv_dot(HexGeom g, HexOp<T> A, const T* __restrict__ p,
                                                 T* __restrict__ q, double* partials,
                                                 DevState* S)
{
    if (S->done) return;
    CELL_INDEX
    double v[1] = {0.0};
    if (act) {
        const T y = apply_row(A, p, c, ci, cj, ck, g);
        q[c] = y;
        v[0] = double(p[c]) * double(y);
    }
    double out[1];
    if (reduce_last<1>(v, partials, &S->counter[0], (int)nb_, bid, out) && threadIdx.x == 0) {
        S->s[S_PQ] = out[0];
        S->s[S_ALPHA] = S->s[S_RZ] / out[0];
    }
}

// ---- x += alpha p ; r -= alpha q ; z = invd r ; rz' = r.z, rr = r.r ; beta ; convergence
template <class T, class TX>
__global__ __launch_bounds__(BS) void k_update(HexGeom g, const T* __restrict__ invd,
                                               TX* __restrict__ x, T* __restrict__ r,
                                               const T* __restrict__ p,
                                               const T* __restrict__ q, double* partials,
                                               DevState* S)
{
    if (S->done) return;
    CELL_INDEX
    const T alpha = T(S->s[S_ALPHA]);
    double v[2] = {0.0, 0.0};
    if (act) {
        x[c] = x[c] + TX(alpha * p[c]);
        const T rn = r[c] - alpha * q[c];
        r[c] = rn;
        const T z = invd[c] * rn;
        v[0] = double(rn) * double(z);
        v[1] = double(rn) * double(rn);
    }
    double out[2];
    if (reduce_last<2>(v, partials, &S->counter[1], (int)nb_, bid, out) && threadIdx.x == 0) {
        S->s[S_BETA] = out[0] / S->s[S_RZ];
        S->s[S_RZ_OLD] = S->s[S_RZ];
        S->s[S_RZ] = out[0];
        S->s[S_RR] = out[1];
        S->iters += 1;
        if (out[1] <= S->s[S_TOL2] || S->iters >= S->max_iter) S->done = 1;
    }
}

// ---- p = invd r + beta p
template <class T>
__global__ __launch_bounds__(BS) void k_pupdate(HexGeom g, const T* __restrict__ invd,
                                                const T* __restrict__ r, T* __restrict__ p,
                                                DevState* S)
{
    if (S->done) return;
    CELL_INDEX
    if (!act) return;
    const T beta = T(S->s[S_BETA]);
    p[c] = invd[c] * r[c] + beta * p[c];
}

// ---- x64 += e32
__global__ __launch_bounds__(BS) void k_add_correction(HexGeom g, double* __restrict__ x,
                                                       const float* __restrict__ e)
{
    CELL_INDEX
    if (!act) return;
    x[c] += double(e[c]);
}

// ---- activation time (myocardiumDomain.C:345-374)
__global__ __launch_bounds__(BS) void k_activation(HexGeom g, const double* __restrict__ Vn,
                                                   const double* __restrict__ Vnew,
                                                   uint8_t* __restrict__ flag,
                                                   double* __restrict__ tact, double t_old,
                                                   double dt)
{
    CELL_INDEX
    if (!act) return;
    const double vnew = Vnew[c];
    if (flag[c] && vnew > 1e-15) {
        const double vn = Vn[c];
        flag[c] = 0;
        tact[c] = t_old + (0.0 - vn) / (vnew - vn) * dt;
    }
}

// ---- explicit Laplacian (operator check): lap = -(K V) + corr
__global__ __launch_bounds__(BS) void k_laplacian(HexGeom g, HexCoefs a,
                                                  const double* __restrict__ V,
                                                  const double* __restrict__ corr,
                                                  double* __restrict__ out)
{
    CELL_INDEX
    if (!act) return;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const double vc = V[c];
    double acc = 0.0;
    const bool u = a.fx == nullptr;
    if (ci < g.nx - 1) acc += (u ? a.ax : a.fx[c]) * (V[c + 1] - vc);
    if (ci > 0) acc += (u ? a.ax : a.fx[c - 1]) * (V[c - 1] - vc);
    if (cj < g.ny - 1) acc += (u ? a.ay : a.fy[c]) * (V[c + sy] - vc);
    if (cj > 0) acc += (u ? a.ay : a.fy[c - sy]) * (V[c - sy] - vc);
    if (ck < g.nz - 1) acc += (u ? a.az : a.fz[c]) * (V[c + sz] - vc);
    if (ck > 0) acc += (u ? a.az : a.fz[c - sz]) * (V[c - sz] - vc);
    if (corr) acc += corr[c];
    out[c] = acc;
}

const char* kernel_name(int id)
{
    static const char* names[K_COUNT] = {"ionic", "spmv_dot", "cg_update", "p_update",
                                         "residual", "tensor_correction", "activation",
                                         "invdiag", "add_correction"};
    return id >= 0 && id < K_COUNT ? names[id] : "?";
}

static unsigned grid_of(int64_t n)
{
    int64_t nb = (n + BS - 1) / BS;
    nb = (nb + 7) / 8 * 8;  // multiple of 8 for the XCD remap
    return (unsigned)nb;
}

hipError_t launch_hex_laplacian(const HexGeom& g, const HexCoefs& k, const double* V,
                                const double* corr, double* out, hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_laplacian, dim3(grid_of(n)), dim3(BS), 0, s, g, k, V, corr, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ solver object
struct ImplicitSolver {
    HexGeom g;
    int64_t n;
    int prec;
    double chiCm, dt;
    hipStream_t stream;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr, *invd = nullptr;  // fp64 operator
    float *fx32 = nullptr, *fy32 = nullptr, *fz32 = nullptr, *invd32 = nullptr;
    double *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;        // fp64 CG
    float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *e32 = nullptr; // fp32 inner CG
    double* partials = nullptr;
    DevState* S = nullptr;
    DevState* hS = nullptr;  // pinned mirror
    double mass_c = -1.0;    // coefft of the current invd
    int64_t total_iters = 0;
    double last_res = 0.0;
    int last_inner = 16;
};

template <class T> static HexOp<T> op_of(const ImplicitSolver* S, double coefft);
template <> HexOp<double> op_of<double>(const ImplicitSolver* S, double c)
{
    return {S->fx, S->fy, S->fz, S->invd, c * S->chiCm / S->dt};
}
template <> HexOp<float> op_of<float>(const ImplicitSolver* S, double c)
{
    return {S->fx32, S->fy32, S->fz32, S->invd32, float(c * S->chiCm / S->dt)};
}

int implicit_create(ImplicitSolver** out, const HexGeom& g, const double* sig, double chiCm,
                    double dt, int prec, hipStream_t s, std::string* err)
{
    ImplicitSolver* S = new ImplicitSolver();
    S->g = g; S->n = (int64_t)g.nx * g.ny * g.nz; S->prec = prec; S->chiCm = chiCm;
    S->dt = dt; S->stream = s;
    const int64_t n = S->n;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    std::vector<double> fx(n, 0.0), fy(n, 0.0), fz(n, 0.0);
    for (int64_t c = 0; c < n; ++c) {
        const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / sz);
        if (i < g.nx - 1) fx[c] = 0.5 * (sig[9 * c + 0] + sig[9 * (c + 1) + 0]) / (g.hx * g.hx);
        if (j < g.ny - 1) fy[c] = 0.5 * (sig[9 * c + 4] + sig[9 * (c + sy) + 4]) / (g.hy * g.hy);
        if (k < g.nz - 1) fz[c] = 0.5 * (sig[9 * c + 8] + sig[9 * (c + sz) + 8]) / (g.hz * g.hz);
    }
    const int64_t nblk = grid_of(n);
    bool ok = true;
    auto A = [&](void** ptr, size_t bytes) { ok &= hipMalloc(ptr, bytes) == hipSuccess; };
    A((void**)&S->fx, 8 * n); A((void**)&S->fy, 8 * n); A((void**)&S->fz, 8 * n);
    A((void**)&S->invd, 8 * n);
    A((void**)&S->b, 8 * n); A((void**)&S->r, 8 * n); A((void**)&S->p, 8 * n);
    A((void**)&S->q, 8 * n);
    A((void**)&S->partials, 8 * 4 * nblk);
    A((void**)&S->S, sizeof(DevState));
    if (prec == 1) {
        A((void**)&S->fx32, 4 * n); A((void**)&S->fy32, 4 * n); A((void**)&S->fz32, 4 * n);
        A((void**)&S->invd32, 4 * n); A((void**)&S->r32, 4 * n); A((void**)&S->p32, 4 * n);
        A((void**)&S->q32, 4 * n); A((void**)&S->e32, 4 * n);
    }
    ok &= hipHostMalloc((void**)&S->hS, sizeof(DevState), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        implicit_destroy(S);
        *err = "implicit solver allocation failed";
        return 1;
    }
    hipMemcpy(S->fx, fx.data(), 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(S->fy, fy.data(), 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(S->fz, fz.data(), 8 * n, hipMemcpyHostToDevice);
    if (prec == 1) {
        std::vector<float> f(n);
        for (int64_t c = 0; c < n; ++c) f[c] = (float)fx[c];
        hipMemcpy(S->fx32, f.data(), 4 * n, hipMemcpyHostToDevice);
        for (int64_t c = 0; c < n; ++c) f[c] = (float)fy[c];
        hipMemcpy(S->fy32, f.data(), 4 * n, hipMemcpyHostToDevice);
        for (int64_t c = 0; c < n; ++c) f[c] = (float)fz[c];
        hipMemcpy(S->fz32, f.data(), 4 * n, hipMemcpyHostToDevice);
    }
    DevState z{};
    hipMemcpy(S->S, &z, sizeof z, hipMemcpyHostToDevice);
    *out = S;
    return 0;
}

void implicit_destroy(ImplicitSolver* S)
{
    if (!S) return;
    void* ptrs[] = {S->fx, S->fy, S->fz, S->invd, S->b, S->r, S->p, S->q, S->partials, S->S,
                    S->fx32, S->fy32, S->fz32, S->invd32, S->r32, S->p32, S->q32, S->e32};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (S->hS) hipHostFree(S->hS);
    delete S;
}

void implicit_stats(const ImplicitSolver* S, int64_t* iters, double* res)
{
    if (iters) *iters = S->total_iters;
    if (res) *res = S->last_res;
}

// set scalars of a fresh CG: tolerance^2 on ||r||^2, iteration budget
__global__ void k_cg_begin(DevState* S, double tol, int max_iter, int relative_to_r0)
{
    const double ref = relative_to_r0 ? S->s[S_R0] : S->s[S_BB];
    S->s[S_TOL2] = tol * tol * ref;
    S->iters = 0;
    S->max_iter = max_iter;
    S->done = S->s[S_RR] <= S->s[S_TOL2] ? 1 : 0;
}

template <class T, class TX>
static hipError_t cg_iterations(ImplicitSolver* S, const HexOp<T>& A, TX* x, T* r, T* p, T* q,
                                int count, Profiler* prof, std::vector<size_t>* marks)
{
    const unsigned grid = grid_of(S->n);
    for (int it = 0; it < count; ++it) {
        if (marks) marks->push_back(prof->mark());
        int tk = prof ? prof->begin(K_SPMV) : -1;
        hipLaunchKernelGGL((k_spmv_dot<T>), dim3(grid), dim3(BS), 0, S->stream, S->g, A, p, q,
                           S->partials, S->S);
        if (prof) { prof->end(tk); tk = prof->begin(K_UPDATE); }
        hipLaunchKernelGGL((k_update<T, TX>), dim3(grid), dim3(BS), 0, S->stream, S->g, A.invd,
                           x, r, p, q, S->partials, S->S);
        if (prof) { prof->end(tk); tk = prof->begin(K_PUPDATE); }
        hipLaunchKernelGGL((k_pupdate<T>), dim3(grid), dim3(BS), 0, S->stream, S->g, A.invd, r,
                           p, S->S);
        if (prof) prof->end(tk);
    }
    return hipGetLastError();
}

// runs CG chunks until the device flag says done; returns iterations used
template <class T, class TX>
static int cg_solve(ImplicitSolver* S, const HexOp<T>& A, TX* x, T* r, T* p, T* q, int max_iter,
                    Profiler* prof, std::string* err)
{
    int chunk = S->last_inner > 2 ? S->last_inner : 2;
    int launched = 0;
    int done_before = 0;
    std::vector<size_t> marks;
    while (true) {
        if (launched + chunk > max_iter) chunk = max_iter - launched;
        marks.clear();
        if (chunk > 0) {
            if (cg_iterations<T, TX>(S, A, x, r, p, q, chunk, prof, prof ? &marks : nullptr) !=
                hipSuccess) {
                *err = "CG kernel launch failed";
                return -1;
            }
            launched += chunk;
        }
        hipMemcpyAsync(S->hS, S->S, sizeof(DevState), hipMemcpyDeviceToHost, S->stream);
        if (hipStreamSynchronize(S->stream) != hipSuccess) {
            *err = "CG synchronisation failed";
            return -1;
        }
        if (prof) {
            // iterations launched after convergence exit at once and are not timed
            const size_t useful = (size_t)(S->hS->iters - done_before);
            if (useful < marks.size()) prof->invalidate_from(marks[useful]);
            prof->resolve(false);
        }
        done_before = S->hS->iters;
        if (S->hS->done || launched >= max_iter) break;
        chunk = 4;
    }
    S->last_inner = S->hS->iters;
    return S->hS->iters;
}

template <class M, class T>
static hipError_t ionic_rhs_t(ImplicitSolver* S, int tissue, void* st, const double* Vn,
                              const double* Vm1, double* Iion, const ExplicitStep& s,
                              const ModelParams& P)
{
    hipLaunchKernelGGL((k_ionic_rhs<M, T>), dim3(grid_of(S->n)), dim3(BS), 0, S->stream, S->g,
                       (T*)st, Vn, Vm1, Iion, S->b, s, P, tissue);
    return hipGetLastError();
}

int implicit_step(ImplicitSolver* S, int model, int prec, int tissue, void* states,
                  const ModelParams& P, const double* Vn, const double* Vm1, double* Vnew,
                  double* Iion, uint8_t* flag, double* tact, const ExplicitStep& st, double tol,
                  int max_iter, Profiler* prof, std::string* err)
{
    const unsigned grid = grid_of(S->n);
    hipError_t e;
    int tk = prof ? prof->begin(K_IONIC) : -1;
    switch (model * 2 + prec) {
    case 0: e = ionic_rhs_t<BuenoOrovio, double>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 1: e = ionic_rhs_t<BuenoOrovio, float>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 2: e = ionic_rhs_t<TNNP, double>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 3: e = ionic_rhs_t<TNNP, float>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 4: e = ionic_rhs_t<ORd, double>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 5: e = ionic_rhs_t<ORd, float>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 6: e = ionic_rhs_t<AlievPanfilov, double>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    case 7: e = ionic_rhs_t<AlievPanfilov, float>(S, tissue, states, Vn, Vm1, Iion, st, P); break;
    default: e = hipErrorInvalidValue;
    }
    if (prof) prof->end(tk);
    if (e != hipSuccess) { *err = std::string("ionic/rhs launch: ") + hipGetErrorString(e); return 4; }
    // 1/diag for this step's ddt coefficient
    if (st.c != S->mass_c) {
        HexOp<double> A = op_of<double>(S, st.c);
        hipLaunchKernelGGL((k_invdiag<double>), dim3(grid), dim3(BS), 0, S->stream, S->g, A, S->invd);
        if (S->prec == 1) {
            HexOp<float> A32 = op_of<float>(S, st.c);
            hipLaunchKernelGGL((k_invdiag<float>), dim3(grid), dim3(BS), 0, S->stream, S->g, A32, S->invd32);
        }
        S->mass_c = st.c;
    }
    hipMemcpyAsync(Vnew, Vn, 8 * S->n, hipMemcpyDeviceToDevice, S->stream);
    const HexOp<double> A = op_of<double>(S, st.c);
    int used = 0;
    if (prec == 0) {
        tk = prof ? prof->begin(K_RESIDUAL) : -1;
        hipLaunchKernelGGL((k_residual<double>), dim3(grid), dim3(BS), 0, S->stream, S->g, A, A,
                           (const double*)Vnew, (const double*)S->b, S->r, S->p, (double*)nullptr,
                           S->partials, S->S, 2);
        if (prof) prof->end(tk);
        hipLaunchKernelGGL(k_cg_begin, dim3(1), dim3(1), 0, S->stream, S->S, tol, max_iter, 0);
        used = cg_solve<double, double>(S, A, Vnew, S->r, S->p, S->q, max_iter, prof, err);
        if (used < 0) return 4;
        S->total_iters += used;
        S->last_res = std::sqrt(S->hS->s[S_RR] / (S->hS->s[S_BB] > 0 ? S->hS->s[S_BB] : 1.0));
        if (!S->hS->done || S->hS->s[S_RR] > S->hS->s[S_TOL2]) {
            if (used >= max_iter) { *err = "PCG did not converge within max_iter"; return 7; }
        }
    } else {
        const HexOp<float> A32 = op_of<float>(S, st.c);
        int outer = 0;
        int total = 0;
        const double inner_tol = 1e-4;
        while (true) {
            // fp64 residual; starts the fp32 correction solve from e = 0
            tk = prof ? prof->begin(K_RESIDUAL) : -1;
            hipLaunchKernelGGL((k_residual<float>), dim3(grid), dim3(BS), 0, S->stream, S->g, A,
                               A32, (const double*)Vnew, (const double*)S->b, S->r32, S->p32,
                               S->e32, S->partials, S->S, 2);
            if (prof) prof->end(tk);
            hipMemcpyAsync(S->hS, S->S, sizeof(DevState), hipMemcpyDeviceToHost, S->stream);
            hipStreamSynchronize(S->stream);
            const double rr = S->hS->s[S_RR], bb = S->hS->s[S_BB] > 0 ? S->hS->s[S_BB] : 1.0;
            S->last_res = std::sqrt(rr / bb);
            if (rr <= tol * tol * bb) break;
            if (total >= max_iter || outer >= 20) {
                *err = "mixed-precision refinement did not converge";
                return 7;
            }
            hipLaunchKernelGGL(k_cg_begin, dim3(1), dim3(1), 0, S->stream, S->S, inner_tol,
                               max_iter - total, 1);
            used = cg_solve<float, float>(S, A32, S->e32, S->r32, S->p32, S->q32, max_iter - total,
                                          prof, err);
            if (used < 0) return 4;
            total += used;
            tk = prof ? prof->begin(K_ADDCORR) : -1;
            hipLaunchKernelGGL(k_add_correction, dim3(grid), dim3(BS), 0, S->stream, S->g, Vnew,
                               (const float*)S->e32);
            if (prof) prof->end(tk);
            ++outer;
        }
        S->total_iters += total;
    }
    tk = prof ? prof->begin(K_ACTIVATION) : -1;
    hipLaunchKernelGGL(k_activation, dim3(grid), dim3(BS), 0, S->stream, S->g, Vn,
                       (const double*)Vnew, flag, tact, st.t_old, st.dt);
    if (prof) prof->end(tk);
    e = hipGetLastError();
    if (e != hipSuccess) { *err = hipGetErrorString(e); return 4; }
    return 0;
}

} // namespace cg
