// lduMatrix solve on an arbitrary mesh (fvMatrix::solve for Vm, monodomainSolver.C:129-135,
// on meshes that are not a hex block).
//
// The reference's lduMatrix holds diag[n] and, per face f in upper-triangular order,
// lowerAddr[f] < upperAddr[f] with upper[f] == lower[f] for the symmetric Vm equation.
// Here the face addressing is rebuilt once into ELL rows (K = largest row degree), stored
// column-major so that slot k of 64 consecutive rows is one contiguous segment:
//   col[k][i]  neighbour cell of row i in slot k (row i itself in padding slots)
//   fidx[k][i] face of that slot (-1 in padding), used to scatter new coefficients
//   val[k][i]  upper[fidx] (0 in padding), diag[i]
// y_i = diag_i x_i + sum_k val[k][i] x[col[k][i]]: coalesced reads of col/val/diag and
// gathers of x, which stays cache resident at these sizes.
//
// Solver: Jacobi PCG in the single-reduction form of Chronopoulos and Gear (the iterates
// of the textbook recurrence) with OpenFOAM's stopping rule: residual = sum|b - A psi| /
// normFactor, normFactor = sum |A psi0 - sumA xRef| + |b - sumA xRef| + 1e-20 (xRef = mean
// of psi0, sumA = row sums), converged below tolerance or below relTol x initial, tested
// before the first iteration; at most maxIter iterations (lduMatrix::solver, PCG::solve).
// Reductions are the per-block partials of cg_reduce.hpp.
#include <algorithm>
#include <string>
#include <vector>
#include "cg_ldu.hpp"
#include "cg_reduce.hpp"

namespace cg {

namespace {

constexpr int EBS = 256;  // threads per block of the elementwise kernels

enum { L_G = 0,   // (r,u), parity sets 0 and 1
       L_L1 = 2,  // sum |r|, parity sets 2 and 3
       L_D = 4,   // (z,u)
       L_NF = 5,  // normalisation terms
       L_XS = 6,  // sum of psi0
       L_NSLOT = 7 };

struct LduState {
    double alpha[2], gamma[2];
    double nf, l1_0, l1;
    long long done_gen;
    int iters;
    int pad;
};

struct EllView {
    int64_t n;
    int K;
    const int32_t* col;
    const double* val;
    const double* diag;
};

// A x for row i (diagonal first, then the slots in order)
CG_DEV double ell_row(const EllView& E, const double* __restrict__ x, int64_t i)
{
    double y = E.diag[i] * x[i];
    for (int k = 0; k < E.K; ++k) {
        const int64_t s = (int64_t)k * E.n + i;
        y += E.val[s] * x[E.col[s]];
    }
    return y;
}

CG_DEV double ell_rowsum(const EllView& E, int64_t i)
{
    double a = E.diag[i];
    for (int k = 0; k < E.K; ++k) a += E.val[(int64_t)k * E.n + i];
    return a;
}

// contiguous chunk of rows per block for the reducing kernels (RG blocks of RBS)
#define ROW_LOOP(N)                                                                        \
    const int64_t chunk_ = (((N) + RG - 1) / RG + RBS - 1) / RBS * RBS;                    \
    const int64_t lo_ = (int64_t)blockIdx.x * chunk_;                                      \
    const int64_t hi_ = lo_ + chunk_ < (N) ? lo_ + chunk_ : (N);                           \
    for (int64_t i = lo_ + threadIdx.x; i < hi_; i += RBS)

__global__ __launch_bounds__(EBS) void k_fill(int64_t n, int K, const int32_t* __restrict__ fidx,
                                              const double* __restrict__ upper,
                                              double* __restrict__ val)
{
    const int64_t s = (int64_t)blockIdx.x * EBS + threadIdx.x;
    if (s >= (int64_t)K * n) return;
    const int32_t f = fidx[s];
    val[s] = f >= 0 ? upper[f] : 0.0;
}

__global__ __launch_bounds__(EBS) void k_amul(EllView E, const double* __restrict__ x,
                                              double* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * EBS + threadIdx.x;
    if (i >= E.n) return;
    y[i] = ell_row(E, x, i);
}

__global__ __launch_bounds__(RBS) void k_mean(int64_t n, const double* __restrict__ psi,
                                              double* __restrict__ part)
{
    double v[1] = {0.0};
    ROW_LOOP(n) v[0] += psi[i];
    block_sum<RBS, 1>(v);
    if (threadIdx.x == 0) part[L_XS * RG + blockIdx.x] = v[0];
}

// r = b - A psi, u = r / diag; partials (r,u), sum|r|, normalisation terms
__global__ __launch_bounds__(RBS) void k_r0(EllView E, const double* __restrict__ psi,
                                            const double* __restrict__ b,
                                            double* __restrict__ r, double* __restrict__ u,
                                            double* __restrict__ part)
{
    double xs[1];
    sum_slots<1>(part, {L_XS}, xs);
    const double xref = xs[0] / double(E.n);
    double v[3] = {0.0, 0.0, 0.0};
    ROW_LOOP(E.n)
    {
        const double ax = ell_row(E, psi, i);
        const double bi = b[i];
        const double ri = bi - ax;
        const double ui = ri / E.diag[i];
        r[i] = ri;
        u[i] = ui;
        const double t = ell_rowsum(E, i) * xref;
        v[0] += ri * ui;
        v[1] += fabs(ri);
        v[2] += fabs(ax - t) + fabs(bi - t);
    }
    block_sum<RBS, 3>(v);
    if (threadIdx.x == 0) {
        part[(L_G + 0) * RG + blockIdx.x] = v[0];
        part[(L_L1 + 0) * RG + blockIdx.x] = v[1];
        part[L_NF * RG + blockIdx.x] = v[2];
    }
}

// iteration k: convergence test on r_k, then z = A u and the partial (z,u)
__global__ __launch_bounds__(RBS) void k_s(EllView E, const double* __restrict__ u,
                                           double* __restrict__ z, double* __restrict__ part,
                                           LduState* S, long long gen, int k, double tol,
                                           double rel, int max_iter)
{
    if (S->done_gen == gen) return;
    double v3[3];
    sum_slots<3>(part, {L_G + (k & 1), L_L1 + (k & 1), L_NF}, v3);
    const double l1 = v3[1], nf = v3[2] + 1e-20;
    const double l10 = k == 0 ? l1 : S->l1_0;
    const bool conv = l1 < tol * nf || (rel > 1e-20 && l1 < rel * l10);
    const bool stop = conv || k >= max_iter;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (k == 0) {
            S->nf = nf;
            S->l1_0 = l1;
        }
        if (stop) {
            S->l1 = l1;
            S->iters = k;
            S->done_gen = gen;
        }
    }
    if (stop) return;
    double v[1] = {0.0};
    ROW_LOOP(E.n)
    {
        const double zi = ell_row(E, u, i);
        z[i] = zi;
        v[0] += zi * u[i];
    }
    block_sum<RBS, 1>(v);
    if (threadIdx.x == 0) part[L_D * RG + blockIdx.x] = v[0];
}

// iteration k: gamma = (r,u), delta = (z,u); p = u + beta p, s = z + beta s,
// psi += alpha p, r -= alpha s, u = r / diag; partials of iteration k+1
__global__ __launch_bounds__(RBS) void k_u(EllView E, double* __restrict__ psi,
                                           double* __restrict__ r, double* __restrict__ u,
                                           double* __restrict__ p, double* __restrict__ s,
                                           const double* __restrict__ z,
                                           double* __restrict__ part, LduState* S,
                                           long long gen, int k)
{
    if (S->done_gen == gen) return;
    double v2[2];
    sum_slots<2>(part, {L_G + (k & 1), L_D}, v2);
    const double gam = v2[0], del = v2[1];
    double alpha, beta = 0.0;
    if (k == 0) {
        alpha = gam / del;
    } else {
        const double gp = S->gamma[(k - 1) & 1], ap = S->alpha[(k - 1) & 1];
        beta = gam / gp;
        alpha = gam / (del - beta * gam / ap);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->alpha[k & 1] = alpha;
        S->gamma[k & 1] = gam;
    }
    const bool first = k == 0;
    double v[2] = {0.0, 0.0};
    ROW_LOOP(E.n)
    {
        const double pc = first ? u[i] : u[i] + beta * p[i];
        const double sc = first ? z[i] : z[i] + beta * s[i];
        p[i] = pc;
        s[i] = sc;
        psi[i] += alpha * pc;
        const double rn = r[i] - alpha * sc;
        const double un = rn / E.diag[i];
        r[i] = rn;
        u[i] = un;
        v[0] += rn * un;
        v[1] += fabs(rn);
    }
    block_sum<RBS, 2>(v);
    if (threadIdx.x == 0) {
        part[(L_G + ((k + 1) & 1)) * RG + blockIdx.x] = v[0];
        part[(L_L1 + ((k + 1) & 1)) * RG + blockIdx.x] = v[1];
    }
}

unsigned blocks_of(int64_t n) { return (unsigned)((n + EBS - 1) / EBS); }

} // namespace

struct LduSolver {
    int64_t n = 0, nf = 0;
    int K = 0;
    int32_t* col = nullptr;
    int32_t* fidx = nullptr;
    double* val = nullptr;
    double* diag = nullptr;
    double* upper = nullptr;  // staging for the face coefficients
    double *psi = nullptr, *b = nullptr, *y = nullptr;  // staging for host arrays
    double *r = nullptr, *u = nullptr, *p = nullptr, *s = nullptr, *z = nullptr;
    double* part = nullptr;
    LduState* S = nullptr;
    LduState* hS = nullptr;
    hipStream_t stream = nullptr;
    long long gen = 0;
    bool coeffs = false;
    EllView view() const { return {n, K, col, val, diag}; }
};

int ldu_create(LduSolver** out, int64_t n, int64_t nf, const int32_t* lower,
               const int32_t* upper, std::string* err)
{
    *out = nullptr;
    std::vector<int> deg((size_t)n, 0);
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t l = lower[f], u = upper[f];
        if (l < 0 || u < 0 || l >= n || u >= n || l == u) {
            *err = "face " + std::to_string(f) + ": cell index out of range or self-connected";
            return 1;
        }
        ++deg[(size_t)l];
        ++deg[(size_t)u];
    }
    int K = 0;
    for (int d : deg) K = std::max(K, d);
    K = std::max(K, 1);
    // slots in face order per row
    std::vector<int32_t> col((size_t)K * n), fidx((size_t)K * n, -1);
    std::vector<int> fill((size_t)n, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < K; ++k) col[(size_t)k * n + i] = (int32_t)i;
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t l = lower[f], u = upper[f];
        int& a = fill[(size_t)l];
        col[(size_t)a * n + l] = (int32_t)u;
        fidx[(size_t)a * n + l] = (int32_t)f;
        ++a;
        int& c = fill[(size_t)u];
        col[(size_t)c * n + u] = (int32_t)l;
        fidx[(size_t)c * n + u] = (int32_t)f;
        ++c;
    }
    LduSolver* L = new LduSolver();
    L->n = n; L->nf = nf; L->K = K;
    bool ok = true;
    auto A = [&](void** ptr, size_t bytes) { ok &= hipMalloc(ptr, bytes) == hipSuccess; };
    A((void**)&L->col, 4 * (size_t)K * n);
    A((void**)&L->fidx, 4 * (size_t)K * n);
    A((void**)&L->val, 8 * (size_t)K * n);
    A((void**)&L->diag, 8 * n);
    A((void**)&L->upper, 8 * std::max<int64_t>(nf, 1));
    A((void**)&L->psi, 8 * n); A((void**)&L->b, 8 * n); A((void**)&L->y, 8 * n);
    A((void**)&L->r, 8 * n); A((void**)&L->u, 8 * n); A((void**)&L->p, 8 * n);
    A((void**)&L->s, 8 * n); A((void**)&L->z, 8 * n);
    A((void**)&L->part, 8 * RG * L_NSLOT);
    A((void**)&L->S, sizeof(LduState));
    ok &= hipHostMalloc((void**)&L->hS, sizeof(LduState), hipHostMallocDefault) == hipSuccess;
    ok &= hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        ldu_destroy(L);
        *err = "device allocation for the lduMatrix failed";
        return 5;
    }
    hipMemcpy(L->col, col.data(), 4 * (size_t)K * n, hipMemcpyHostToDevice);
    hipMemcpy(L->fidx, fidx.data(), 4 * (size_t)K * n, hipMemcpyHostToDevice);
    LduState z{};
    z.done_gen = -1;
    hipMemcpy(L->S, &z, sizeof z, hipMemcpyHostToDevice);
    *out = L;
    return 0;
}

void ldu_destroy(LduSolver* L)
{
    if (!L) return;
    if (L->stream) hipStreamSynchronize(L->stream);
    void* ptrs[] = {L->col, L->fidx, L->val, L->diag, L->upper, L->psi, L->b, L->y,
                    L->r,   L->u,    L->p,   L->s,    L->z,     L->part, L->S};
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (L->hS) hipHostFree(L->hS);
    if (L->stream) hipStreamDestroy(L->stream);
    delete L;
}

int64_t ldu_rows(const LduSolver* L) { return L->n; }
int ldu_width(const LduSolver* L) { return L->K; }

int ldu_set_coeffs(LduSolver* L, const double* diag, const double* upper, bool device)
{
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(L->diag, diag, 8 * L->n, kind, L->stream) != hipSuccess) return 4;
    if (L->nf > 0 && hipMemcpyAsync(L->upper, upper, 8 * L->nf, kind, L->stream) != hipSuccess)
        return 4;
    const int64_t slots = (int64_t)L->K * L->n;
    hipLaunchKernelGGL(k_fill, dim3(blocks_of(slots)), dim3(EBS), 0, L->stream, L->n, L->K,
                       L->fidx, L->upper, L->val);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(L->stream) != hipSuccess)
        return 4;
    L->coeffs = true;
    return 0;
}

int ldu_amul(LduSolver* L, const double* x, double* y, bool device, hipEvent_t ev0,
             hipEvent_t ev1)
{
    const double* dx = x;
    double* dy = y;
    if (!device) {
        if (hipMemcpyAsync(L->psi, x, 8 * L->n, hipMemcpyHostToDevice, L->stream) != hipSuccess)
            return 4;
        dx = L->psi;
        dy = L->y;
    }
    if (ev0) hipEventRecord(ev0, L->stream);
    hipLaunchKernelGGL(k_amul, dim3(blocks_of(L->n)), dim3(EBS), 0, L->stream, L->view(), dx, dy);
    if (ev1) hipEventRecord(ev1, L->stream);
    if (!device && hipMemcpyAsync(y, dy, 8 * L->n, hipMemcpyDeviceToHost, L->stream) != hipSuccess)
        return 4;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(L->stream) != hipSuccess)
        return 4;
    return 0;
}

int ldu_solve(LduSolver* L, double* psi, const double* b, bool device, double tol, double rel,
              int max_iter, int* iters, double* init_res, double* final_res)
{
    double* dpsi = psi;
    const double* db = b;
    if (!device) {
        if (hipMemcpyAsync(L->psi, psi, 8 * L->n, hipMemcpyHostToDevice, L->stream) != hipSuccess ||
            hipMemcpyAsync(L->b, b, 8 * L->n, hipMemcpyHostToDevice, L->stream) != hipSuccess)
            return 4;
        dpsi = L->psi;
        db = L->b;
    }
    const EllView E = L->view();
    const long long gen = ++L->gen;
    hipLaunchKernelGGL(k_mean, dim3(RG), dim3(RBS), 0, L->stream, L->n, (const double*)dpsi, L->part);
    hipLaunchKernelGGL(k_r0, dim3(RG), dim3(RBS), 0, L->stream, E, (const double*)dpsi, db,
                       L->r, L->u, L->part);
    hipLaunchKernelGGL(k_s, dim3(RG), dim3(RBS), 0, L->stream, E, (const double*)L->u, L->z,
                       L->part, L->S, gen, 0, tol, rel, max_iter);
    int k = 0;
    for (;;) {
        // a chunk of iterations, then the 64-byte state; iterations after convergence
        // return at once
        for (int q = 0; q < 8; ++q, ++k) {
            hipLaunchKernelGGL(k_u, dim3(RG), dim3(RBS), 0, L->stream, E, dpsi, L->r, L->u,
                               L->p, L->s, (const double*)L->z, L->part, L->S, gen, k);
            hipLaunchKernelGGL(k_s, dim3(RG), dim3(RBS), 0, L->stream, E, (const double*)L->u,
                               L->z, L->part, L->S, gen, k + 1, tol, rel, max_iter);
        }
        if (hipMemcpyAsync(L->hS, L->S, sizeof(LduState), hipMemcpyDeviceToHost, L->stream) !=
                hipSuccess ||
            hipStreamSynchronize(L->stream) != hipSuccess)
            return 4;
        if (L->hS->done_gen == gen) break;
        if (k > max_iter + 16) return 4;  // the state stopped advancing
    }
    if (!device &&
        (hipMemcpyAsync(psi, dpsi, 8 * L->n, hipMemcpyDeviceToHost, L->stream) != hipSuccess ||
         hipStreamSynchronize(L->stream) != hipSuccess))
        return 4;
    const LduState& h = *L->hS;
    if (iters) *iters = h.iters;
    if (init_res) *init_res = h.l1_0 / h.nf;
    if (final_res) *final_res = h.l1 / h.nf;
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

} // namespace cg
