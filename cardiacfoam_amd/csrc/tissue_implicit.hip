// Implicit monodomain step on the hex block (solveDiffusionImplicit,
// monodomainSolver.C:121-150).
//
// Matrix per unit volume (fvm::laplacian, orthogonal part; fvm::ddt with coefficient c):
//   (c chiCm/dt + sum_f a_f) V_P - sum_f a_f V_N = rhs,   a_f = sigma_nn,f / h_d^2
// stored as three face-coefficient planes fx, fy, fz (face to +x/+y/+z, 0 on the wall)
// and 1/diag, applied as a 7-point stencil without index arrays.
// Right-hand side, built by the ionic kernel so Iion never takes a pass of its own:
//   rhs = chiCm/dt (c0 V^n - c00 V^{n-1}) - chiCm Iion + source
// to which the first residual pass adds the explicit tensor correction
//   div(SfGammaCorr . interp grad V^n)   (19-point stencil over V^n and sigma xy, xz, yz).
//
// Solver: Jacobi-preconditioned conjugate gradients in the single-reduction form of
// Chronopoulos and Gear (J. Comput. Appl. Math. 25, 1989): the same iterates as the
// textbook recurrence, with both dot products of an iteration taken in one pass.
//   fp64 path : CG in fp64 to ||b - A x||_2 <= tol ||b||_2.
//   mixed path: iterative refinement; residual and solution in fp64, each correction from
//               an fp32 CG (fp32 operator and vectors, fp64 dot products) to 1e-4 relative.
// Kernels of one step:
//   residual (first)  x0 = V^n, b += tensor correction, r = b - A x0, u = M^-1 r
//   per iteration k   S: convergence test on r_k; w = A u, partial (w,u)
//                     U: alpha, beta; p = u + beta p, s = w + beta s, x += alpha p,
//                        r -= alpha s, u = M^-1 r, partials (r,u), (r,r)
//   residual (mixed)  x' = x + e (ping-pong buffers), r = b - A x', next correction start
//   activation        first crossing of V > 1e-15 V, copy of the solution into V^{n+1}
// Scalars are never stored for readers of the same launch: every block recomputes them
// from per-block partial sums (fixed summation order, so all blocks agree); partial slots
// read and written in one launch are double-buffered by iteration parity.  Each CG solve
// carries a generation number, so "solve g is finished" needs no reset.  No atomics, no
// device-scope fences.
#include <cmath>
#include <cstring>
#include <vector>
#include "cg_implicit.hpp"
#include "cg_hexops.hpp"
#include "cg_reduce.hpp"

namespace cg {

constexpr int BS = 256;    // elementwise kernels

// partial-sum slots (RG doubles each)
enum { P_G = 0,      // (r,u), parity sets 0 and 1
       P_RR = 2,     // (r,r) in the CG precision, parity sets 2 and 3
       P_D = 4,      // (w,u)
       P_RR0 = 5,    // (r,r) of the fp64 residual
       P_BB = 6,     // (b,b)
       P_L1 = 7,     // sum |r| (OpenFOAM norm), parity sets 7 and 8
       P_NF = 9,     // normalisation factor terms of the solve's first residual
       P_XS = 10,    // sum of the initial guess (its mean is OpenFOAM's xRef)
       P_NSLOT = 11 };

enum { S_RR, S_BB, S_R0, S_NSCAL };
struct DevState {
    double s[S_NSCAL];        // last (r,r), (b,b), fp64 (r0,r0)
    double alpha[2], gamma[2];  // by iteration parity, written by block 0 of the update
    long long done_gen;       // generation of the last finished CG solve
    long long outer_gen;      // generation whose start already met the outer tolerance
    long long conv_gen;       // generation of the last solve that met the step tolerance
    int iters;                // iterations of the current solve
    int pad;
    long long iters_sum;      // iterations of all solves, added by each residual pass
    double nf;                // OpenFOAM normalisation factor of the current solve
    double l1_0, l1;          // sum |r| at the start and at the end of the current solve
    double init_res;          // initial residual of the outer corrector (in its norm)
};

static_assert(sizeof(DevState) % 8 == 0 && sizeof(DevState) / 8 <= 64,
              "the state is published by the lanes of one wave, one 8-byte word each");

template <class T> struct HexOp {
    const T* fx;
    const T* fy;
    const T* fz;
    const T* invd;
    T mass;  // c chiCm / dt
};

// (A x)_c = mass x_c + sum_f a_f (x_c - x_nb), x read through a loader (plain array, or
// fp64 x plus fp32 correction).  Branch-free: wall neighbours are clamped to c, so their
// difference is 0 whatever the coefficient, and all 13 loads of a row issue together.
template <class T, class LD>
CG_DEV T apply_row(const HexOp<T>& A, const LD& x, int64_t c, const IJK& q, const HexGeom& g)
{
    const Nbr b = nbr_of(c, q, g);
    const T xc = T(x(c));
    const T fl = A.fx[c] * (xc - T(x(b.xp))) + A.fx[b.xm] * (xc - T(x(b.xm))) +
                 A.fy[c] * (xc - T(x(b.yp))) + A.fy[b.ym] * (xc - T(x(b.ym))) +
                 A.fz[c] * (xc - T(x(b.zp))) + A.fz[b.zm] * (xc - T(x(b.zm)));
    return A.mass * xc + fl;
}

// ---- four consecutive rows (c0..c0+3, one x-row segment; nx % 4 == 0) of A x with
// 16/32-byte vector loads: 7 vector loads of x and face planes plus two x scalars per quad
template <class T> using V4 = T __attribute__((ext_vector_type(4)));
template <class T> CG_DEV V4<T> ld4(const T* p) { return *reinterpret_cast<const V4<T>*>(p); }

// x loaders: one value (operator()) or four consecutive ones from a 32-byte aligned c (q4)
template <class T> struct PlainLoad {
    const T* __restrict__ x;
    CG_DEV T operator()(int64_t c) const { return x[c]; }
    CG_DEV V4<T> q4(int64_t c) const { return ld4(x + c); }
};
struct ExtrapLoad {  // linear extrapolation 2 x - x_prev (initial guess of a step)
    const double* __restrict__ x;
    const double* __restrict__ xp;
    CG_DEV double operator()(int64_t c) const { return 2.0 * x[c] - xp[c]; }
    CG_DEV V4<double> q4(int64_t c) const { return 2.0 * ld4(x + c) - ld4(xp + c); }
};
struct CorrectedLoad {  // fp64 x plus the fp32 correction e
    const double* __restrict__ x;
    const float* __restrict__ e;
    CG_DEV double operator()(int64_t c) const { return x[c] + double(e[c]); }
    CG_DEV V4<double> q4(int64_t c) const
    {
        return ld4(x + c) + __builtin_convertvector(ld4(e + c), V4<double>);
    }
};

template <class T, class LD>
CG_DEV V4<T> apply_quad_ld(const HexOp<T>& A, const LD& x, int64_t c0, const IJK& q,
                           const HexGeom& g)
{
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t ym = q.j > 0 ? c0 - sy : c0, yp = q.j < g.ny - 1 ? c0 + sy : c0;
    const int64_t zm = q.k > 0 ? c0 - sz : c0, zp = q.k < g.nz - 1 ? c0 + sz : c0;
    const int64_t xl = q.i > 0 ? c0 - 1 : c0, xr = q.i + 4 < g.nx ? c0 + 4 : c0 + 3;
    const V4<T> xc = x.q4(c0);
    const V4<T> xym = x.q4(ym), xyp = x.q4(yp), xzm = x.q4(zm), xzp = x.q4(zp);
    const V4<T> fx = ld4(A.fx + c0), fyp = ld4(A.fy + c0), fym = ld4(A.fy + ym);
    const V4<T> fzp = ld4(A.fz + c0), fzm = ld4(A.fz + zm);
    const T x_l = x(xl), x_r = x(xr), fx_l = A.fx[xl];
    // x-direction neighbours of the four rows; fx of the +x wall face is 0
    const V4<T> xxm = {x_l, xc.x, xc.y, xc.z};
    const V4<T> xxp = {xc.y, xc.z, xc.w, x_r};
    const V4<T> fxm = {fx_l, fx.x, fx.y, fx.z};
    return A.mass * xc + (fx * (xc - xxp) + fxm * (xc - xxm)) + (fyp * (xc - xyp) + fym * (xc - xym)) +
           (fzp * (xc - xzp) + fzm * (xc - xzm));
}

template <class T>
CG_DEV V4<T> apply_quad(const HexOp<T>& A, const T* __restrict__ x, int64_t c0, const IJK& q,
                        const HexGeom& g)
{
    return apply_quad_ld(A, PlainLoad<T>{x}, c0, q, g);
}

// The explicit tensor correction of hex_tensor_correction for four consecutive cells when
// only the xy conductivity is non-zero (fibres in the xy plane): the rows j-1, j, j+1 of V
// and of sigma_xy as 32-byte loads plus their end values, the y-gradients of the cells and
// their x-neighbours and the x-gradients of the cells and their y-neighbours taken from
// those rows.  Same walls (clamped neighbours) and the same expression order per cell.
CG_DEV V4<double> quad_correction_xy(const double* __restrict__ V,
                                     const double* __restrict__ oxy, int64_t c0, const IJK& q,
                                     const HexGeom& g)
{
    const int64_t sy = g.nx;
    const int64_t rm = q.j > 0 ? c0 - sy : c0, rp = q.j < g.ny - 1 ? c0 + sy : c0;
    const int64_t xl = q.i > 0 ? -1 : 0, xr = q.i + 4 < g.nx ? 4 : 3;
    const V4<double> v0 = ld4(V + c0), vm = ld4(V + rm), vp = ld4(V + rp);
    const double v0l = V[c0 + xl], v0r = V[c0 + xr], vml = V[rm + xl], vmr = V[rm + xr];
    const double vpl = V[rp + xl], vpr = V[rp + xr];
    const V4<double> o0 = ld4(oxy + c0), om = ld4(oxy + rm), op = ld4(oxy + rp);
    const double o0l = oxy[c0 + xl], o0r = oxy[c0 + xr];
    const double rhx = 0.5 / g.hx, rhy = 0.5 / g.hy;
    // faces -x/+x (d = x, e = y): y-gradients at the cells and their x-neighbours
    const V4<double> gy = (vp - vm) * rhy;
    const double gyl = (vpl - vml) * rhy, gyr = (vpr - vmr) * rhy;
    const V4<double> gyxm = {gyl, gy.x, gy.y, gy.z}, gyxp = {gy.y, gy.z, gy.w, gyr};
    const V4<double> oxm = {o0l, o0.x, o0.y, o0.z}, oxp = {o0.y, o0.z, o0.w, o0r};
    const V4<double> qx = (o0 + oxp) * (gy + gyxp) - (o0 + oxm) * (gy + gyxm);
    // faces -y/+y (d = y, e = x): x-gradients at the cells and their y-neighbours
    auto ddx = [&](const V4<double>& r, double l, double rr) {
        const V4<double> up = {r.y, r.z, r.w, rr}, dn = {l, r.x, r.y, r.z};
        return (up - dn) * rhx;
    };
    const V4<double> gx = ddx(v0, v0l, v0r), gxm = ddx(vm, vml, vmr), gxp = ddx(vp, vpl, vpr);
    const V4<double> qy = (o0 + op) * (gx + gxp) - (o0 + om) * (gx + gxm);
    return 0.25 * qx * (2.0 * rhx) + 0.25 * qy * (2.0 * rhy);
}

// one cell per lane (elementwise kernels)
#define CELL_INDEX                                                                         \
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);                                 \
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;                                         \
    const int64_t c = (int64_t)bid * BS + threadIdx.x;                                     \
    if (c >= n) return;                                                                    \
    const IJK cc = ijk_of((uint32_t)c, g);

// contiguous chunk of cells per block (reducing kernels on RG blocks of RBS threads)
#define RANGE_LOOP(g)                                                                      \
    const int64_t n = (int64_t)(g).nx * (g).ny * (g).nz;                                   \
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);                                 \
    const int64_t chunk_ = ((n + RG - 1) / RG + RBS - 1) / RBS * RBS;                      \
    const int64_t lo_ = (int64_t)bid * chunk_;                                             \
    const int64_t hi_ = lo_ + chunk_ < n ? lo_ + chunk_ : n;                               \
    for (int64_t c = lo_ + threadIdx.x; c < hi_; c += RBS)

// the same chunks in quads of 4 consecutive cells (c0 = 4 qi), for nx % 4 == 0
#define RANGE_LOOP_Q(g)                                                                    \
    const int64_t nq_ = (int64_t)(g).nx * (g).ny * (g).nz / 4;                             \
    const int64_t qchunk_ = ((nq_ + RG - 1) / RG + RBS - 1) / RBS * RBS;                   \
    const int64_t qlo_ = (int64_t)bid * qchunk_;                                           \
    const int64_t qhi_ = qlo_ + qchunk_ < nq_ ? qlo_ + qchunk_ : nq_;                      \
    for (int64_t qi = qlo_ + threadIdx.x; qi < qhi_; qi += RBS)

template <class T> CG_DEV double dot4(const V4<T>& a, const V4<T>& b)
{
    return (double(a.x) * double(b.x) + double(a.y) * double(b.y)) +
           (double(a.z) * double(b.z) + double(a.w) * double(b.w));
}
template <class T> CG_DEV void st4(T* p, const V4<T>& v) { *reinterpret_cast<V4<T>*>(p) = v; }

// ---- 1/diag for a given mass coefficient
template <class T>
__global__ __launch_bounds__(BS) void k_invdiag(HexGeom g, HexOp<T> A, T* __restrict__ invd)
{
    CELL_INDEX
    const Nbr b = nbr_of(c, cc, g);
    const T fxm = A.fx[b.xm], fym = A.fy[b.ym], fzm = A.fz[b.zm];
    const T d = A.mass + ((A.fx[c] + (cc.i > 0 ? fxm : T(0))) +
                          (A.fy[c] + (cc.j > 0 ? fym : T(0))) +
                          (A.fz[c] + (cc.k > 0 ? fzm : T(0))));
    invd[c] = T(1) / d;
}

// ---- fp64 residual that starts CG solve `gen` in precision TP.
//   x = xin (+ e when e != null and solve gen-1 did not stop at its outer test), or
//   x = 2 xin - xprev when xprev != null (extrapolated initial guess of a step);
//   xout = x when xout != null (xout never aliases xin or e);
//   b += tensor correction of xin when od != null (first pass of a step);
//   r = b - A x;  rt = TP(r), ut = M^-1 rt.
//   partials: (rt,ut) -> G set 0, (rt,rt) -> RR set 0, (r,r) -> RR0, (b,b) -> BB.
//   Block 0 adds the iterations of the previous solve to the running sum.
//   OF: also sum |r| and the normalisation terms |A x - mass xref| + |b - mass xref|
//   (lduMatrix::solver::normFactor; the row sums of A are the mass coefficient).
template <class TP, class LD, bool CORR, bool OF>
CG_DEV void residual_range(const HexGeom& g, const HexOp<double>& A, const TP* __restrict__ invd_t,
                           const LD& X, const double* __restrict__ xin, double* __restrict__ xout,
                           const double* bsrc, double* bout, const double* __restrict__ od,
                           unsigned od_mask, TP* __restrict__ rt, TP* __restrict__ ut,
                           double xref, double (&v)[6])
{
    const double mref = A.mass * xref;
    RANGE_LOOP(g)
    {
        const IJK q = ijk_of((uint32_t)c, g);
        double bc = bsrc[c];
        if (CORR) bc += hex_tensor_correction(xin, od, n, c, q.i, q.j, q.k, g, od_mask);
        if (bout) bout[c] = bc;
        if (xout) xout[c] = X(c);
        const double ax = apply_row(A, X, c, q, g);
        const double r = bc - ax;
        const TP rr = TP(r);
        const TP u = invd_t[c] * rr;
        rt[c] = rr;
        ut[c] = u;
        v[0] += double(rr) * double(u);
        v[1] += double(rr) * double(rr);
        v[2] += r * r;
        v[3] += bc * bc;
        if (OF) {
            v[4] += fabs(r);
            v[5] += fabs(ax - mref) + fabs(bc - mref);
        }
    }
}

// the same on quads of cells (nx % 4 == 0), the correction limited to sigma_xy (CORRXY)
template <class TP, class LD, bool CORRXY, bool OF>
CG_DEV void residual_range_q(const HexGeom& g, const HexOp<double>& A,
                             const TP* __restrict__ invd_t, const LD& X,
                             const double* __restrict__ xin, double* __restrict__ xout,
                             const double* bsrc, double* bout, const double* __restrict__ oxy,
                             TP* __restrict__ rt, TP* __restrict__ ut, double xref,
                             double (&v)[6], uint32_t bid)
{
    const double mref = A.mass * xref;
    RANGE_LOOP_Q(g)
    {
        const int64_t c0 = qi * 4;
        const IJK q = ijk_of((uint32_t)c0, g);
        V4<double> bc = ld4(bsrc + c0);
        if (CORRXY) bc += quad_correction_xy(xin, oxy, c0, q, g);
        if (bout) st4(bout + c0, bc);
        if (xout) st4(xout + c0, X.q4(c0));
        const V4<double> ax = apply_quad_ld(A, X, c0, q, g);
        const V4<double> r = bc - ax;
        const V4<TP> rr = __builtin_convertvector(r, V4<TP>);
        const V4<TP> u = ld4(invd_t + c0) * rr;
        st4(rt + c0, rr);
        st4(ut + c0, u);
        v[0] += dot4(rr, u);
        v[1] += dot4(rr, rr);
        v[2] += dot4(r, r);
        v[3] += dot4(bc, bc);
        if (OF) {
            v[4] += (fabs(r.x) + fabs(r.y)) + (fabs(r.z) + fabs(r.w));
            const V4<double> t1 = ax - mref, t2 = bc - mref;
            v[5] += ((fabs(t1.x) + fabs(t2.x)) + (fabs(t1.y) + fabs(t2.y))) +
                    ((fabs(t1.z) + fabs(t2.z)) + (fabs(t1.w) + fabs(t2.w)));
        }
    }
}

// ---- sum of the initial guess x (OpenFOAM's xRef is its mean)
template <class LD>
__global__ __launch_bounds__(RBS) void k_mean(HexGeom g, LD X, double* __restrict__ part)
{
    double v[1] = {0.0};
    RANGE_LOOP(g) v[0] += X(c);
    block_sum<RBS, 1>(v);
    if (threadIdx.x == 0) part[P_XS * RG + bid] = v[0];
}

//   bsrc: right-hand side read; bout: where b plus the correction goes (null: nowhere,
//   may equal bsrc).  OF: partial sum |r| -> L1 set 0 and, with of_nf (first residual of a
//   solve), the normalisation terms -> NF with xref = mean from the k_mean partials.
// QUAD (nx % 4 == 0, correction absent or sigma_xy only): four cells per lane.
template <class TP, bool OF, bool QUAD>
__global__ __launch_bounds__(RBS) void k_residual(HexGeom g, HexOp<double> A,
                                                  const TP* __restrict__ invd_t,
                                                  const double* __restrict__ xin,
                                                  const double* __restrict__ vcor,
                                                  const double* __restrict__ xprev,
                                                  const float* __restrict__ e,
                                                  double* __restrict__ xout,
                                                  const double* bsrc, double* bout,
                                                  const double* __restrict__ od, unsigned od_mask,
                                                  TP* __restrict__ rt, TP* __restrict__ ut,
                                                  double* __restrict__ part, DevState* S,
                                                  long long gen, int of_nf)
{
    const bool add_e = e != nullptr && S->outer_gen != gen - 1;
    double xref = 0.0;
    if (OF && of_nf) {
        double xs[1];
        sum_slots<1>(part, {P_XS}, xs);
        xref = xs[0] / double((int64_t)g.nx * g.ny * g.nz);
    }
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double* vc = vcor ? vcor : xin;  // field of the tensor correction
    const PlainLoad<double> X0{xin};
    const ExtrapLoad XE{xin, xprev};
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    if constexpr (QUAD) {
#define CG_RESIDUAL(LD, XL, CORR)                                                          \
    residual_range_q<TP, LD, CORR, OF>(g, A, invd_t, XL, vc, xout, bsrc, bout, od, rt, ut,   \
                                       xref, v, bid)
        if (od && xprev) CG_RESIDUAL(ExtrapLoad, XE, true);
        else if (od) CG_RESIDUAL(PlainLoad<double>, X0, true);
        else if (xprev) CG_RESIDUAL(ExtrapLoad, XE, false);
        else if (add_e) CG_RESIDUAL(CorrectedLoad, (CorrectedLoad{xin, e}), false);
        else CG_RESIDUAL(PlainLoad<double>, X0, false);
#undef CG_RESIDUAL
    } else {
#define CG_RESIDUAL(LD, XL, CORR)                                                          \
    residual_range<TP, LD, CORR, OF>(g, A, invd_t, XL, vc, xout, bsrc, bout, od, od_mask,  \
                                     rt, ut, xref, v)
        if (od && xprev) CG_RESIDUAL(ExtrapLoad, XE, true);
        else if (od) CG_RESIDUAL(PlainLoad<double>, X0, true);
        else if (xprev) CG_RESIDUAL(ExtrapLoad, XE, false);
        else if (add_e) CG_RESIDUAL(CorrectedLoad, (CorrectedLoad{xin, e}), false);
        else CG_RESIDUAL(PlainLoad<double>, X0, false);
#undef CG_RESIDUAL
    }
    block_sum<RBS, 6>(v);
    if (threadIdx.x == 0) {
        part[(P_G + 0) * RG + bid] = v[0];
        part[(P_RR + 0) * RG + bid] = v[1];
        part[P_RR0 * RG + bid] = v[2];
        part[P_BB * RG + bid] = v[3];
        if (OF) {
            part[(P_L1 + 0) * RG + bid] = v[4];
            if (of_nf) part[P_NF * RG + bid] = v[5];
        }
        if (blockIdx.x == 0) {
            S->iters_sum += S->iters;
            S->iters = 0;
        }
    }
}

// ---- S-kernel of iteration k of solve gen: convergence test on r_k, then w = A u_k and
// the partial (w,u).  Converged when (r,r) <= tol2 * ref, ref = (r0,r0) (relative_to_r0)
// or (b,b), or when (r,r) <= abs2 (b,b) (an inner solve that already meets the step
// tolerance); at k == 0 with outer_tol2 > 0 also when the fp64 residual already meets
// (r0,r0) <= outer_tol2 (b,b) (the refinement is finished).
//
// NORM 0: the tests above in relative L2.
// NORM 1 (fp64 CG in OpenFOAM's norm): converged when sum|r| < of_tol nf, or
//   of_rel > 1e-20 and sum|r| < of_rel sum|r_0| (SolverPerformance::checkConvergence).
// NORM 2 (mixed, OpenFOAM's norm): the k == 0 outer test is sum|r| < of_tol nf on the fp64
//   residual; the inner test stays relative L2.
// record_init (k == 0 of an outer corrector's first solve): S->init_res = the initial
// residual, sum|r|/nf (NORM 1, 2) or ||r||/||b|| (NORM 0).
template <class T, bool QUAD, int NORM>
__global__ __launch_bounds__(RBS) void k_cg_s(HexGeom g, HexOp<T> A, const T* __restrict__ u,
                                              T* __restrict__ w, double* __restrict__ part,
                                              DevState* S, long long gen, int k, double tol2,
                                              int relative_to_r0, double abs2,
                                              double outer_tol2, int max_iter, double of_tol,
                                              double of_rel, int record_init, DevState* pub)
{
    // pub (the host mirror, or null): block 0 copies the state there once it is final for
    // this launch, one 8-byte word per lane
    auto publish = [&]() {
        if (pub && blockIdx.x == 0 && threadIdx.x < sizeof(DevState) / 8)
            reinterpret_cast<volatile unsigned long long*>(pub)[threadIdx.x] =
                reinterpret_cast<const volatile unsigned long long*>(S)[threadIdx.x];
    };
    if (S->done_gen == gen) {
        publish();
        return;
    }
    double rr, rr0, bb, l1 = 0.0, nf = 1.0;
    if (NORM == 1 || (NORM == 2 && k == 0)) {
        double v5[5];
        sum_slots<5>(part, {P_RR + (k & 1), P_RR0, P_BB, P_L1 + (k & 1), P_NF}, v5);
        rr = v5[0]; rr0 = v5[1]; bb = v5[2]; l1 = v5[3]; nf = v5[4] + 1e-20;
    } else {
        double v3[3];
        sum_slots<3>(part, {P_RR + (k & 1), P_RR0, P_BB}, v3);
        rr = v3[0]; rr0 = v3[1]; bb = v3[2];
    }
    bool outer, conv;
    if (NORM == 1) {
        const double l10 = k == 0 ? l1 : S->l1_0;
        outer = false;
        conv = l1 < of_tol * nf || (of_rel > 1e-20 && l1 < of_rel * l10);
    } else {
        outer = k == 0 && outer_tol2 > 0.0 &&
                (NORM == 2 ? l1 < of_tol * nf : rr0 <= outer_tol2 * bb);
        conv = rr <= tol2 * (relative_to_r0 ? rr0 : bb) || rr <= abs2 * bb;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (k == 0) {
            S->s[S_R0] = rr0;
            S->s[S_BB] = bb;
            if (NORM != 0) {
                S->nf = nf;
                S->l1_0 = l1;
            }
            if (record_init)
                S->init_res = NORM != 0 ? l1 / nf : sqrt(rr0 / (bb > 0.0 ? bb : 1.0));
        }
        if (outer || conv || k >= max_iter) {
            S->s[S_RR] = rr;
            S->l1 = l1;
            S->iters = k;
            if (outer) S->outer_gen = gen;
            if (outer || (conv && !relative_to_r0)) S->conv_gen = gen;
            S->done_gen = gen;
        }
    }
    if (pub && blockIdx.x == 0) {
        __syncthreads();   // thread 0's updates of S before the copy by the block's lanes
        publish();
    }
    if (outer || conv || k >= max_iter) return;
    double v[1] = {0.0};
    if constexpr (QUAD) {
        const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
        RANGE_LOOP_Q(g)
        {
            const int64_t c0 = qi * 4;
            const V4<T> y = apply_quad(A, u, c0, ijk_of((uint32_t)c0, g), g);
            st4(w + c0, y);
            v[0] += dot4(ld4(u + c0), y);
        }
    } else {
        const PlainLoad<T> U{u};
        RANGE_LOOP(g)
        {
            const IJK ix = ijk_of((uint32_t)c, g);
            const T y = apply_row(A, U, c, ix, g);
            w[c] = y;
            v[0] += double(u[c]) * double(y);
        }
    }
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    block_sum<RBS, 1>(v);
    if (threadIdx.x == 0) part[P_D * RG + bid] = v[0];
}

// ---- U-kernel of iteration k: gamma_k = (r,u), delta_k = (w,u);
//   beta = gamma_k / gamma_{k-1} (0 at k = 0),
//   alpha = gamma_k / (delta_k - beta gamma_k / alpha_{k-1})   (gamma/delta at k = 0);
//   p = u + beta p, s = w + beta s, x += alpha p (x = alpha p at k = 0 when x_from_zero),
//   r -= alpha s, u = M^-1 r; partials (r,u), (r,r) of iteration k+1.
//   L1: also the partial sum |r| of iteration k+1 (OpenFOAM norm, fp64 path).
template <class T, class TX, bool QUAD, bool L1>
__global__ __launch_bounds__(RBS) void k_cg_u(HexGeom g, const T* __restrict__ invd,
                                              TX* __restrict__ x, T* __restrict__ r,
                                              T* __restrict__ u, T* __restrict__ p,
                                              T* __restrict__ s, const T* __restrict__ w,
                                              double* __restrict__ part, DevState* S,
                                              long long gen, int k, int x_from_zero)
{
    if (S->done_gen == gen) return;
    double v2[2];
    sum_slots<2>(part, {P_G + (k & 1), P_D}, v2);
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
        S->iters = k + 1;
    }
    const T a = T(alpha), bt = T(beta);
    const bool first = k == 0;
    const bool xz = first && x_from_zero;
    double v[3] = {0.0, 0.0, 0.0};
    if constexpr (QUAD) {
        const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
        RANGE_LOOP_Q(g)
        {
            const int64_t c0 = qi * 4;
            const V4<T> u4 = ld4(u + c0), w4 = ld4(w + c0);
            const V4<T> pc = first ? u4 : u4 + bt * ld4(p + c0);
            const V4<T> sc = first ? w4 : w4 + bt * ld4(s + c0);
            st4(p + c0, pc);
            st4(s + c0, sc);
            const V4<TX> ap = __builtin_convertvector(a * pc, V4<TX>);
            st4(x + c0, xz ? ap : ld4(x + c0) + ap);
            const V4<T> rn = ld4(r + c0) - a * sc;
            const V4<T> un = ld4(invd + c0) * rn;
            st4(r + c0, rn);
            st4(u + c0, un);
            v[0] += dot4(rn, un);
            v[1] += dot4(rn, rn);
            if (L1)
                v[2] += (fabs(double(rn.x)) + fabs(double(rn.y))) +
                        (fabs(double(rn.z)) + fabs(double(rn.w)));
        }
    } else {
        RANGE_LOOP(g)
        {
            const T pc = first ? u[c] : u[c] + bt * p[c];
            const T sc = first ? w[c] : w[c] + bt * s[c];
            p[c] = pc;
            s[c] = sc;
            x[c] = xz ? TX(a * pc) : TX(x[c] + TX(a * pc));
            const T rn = r[c] - a * sc;
            const T un = invd[c] * rn;
            r[c] = rn;
            u[c] = un;
            v[0] += double(rn) * double(un);
            v[1] += double(rn) * double(rn);
            if (L1) v[2] += fabs(double(rn));
        }
    }
    if constexpr (L1) block_sum<RBS, 3>(v);
    else {
        double v2[2] = {v[0], v[1]};
        block_sum<RBS, 2>(v2);
        v[0] = v2[0];
        v[1] = v2[1];
    }
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    if (threadIdx.x == 0) {
        part[(P_G + ((k + 1) & 1)) * RG + bid] = v[0];
        part[(P_RR + ((k + 1) & 1)) * RG + bid] = v[1];
        if (L1) part[(P_L1 + ((k + 1) & 1)) * RG + bid] = v[2];
    }
}

// ---- activation time (myocardiumDomain.C:345-374) and V^{n+1} = x when x is elsewhere;
// with a gate it runs only when the solve of that generation met the step tolerance
__global__ __launch_bounds__(BS) void k_activation(HexGeom g, const double* __restrict__ Vn,
                                                   const double* x,  // may be Vnew
                                                   double* Vnew,
                                                   uint8_t* __restrict__ flag,
                                                   double* __restrict__ tact, double t_old,
                                                   double dt, double vthr,
                                                   const long long* gate, long long gate_val)
{
    if (gate && *gate != gate_val) return;
    CELL_INDEX
    (void)cc;
    const double vnew = x[c];
    if (x != Vnew) Vnew[c] = vnew;
    if (flag[c] && vnew > vthr + 1e-15) {
        const double vn = Vn[c];
        flag[c] = 0;
        tact[c] = t_old + (vthr - vn) / (vnew - vn) * dt;
    }
}

template <class T>
__global__ __launch_bounds__(BS) void k_apply4(HexGeom g, HexOp<T> A, const T* __restrict__ x,
                                               T* __restrict__ y)
{
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t c0 = ((int64_t)bid * BS + threadIdx.x) * 4;
    if (c0 >= n) return;
    const IJK q = ijk_of((uint32_t)c0, g);
    *reinterpret_cast<V4<T>*>(y + c0) = apply_quad(A, x, c0, q, g);
}

// ---- y = A x with the implicit operator (one row per lane, no reduction)
template <class T>
__global__ __launch_bounds__(BS) void k_apply(HexGeom g, HexOp<T> A, const T* __restrict__ x,
                                              T* __restrict__ y)
{
    CELL_INDEX
    y[c] = apply_row(A, PlainLoad<T>{x}, c, cc, g);
}

// ---- explicit Laplacian (operator check)
__global__ __launch_bounds__(BS) void k_laplacian(HexGeom g, HexCoefs a,
                                                  const double* __restrict__ V,
                                                  const double* __restrict__ od,
                                                  unsigned od_mask, double* __restrict__ out)
{
    CELL_INDEX
    const double acc0 = a.fx ? hex_laplacian<false>(V, c, cc, g, a) : hex_laplacian<true>(V, c, cc, g, a);
    double acc = acc0;
    if (od) acc += hex_tensor_correction(V, od, n, c, cc.i, cc.j, cc.k, g, od_mask);
    out[c] = acc;
}

const char* kernel_name(int id)
{
    static const char* names[K_COUNT] = {"ionic",    "spmv_dot",   "cg_update", "residual_first",
                                         "residual", "activation", "invdiag",    "apply",
                                         "pseudo_ecg"};
    return id >= 0 && id < K_COUNT ? names[id] : "?";
}

static unsigned grid_of(int64_t n)
{
    int64_t nb = (n + BS - 1) / BS;
    nb = (nb + 7) / 8 * 8;  // multiple of 8 for the XCD remap
    return (unsigned)nb;
}

hipError_t launch_hex_laplacian(const HexGeom& g, const HexCoefs& k, const double* V,
                                const double* od, unsigned od_mask, double* out, hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_laplacian, dim3(grid_of(n)), dim3(BS), 0, s, g, k, V, od, od_mask, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ solver object
template <class T> struct CgVectors {
    T *r = nullptr, *u = nullptr, *p = nullptr, *s = nullptr, *w = nullptr;
};

struct ImplicitSolver {
    HexGeom g;
    int64_t n;
    int prec;
    double chiCm, dt;
    hipStream_t stream;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr, *invd = nullptr;  // fp64 operator
    float *fx32 = nullptr, *fy32 = nullptr, *fz32 = nullptr, *invd32 = nullptr;
    double* od = nullptr;                                                  // sigma xy xz yz
    unsigned od_mask = 0;
    double* b = nullptr;
    double* xalt = nullptr;  // second fp64 solution buffer of the refinement (mixed)
    CgVectors<double> v64;
    CgVectors<float> v32;
    float* e32 = nullptr;
    double* partials = nullptr;
    DevState* S = nullptr;
    DevState* hS = nullptr;  // pinned, coherent host mirror of S
    DevState* hS_dev = nullptr;      // the mirror's device address: the pass's last S-kernel
                                     // writes S there, so no copy kernel runs
    hipEvent_t read_done = nullptr;  // after the kernel that published S
    double mass_c = -1.0;    // coefft of the current invd
    bool quad = false;       // nx % 4 == 0: CG kernels on quads of cells (vector loads)
    long long gen = 0;       // generation of the last CG solve started
    int64_t total_iters = 0;
    double last_res = 0.0;
    int last_inner = 16;
    long long iters_sum_seen = 0;  // DevState::iters_sum at the previous read
    SolveControls ctl;
    double* bc = nullptr;    // b plus the correction of the current outer corrector (mixed)
    int last_solves = 0;
    double last_init = 0.0, last_final = 0.0;
};

int implicit_set_controls(ImplicitSolver* S, const SolveControls& c)
{
    if (c.norm != 0 && c.norm != 1) return 1;
    S->ctl = c;
    if (c.n_outer > 1 && S->prec == 1 && S->od && !S->bc &&
        hipMalloc((void**)&S->bc, 8 * S->n) != hipSuccess)
        return 5;
    return 0;
}

void implicit_report(const ImplicitSolver* S, int* solves, double* init_res, double* final_res)
{
    if (solves) *solves = S->last_solves;
    if (init_res) *init_res = S->last_init;
    if (final_res) *final_res = S->last_final;
}

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
    S->quad = g.nx % 4 == 0;
    const int64_t n = S->n;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    std::vector<double> fx(n, 0.0), fy(n, 0.0), fz(n, 0.0), od(3 * n, 0.0);
    for (int64_t c = 0; c < n; ++c) {
        const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / sz);
        if (i < g.nx - 1) fx[c] = 0.5 * (sig[9 * c + 0] + sig[9 * (c + 1) + 0]) / (g.hx * g.hx);
        if (j < g.ny - 1) fy[c] = 0.5 * (sig[9 * c + 4] + sig[9 * (c + sy) + 4]) / (g.hy * g.hy);
        if (k < g.nz - 1) fz[c] = 0.5 * (sig[9 * c + 8] + sig[9 * (c + sz) + 8]) / (g.hz * g.hz);
        od[c] = sig[9 * c + 1];
        od[n + c] = sig[9 * c + 2];
        od[2 * n + c] = sig[9 * c + 5];
        for (int q = 0; q < 3; ++q)
            if (od[q * n + c] != 0.0) S->od_mask |= 1u << q;
    }
    const bool offdiag = S->od_mask != 0;
    bool ok = true;
    auto A = [&](void** ptr, size_t bytes) { ok &= hipMalloc(ptr, bytes) == hipSuccess; };
    A((void**)&S->fx, 8 * n); A((void**)&S->fy, 8 * n); A((void**)&S->fz, 8 * n);
    A((void**)&S->invd, 8 * n);
    A((void**)&S->b, 8 * n);
    A((void**)&S->partials, 8 * RG * P_NSLOT);
    A((void**)&S->S, sizeof(DevState));
    if (offdiag) A((void**)&S->od, 8 * 3 * n);
    if (prec == 1) {
        A((void**)&S->fx32, 4 * n); A((void**)&S->fy32, 4 * n); A((void**)&S->fz32, 4 * n);
        A((void**)&S->invd32, 4 * n);
        A((void**)&S->v32.r, 4 * n); A((void**)&S->v32.u, 4 * n); A((void**)&S->v32.p, 4 * n);
        A((void**)&S->v32.s, 4 * n); A((void**)&S->v32.w, 4 * n);
        A((void**)&S->e32, 4 * n);
        A((void**)&S->xalt, 8 * n);
    } else {
        A((void**)&S->v64.r, 8 * n); A((void**)&S->v64.u, 8 * n); A((void**)&S->v64.p, 8 * n);
        A((void**)&S->v64.s, 8 * n); A((void**)&S->v64.w, 8 * n);
    }
    ok &= hipHostMalloc((void**)&S->hS, sizeof(DevState),
                        hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess;
    ok &= S->hS && hipHostGetDevicePointer((void**)&S->hS_dev, S->hS, 0) == hipSuccess;
    ok &= hipEventCreateWithFlags(&S->read_done, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        implicit_destroy(S);
        *err = "implicit solver allocation failed";
        return 1;
    }
    hipMemcpy(S->fx, fx.data(), 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(S->fy, fy.data(), 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(S->fz, fz.data(), 8 * n, hipMemcpyHostToDevice);
    if (offdiag) hipMemcpy(S->od, od.data(), 8 * 3 * n, hipMemcpyHostToDevice);
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
    z.done_gen = -1;
    z.outer_gen = -1;
    z.conv_gen = -1;
    hipMemcpy(S->S, &z, sizeof z, hipMemcpyHostToDevice);
    *out = S;
    return 0;
}

void implicit_destroy(ImplicitSolver* S)
{
    if (!S) return;
    void* ptrs[] = {S->bc,     S->fx,     S->fy,     S->fz,     S->invd,     S->b,        S->xalt,
                    S->partials, S->S,    S->od,     S->fx32,     S->fy32,     S->fz32,
                    S->invd32, S->v32.r,  S->v32.u,  S->v32.p,    S->v32.s,    S->v32.w,
                    S->e32,    S->v64.r,  S->v64.u,  S->v64.p,    S->v64.s,    S->v64.w};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (S->hS) hipHostFree(S->hS);
    if (S->read_done) hipEventDestroy(S->read_done);
    delete S;
}

void implicit_stats(const ImplicitSolver* S, int64_t* iters, double* res)
{
    if (iters) *iters = S->total_iters;
    if (res) *res = S->last_res;
}

int implicit_apply(ImplicitSolver* S, int prec, double coefft, const void* x, void* y,
                   Profiler* prof)
{
    if (prec == 1 && !S->fx32) return 9;
    const int tk = prof ? prof->begin(K_APPLY) : -1;
    // quads need whole x-rows of multiples of 4 and 16/32-byte aligned vectors
    const int elt = prec == 0 ? 8 : 4;
    const bool quad = S->g.nx % 4 == 0 && ((uintptr_t)x % (4 * elt)) == 0 &&
                      ((uintptr_t)y % (4 * elt)) == 0;
    const unsigned grid = quad ? grid_of((S->n + 3) / 4) : grid_of(S->n);
    if (prec == 0 && quad)
        hipLaunchKernelGGL((k_apply4<double>), dim3(grid), dim3(BS), 0, S->stream, S->g,
                           op_of<double>(S, coefft), (const double*)x, (double*)y);
    else if (prec == 0)
        hipLaunchKernelGGL((k_apply<double>), dim3(grid), dim3(BS), 0, S->stream, S->g,
                           op_of<double>(S, coefft), (const double*)x, (double*)y);
    else if (quad)
        hipLaunchKernelGGL((k_apply4<float>), dim3(grid), dim3(BS), 0, S->stream, S->g,
                           op_of<float>(S, coefft), (const float*)x, (float*)y);
    else
        hipLaunchKernelGGL((k_apply<float>), dim3(grid), dim3(BS), 0, S->stream, S->g,
                           op_of<float>(S, coefft), (const float*)x, (float*)y);
    if (prof) prof->end(tk);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

namespace {

// launch bookkeeping for the per-kernel timings: launches that exited on a finished
// solve are dropped once the host knows how many iterations ran
struct CgLaunch {
    size_t rec;
    long long gen;
    int k;
    bool s_kernel;
};

template <class T> struct CgRun {
    ImplicitSolver* S;
    HexOp<T> A;
    CgVectors<T> v;
    Profiler* prof;
    std::vector<CgLaunch>* log;

    // norm of the convergence tests: 0 relative L2; 1 OpenFOAM's on the fp64 CG; 2
    // OpenFOAM's on the outer test of the refinement (mixed)
    int norm_kind() const { return S->ctl.norm == 0 ? 0 : (sizeof(T) == 8 ? 1 : 2); }

    // publish: the launch that ends a pass before the host reads the state
    void s_kernel(long long gen, int k, double tol2, int rel, double abs2, double outer_tol2,
                  int max_iter, double of_tol = 0.0, int record_init = 0, bool publish = false)
    {
        const int tk = prof ? prof->begin(K_SPMV) : -1;
        const double of_rel = S->ctl.rel_tol;
        DevState* pub = publish ? S->hS_dev : nullptr;
#define CG_S(Q, N)                                                                         \
    hipLaunchKernelGGL((k_cg_s<T, Q, N>), dim3(RG), dim3(RBS), 0, S->stream, S->g, A,      \
                       (const T*)v.u, v.w, S->partials, S->S, gen, k, tol2, rel, abs2,     \
                       outer_tol2, max_iter, of_tol, of_rel, record_init, pub)
        switch (norm_kind() + (S->quad ? 3 : 0)) {
        case 0: CG_S(false, 0); break;
        case 1: CG_S(false, 1); break;
        case 2: CG_S(false, 2); break;
        case 3: CG_S(true, 0); break;
        case 4: CG_S(true, 1); break;
        default: CG_S(true, 2); break;
        }
#undef CG_S
        if (prof) {
            prof->end(tk);
            log->push_back({(size_t)tk, gen, k, true});
        }
    }
    template <class TX> void u_kernel(long long gen, int k, TX* x, int x_from_zero)
    {
        const int tk = prof ? prof->begin(K_UPDATE) : -1;
        const bool l1 = norm_kind() == 1;
#define CG_U(Q, L)                                                                         \
    hipLaunchKernelGGL((k_cg_u<T, TX, Q, L>), dim3(RG), dim3(RBS), 0, S->stream, S->g,     \
                       A.invd, x, v.r, v.u, v.p, v.s, (const T*)v.w, S->partials, S->S,    \
                       gen, k, x_from_zero)
        if (S->quad) {
            if (l1) CG_U(true, true);
            else CG_U(true, false);
        } else {
            if (l1) CG_U(false, true);
            else CG_U(false, false);
        }
#undef CG_U
        if (prof) {
            prof->end(tk);
            log->push_back({(size_t)tk, gen, k, false});
        }
    }
    // the sum of the initial guess (xin, or 2 xin - xprev) for OpenFOAM's xRef
    void mean(const double* xin, const double* xprev)
    {
        if (xprev)
            hipLaunchKernelGGL((k_mean<ExtrapLoad>), dim3(RG), dim3(RBS), 0, S->stream, S->g,
                               ExtrapLoad{xin, xprev}, S->partials);
        else
            hipLaunchKernelGGL((k_mean<PlainLoad<double>>), dim3(RG), dim3(RBS), 0, S->stream,
                               S->g, PlainLoad<double>{xin}, S->partials);
    }
    void residual(int kid, long long gen, const double* xin, const double* xprev,
                  const float* e, double* xout, const double* od, const HexOp<double>& A64,
                  const double* bsrc, double* bout, int of_nf, const double* vcor = nullptr)
    {
        const int tk = prof ? prof->begin(kid) : -1;
        // quads need nx % 4 == 0 and no correction beyond sigma_xy
        const bool quad = S->quad && (!od || S->od_mask == 1u);
#define CG_R(OF, Q, NF)                                                                    \
    hipLaunchKernelGGL((k_residual<T, OF, Q>), dim3(RG), dim3(RBS), 0, S->stream, S->g, A64, \
                       (const T*)A.invd, xin, vcor, xprev, e, xout, bsrc, bout, od, S->od_mask, \
                       v.r,                                                                \
                       v.u, S->partials, S->S, gen, NF)
        if (S->ctl.norm == 1) {
            if (quad) CG_R(true, true, of_nf);
            else CG_R(true, false, of_nf);
        } else {
            if (quad) CG_R(false, true, 0);
            else CG_R(false, false, 0);
        }
#undef CG_R
        if (prof) prof->end(tk);
    }
};

// drop the timing records of launches of solve `gen` beyond its `iters` iterations
// (S-kernels with k >= iters only tested convergence, U-kernels with k >= iters did nothing)
void drop_idle(Profiler* prof, std::vector<CgLaunch>& log, long long gen, int iters)
{
    if (!prof) return;
    for (auto& l : log)
        if (l.gen == gen && l.k >= iters) prof->invalidate_range(l.rec, l.rec + 1);
}

} // namespace

int implicit_step(ImplicitSolver* S, int model, int prec, int tissue, void* states,
                  const ModelParams& P, const double* Vn, const double* Vm1, double* Vnew,
                  double* Iion, uint8_t* flag, double* tact, const ExplicitStep& st, double tol,
                  int max_iter, Profiler* prof, bool ionic_queued, const NextIonic* next,
                  bool extrapolate, std::string* err)
{
    const unsigned grid = grid_of(S->n);
    int tk;
    hipError_t e;
    // With one corrector per step the ionic kernel (mixed path, extrapolating) also writes
    // the initial guess 2 V^n - V^{n-1} into xalt, so the first residual reads one x.
    // Outer correctors keep the extrapolation in the residual.
    const bool fuse = S->ctl.n_outer <= 1;
    const double* od_res = S->od;
    const bool x0_fused = fuse && prec == 1 && extrapolate;
    if (!ionic_queued) {
        tk = prof ? prof->begin(K_IONIC) : -1;
        e = launch_ionic_rhs(model, prec, tissue, S->g, states, Vn, Vm1,
                             x0_fused ? S->xalt : nullptr, Iion, S->b, st, P, S->stream);
        if (prof) prof->end(tk);
        if (e != hipSuccess) { *err = std::string("ionic/rhs launch: ") + hipGetErrorString(e); return 4; }
    }
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
    const HexOp<double> A = op_of<double>(S, st.c);
    const double inner_tol = 1e-4;
    std::vector<CgLaunch> log;
    int used_total = 0;
    const long long* conv_gen = &S->S->conv_gen;
    const bool of = S->ctl.norm == 1;
    const int n_outer = S->ctl.n_outer > 1 ? S->ctl.n_outer : 1;

    // Before each read: the step's activation (x into V^{n+1}) and the next step's ionic
    // kernel, both gated on "solve g met the tolerance".  The pass's last S-kernel has
    // already written the state into the host mirror; the host waits for the event behind
    // it only, and the queued ionic kernel of the next step keeps the device busy
    // meanwhile.  Gated launches of a pass that did not converge do nothing and are dropped
    // from the timings.  Outer correctors before the last one read back without the gated
    // launches (the step is not finished).
    auto finish_and_read = [&](long long g, const double* x, size_t* mark, bool last) -> bool {
        *mark = prof ? prof->mark() : 0;
        // with the solution already in V^{n+1}, the next step's ionic kernel does the
        // activation update (nothing to copy); otherwise k_activation copies and updates
        const bool act_in_next = last && next && x == Vnew;
        if (last && !act_in_next) {
            int t = prof ? prof->begin(K_ACTIVATION) : -1;
            hipLaunchKernelGGL(k_activation, dim3(grid), dim3(BS), 0, S->stream, S->g, Vn, x,
                               Vnew, flag, tact, st.t_old, st.dt, st.act_v, conv_gen, g);
            if (prof) prof->end(t);
        }
        hipEventRecord(S->read_done, S->stream);
        if (last && next) {
            // the next step (which extrapolates) gets the same fusion
            ExplicitStep sn = next->st;
            sn.gate = conv_gen;
            sn.gate_val = g;
            if (act_in_next) {
                sn.act_flag = flag;
                sn.act_tact = tact;
                sn.act_t_old = st.t_old;
                sn.act_dt = st.dt;
                sn.act_v = st.act_v;
            }
            int t = prof ? prof->begin(K_IONIC) : -1;
            launch_ionic_rhs(model, prec, tissue, S->g, states, Vnew, Vn,
                             fuse && prec == 1 ? S->xalt : nullptr, Iion, S->b, sn, P,
                             S->stream);
            if (prof) prof->end(t);
        }
        if (hipGetLastError() != hipSuccess) return false;
        return hipEventSynchronize(S->read_done) == hipSuccess;
    };

    // PIMPLE outer correctors (monodomainSolver.C:139-150): each corrector re-evaluates the
    // explicit tensor correction from the latest V and solves from it; after the first, a
    // previous initial residual below residualControl makes the corrector the last one
    double prev_init = 0.0;
    int solves = 0;
    if (prec == 0) {
        // fp64 CG per corrector; x0 = V^n (copied into V^{n+1} by the first residual)
        CgRun<double> cg{S, A, S->v64, prof, &log};
        for (int outer = 1;; ++outer) {
            const bool last = n_outer == 1 || outer >= n_outer ||
                              (outer > 1 && prev_init < S->ctl.residual_control);
            const long long gen = ++S->gen;
            const double* xin = outer == 1 ? Vn : Vnew;
            if (of) cg.mean(xin, nullptr);
            cg.residual(K_RESIDUAL0, gen, xin, nullptr, nullptr, outer == 1 ? Vnew : nullptr,
                        outer == 1 ? od_res : S->od, A, S->b, nullptr, 1);
            cg.s_kernel(gen, 0, tol * tol, 0, 0.0, 0.0, max_iter, tol, 1);
            int k = 0;
            int chunk = S->last_inner + 1 > 2 ? S->last_inner + 1 : 2;
            for (;;) {
                for (int q = 0; q < chunk; ++q, ++k) {
                    cg.u_kernel(gen, k, Vnew, 0);
                    cg.s_kernel(gen, k + 1, tol * tol, 0, 0.0, 0.0, max_iter, tol, 0,
                                q == chunk - 1);
                }
                size_t gated;
                if (!finish_and_read(gen, Vnew, &gated, last)) {
                    *err = "CG launch or read-back failed";
                    return 4;
                }
                const DevState& h = *S->hS;
                if (h.done_gen == gen) {
                    drop_idle(prof, log, gen, h.iters);
                    const bool ok = h.conv_gen == gen;
                    if (prof) {
                        if (!ok && last) prof->invalidate_range(gated, prof->mark());
                        prof->resolve(false);
                    }
                    used_total += h.iters;
                    S->last_inner = h.iters > 0 ? h.iters : S->last_inner;
                    const double bb = h.s[S_BB] > 0 ? h.s[S_BB] : 1.0;
                    S->last_res = of ? h.l1 / h.nf : std::sqrt(h.s[S_RR] / bb);
                    prev_init = h.init_res;
                    if (!ok) {
                        *err = "PCG did not converge within max_iter";
                        return 7;
                    }
                    break;
                }
                if (prof) prof->invalidate_range(gated, prof->mark());
                if (k > max_iter + 64) {   // the device stops a solve at max_iter
                    *err = "solver state read-back did not report the end of the solve";
                    return 4;
                }
                chunk = 4;  // not finished: continue the same solve
            }
            log.clear();
            ++solves;
            if (last) break;
        }
    } else {
        // mixed: refinement passes per corrector; x starts in V^n or, extrapolated, in xalt,
        // then alternates Vnew / xalt.  An inner solve stops at 1e-4 relative or (relative
        // L2 only) as soon as its residual meets half the step tolerance.
        CgRun<float> cg{S, op_of<float>(S, st.c), S->v32, prof, &log};
        const double abs2 = of ? 0.0 : 0.25 * tol * tol;
        // b plus the correction lives in b itself unless later correctors need b again
        double* bcur = od_res ? (n_outer > 1 ? S->bc : S->b) : nullptr;
        const double* bref = bcur ? bcur : S->b;
        const double* cur = Vn;
        for (int outer = 1;; ++outer) {
            const bool last = n_outer == 1 || outer >= n_outer ||
                              (outer > 1 && prev_init < S->ctl.residual_control);
            long long gen = ++S->gen;
            if (outer == 1 && x0_fused) {
                // x0 = 2 V^n - V^{n-1} already in xalt (ionic kernel), so the first
                // refinement lands in V^{n+1} itself
                if (of) cg.mean(S->xalt, nullptr);
                cg.residual(K_RESIDUAL0, gen, S->xalt, nullptr, nullptr, nullptr, od_res, A,
                            S->b, bcur, 1, Vn);
                cur = S->xalt;
            } else if (outer == 1) {
                // x0 = 2 V^n - V^{n-1} (V^n on the first step), written to xalt
                const double* xprev = extrapolate ? Vm1 : nullptr;
                if (of) cg.mean(Vn, xprev);
                cg.residual(K_RESIDUAL0, gen, Vn, xprev, nullptr, xprev ? S->xalt : nullptr,
                            od_res, A, S->b, bcur, 1);
                cur = xprev ? S->xalt : Vn;
            } else {
                if (of) cg.mean(cur, nullptr);
                cg.residual(K_RESIDUAL0, gen, cur, nullptr, nullptr, nullptr, S->od, A, S->b,
                            bcur, 1);
            }
            cg.s_kernel(gen, 0, inner_tol * inner_tol, 1, abs2, tol * tol, max_iter, tol, 1);
            bool init_read = false;
            for (int pass = 0;; ++pass) {
                const int chunk = S->last_inner + 1 > 2 ? S->last_inner + 1 : 2;
                for (int k = 0; k < chunk; ++k) {
                    cg.u_kernel(gen, k, S->e32, 1);
                    cg.s_kernel(gen, k + 1, inner_tol * inner_tol, 1, abs2, tol * tol, max_iter);
                }
                // x' = x + e into the other buffer, fp64 residual, start of the next solve
                double* nxt = cur == Vnew ? S->xalt : Vnew;
                const long long ngen = ++S->gen;
                cg.residual(K_RESIDUAL, ngen, cur, nullptr, S->e32, nxt, nullptr, A, bref,
                            nullptr, 0);
                cg.s_kernel(ngen, 0, inner_tol * inner_tol, 1, abs2, tol * tol, max_iter, tol, 0,
                            true);
                cur = nxt;
                size_t gated;
                if (!finish_and_read(ngen, cur, &gated, last)) {
                    *err = "CG launch or read-back failed";
                    return 4;
                }
                const DevState& h = *S->hS;
                if (!init_read) {  // recorded by the corrector's first solve
                    prev_init = h.init_res;
                    init_read = true;
                }
                // iterations of solve gen, folded into the running sum by the residual pass
                const int run = (int)(h.iters_sum - S->iters_sum_seen);
                S->iters_sum_seen = h.iters_sum;
                drop_idle(prof, log, gen, run);
                const bool finished = h.outer_gen == ngen;
                if (finished) drop_idle(prof, log, ngen, 0);
                else if (prof && last) prof->invalidate_range(gated, prof->mark());
                if (prof) prof->resolve(false);
                log.clear();
                used_total += run;
                if (run > 0) S->last_inner = run;
                const double bb = h.s[S_BB] > 0 ? h.s[S_BB] : 1.0;
                S->last_res = of ? h.l1_0 / h.nf : std::sqrt(h.s[S_R0] / bb);
                if (finished) break;
                if (used_total >= max_iter * n_outer || pass >= 50) {
                    *err = "mixed-precision refinement did not converge";
                    return 7;
                }
                gen = ngen;
            }
            ++solves;
            if (last) break;
        }
    }
    S->last_solves = solves;
    S->last_init = prev_init;
    S->last_final = S->last_res;
    S->total_iters += used_total;
    e = hipGetLastError();
    if (e != hipSuccess) { *err = hipGetErrorString(e); return 4; }
    return 0;
}

} // namespace cg
