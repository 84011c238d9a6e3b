// Adaptive embedded Runge-Kutta-Fehlberg 4(5) per cell, the integrator the reference
// selects with `solver RKF45` in the ionic model's dictionary (ionicModel.H:233-242;
// called from BuenoOrovio.C:132, TNNP.C:143, ORd.C:131, AlievPanfilov.C:132).
// The algorithm is OpenFOAM's (ESI v2312..v2512, not part of the reference tree):
// ODESolver::solve(xStart, xEnd, y, dxTry) driving adaptiveSolver::solve with the RKF45
// stage scheme, restated here from the published Fehlberg method and that controller:
//   - Fehlberg 4(5) tableau; the fifth-order solution is propagated and the error is its
//     difference to the fourth-order one;
//   - error norm max_i |err_i| / (absTol + relTol max(|y0_i|, |y_i|));
//   - a step is retried while err > 1, dx scaled by max(0.9 err^-1/4, 0.2); the
//     derivative at the step start is reused across retries;
//   - after acceptance the next trial is min(0.9 err^-1/5, 10) dx, or 9 dx when
//     err <= (10/0.9)^-5;
//   - the step that would pass xEnd is truncated to land on it, and the trial size from
//     before the truncation is handed back (the per-cell `step` the model keeps).
// OpenFOAM's `Euler` ODESolver (`solver Euler;`) runs under the same controller with a
// one-stage step: y = y0 + dx dy/dx(x0), and the whole increment dx dy/dx(x0) as the
// error estimate, so each accepted step changes no state by more than its tolerance.
// One lane integrates one cell; lanes of a wave run different numbers of sub-steps.
// METHOD 1: RKF45; 2: Euler.
#pragma once
#include "cg_common.hpp"

namespace cg {

// f(t, y, dydt): right-hand side at model time t.  ref (optional): y holds offsets from
// ref (the FP32 offset storage); the error norm is taken on the values ref + y.
template <int NS, class T, class F>
CG_HD double rkf45_trial(F& f, double x0, const T (&y0)[NS], const T (&d0)[NS], double dxd,
                          T (&y)[NS], T at, T rt, const T* ref = nullptr)
{
    const T dx = T(dxd);
    T yt[NS], k2[NS], k3[NS], k4[NS], k5[NS], k6[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) yt[i] = y0[i] + T(1.0 / 4.0) * dx * d0[i];
    f(x0 + (1.0 / 4.0) * dxd, yt, k2);
#pragma unroll
    for (int i = 0; i < NS; ++i)
        yt[i] = y0[i] + dx * (T(3.0 / 32.0) * d0[i] + T(9.0 / 32.0) * k2[i]);
    f(x0 + (3.0 / 8.0) * dxd, yt, k3);
#pragma unroll
    for (int i = 0; i < NS; ++i)
        yt[i] = y0[i] + dx * (T(1932.0 / 2197.0) * d0[i] + T(-7200.0 / 2197.0) * k2[i] +
                              T(7296.0 / 2197.0) * k3[i]);
    f(x0 + (12.0 / 13.0) * dxd, yt, k4);
#pragma unroll
    for (int i = 0; i < NS; ++i)
        yt[i] = y0[i] + dx * (T(439.0 / 216.0) * d0[i] + T(-8.0) * k2[i] +
                              T(3680.0 / 513.0) * k3[i] + T(-845.0 / 4104.0) * k4[i]);
    f(x0 + dxd, yt, k5);
#pragma unroll
    for (int i = 0; i < NS; ++i)
        yt[i] = y0[i] + dx * (T(-8.0 / 27.0) * d0[i] + T(2.0) * k2[i] +
                              T(-3544.0 / 2565.0) * k3[i] + T(1859.0 / 4104.0) * k4[i] +
                              T(-11.0 / 40.0) * k5[i]);
    f(x0 + (1.0 / 2.0) * dxd, yt, k6);
    constexpr double b1 = 16.0 / 135.0, b3 = 6656.0 / 12825.0, b4 = 28561.0 / 56430.0,
                     b5 = -9.0 / 50.0, b6 = 2.0 / 55.0;
    constexpr double e1 = 25.0 / 216.0 - b1, e3 = 1408.0 / 2565.0 - b3,
                     e4 = 2197.0 / 4104.0 - b4, e5 = -1.0 / 5.0 - b5, e6 = -b6;
    T maxErr = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        y[i] = y0[i] + dx * (T(b1) * d0[i] + T(b3) * k3[i] + T(b4) * k4[i] + T(b5) * k5[i] +
                             T(b6) * k6[i]);
        const T err = dx * (T(e1) * d0[i] + T(e3) * k3[i] + T(e4) * k4[i] + T(e5) * k5[i] +
                            T(e6) * k6[i]);
        const T v0 = ref ? ref[i] + y0[i] : y0[i], v1 = ref ? ref[i] + y[i] : y[i];
        const T tol = at + rt * (fabs(v0) > fabs(v1) ? fabs(v0) : fabs(v1));
        const T q = fabs(err) / tol;
        maxErr = q > maxErr ? q : maxErr;
    }
    return double(maxErr);
}

// Euler::solve(x0, y0, dydx0, dx, y): err = dx dydx0, y = y0 + err
template <int NS, class T>
CG_HD double euler_trial(const T (&y0)[NS], const T (&d0)[NS], double dxd, T (&y)[NS], T at,
                         T rt, const T* ref = nullptr)
{
    const T dx = T(dxd);
    T maxErr = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const T err = dx * d0[i];
        y[i] = y0[i] + err;
        const T v0 = ref ? ref[i] + y0[i] : y0[i], v1 = ref ? ref[i] + y[i] : y[i];
        const T tol = at + rt * (fabs(v0) > fabs(v1) ? fabs(v0) : fabs(v1));
        const T q = fabs(err) / tol;
        maxErr = q > maxErr ? q : maxErr;
    }
    return double(maxErr);
}

// one accepted step from x; dxTry in: trial size, out: next trial size
template <int METHOD, int NS, class T, class F>
CG_HD bool adaptive_step(F& f, double& x, T (&y)[NS], double& dxTry, T at, T rt,
                         const T* ref = nullptr)
{
    T d0[NS], yt[NS];
    f(x, y, d0);
    double dx = dxTry, err;
    for (;;) {
        if constexpr (METHOD == 2) err = euler_trial<NS>(y, d0, dx, yt, at, rt, ref);
        else err = rkf45_trial<NS>(f, x, y, d0, dx, yt, at, rt, ref);
        if (!(err > 1.0)) break;
        const double scale = 0.9 * pow(err, -0.25);
        dx *= scale > 0.2 ? scale : 0.2;
        if (dx < 1e-300) return false;  // step-size underflow
    }
    x += dx;
#pragma unroll
    for (int i = 0; i < NS; ++i) y[i] = yt[i];
    if (err > pow(10.0 / 0.9, -1.0 / 0.2)) {
        const double scale = 0.9 * pow(err, -0.2);
        dxTry = (scale < 10.0 ? scale : 10.0) * dx;
    } else {
        dxTry = 0.9 * 10.0 * dx;
    }
    return true;
}

// integrate y over [xStart, xEnd] (ODESolver::solve); false on step-size underflow or
// max_steps
template <int METHOD, int NS, class T, class F>
CG_HD bool ode_solve(F& f, T (&y)[NS], double xStart, double xEnd, double& dxTry, T at,
                     T rt, long long max_steps, const T* ref = nullptr)
{
    static_assert(METHOD == 1 || METHOD == 2, "adaptive method: 1 RKF45, 2 Euler");
    double x = xStart;
    bool last = false;
    for (long long k = 0; k < max_steps; ++k) {
        const double dxTry0 = dxTry;
        if ((x + dxTry - xEnd) * (x + dxTry - xStart) > 0) {
            last = true;
            dxTry = xEnd - x;
        }
        if (!adaptive_step<METHOD, NS>(f, x, y, dxTry, at, rt, ref)) return false;
        if ((x - xEnd) * (xEnd - xStart) >= 0) {
            if (k > 0 && last) dxTry = dxTry0;
            return true;
        }
    }
    return false;
}

} // namespace cg
