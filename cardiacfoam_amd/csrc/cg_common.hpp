// Shared definitions for the gfx950 kernels of the monodomain tissue step.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

#define CG_DEV __device__ __forceinline__
#define CG_HD __host__ __device__ __forceinline__

namespace cg {

// Heaviside of src/ionicModels/BuenoOrovio/heaviside.H:39-49 (strict x > 0).
template <class T> CG_HD bool pos(T x) { return x > T(0); }

// Precision-generic math: the FP64 path calls the double routines, the FP32 path
// the single ones (hardware v_exp_f32 based), never mixing the two in one formula.
CG_HD double cexp(double x) { return exp(x); }
CG_HD float cexp(float x) { return expf(x); }
CG_HD double clog(double x) { return log(x); }
CG_HD float clog(float x) { return logf(x); }
CG_HD double ctanh(double x) { return tanh(x); }
CG_HD float ctanh(float x) { return tanhf(x); }
CG_HD double csqrt(double x) { return sqrt(x); }
CG_HD float csqrt(float x) { return sqrtf(x); }
CG_HD double cpow(double x, double y) { return pow(x, y); }
CG_HD float cpow(float x, float y) { return powf(x, y); }

// GHK flux quotient v / (exp(k v) - 1) of the generated ICaL/ICab/INab expressions
// (TNNP_2004.H, ORd_2011.H), with ekv = exp(k v) already computed.  FP64 keeps the plain
// quotient, the reference's expression (singular only at v = 0 exactly).  In FP32,
// exp(k v) - 1 cancels for small |k v| and is exactly 0 next to v = 0 (a membrane potential
// crossing 0 mV then gives inf); there the quotient comes from the series
// x / (e^x - 1) = 1 - x/2 + x^2/12 - x^4/720, x = k v, truncation < 1e-12 for |x| < 0.05.
CG_HD double ghk_quot(double v, double, double ekv) { return v / (ekv - 1.0); }
CG_HD float ghk_quot(float v, float k, float ekv)
{
    const float x = k * v;
    if (fabsf(x) < 0.05f) {
        const float x2 = x * x;
        return (1.0f - 0.5f * x + x2 * (1.0f / 12.0f - x2 * (1.0f / 720.0f))) / k;
    }
    return v / (ekv - 1.0f);
}
template <class T> constexpr bool is_fp32() { return sizeof(T) == 4; }

template <class T> CG_HD T sq(T x) { return x * x; }
template <class T> CG_HD T cube(T x) { return x * x * x; }

// Rush-Larsen update of one gate over dt: xinf + (x - xinf) exp(-dt/tau)
// (batchedKernelExecution.H:590-596).  FP32 writes it as the increment
// x + (xinf - x) (1 - exp(-dt/tau)), the form the offset storage of the mixed path adds to
// its stored words (RlStep below); 1 - exp(-e) is exact in FP32 once exp(-e) is rounded.
// That rounding (absolute 6e-8 near 1) still biases the increments of slow gates
// (e = dt/tau ~ 1e-5) by up to 0.6 %; a series for small e removes it (measured on the
// 2.1M-cell ORd run over 200 ms: Vm L-inf against FP64 0.20 -> 0.06 mV) but costs 21 % of
// the ionic kernel's time, and both stay well inside the 0.5 mV bound, so the plain
// exponential is kept in the tissue kernels.
CG_HD float rl_frac(float e) { return 1.0f - cexp(-e); }
// The same with FP32 rounding relative to the result: below e = 0.05 the series
// e (1 - e/2 (1 - e/3 (1 - e/4))), whose truncation is < 6e-8 relative there.  The
// single-cell paths (batches, run_cells) are bound by arithmetic, not memory, and over
// thousands of beats the 0.6 % bias of the plain form moves the action potentials of
// cells paced near refractoriness; they use this form.
CG_HD float rl_frac_precise(float e)
{
    if (e < 0.05f) return e * (1.0f - e * (0.5f - e * (1.0f / 6.0f - e * (1.0f / 24.0f))));
    return 1.0f - cexp(-e);
}
template <class T> CG_HD T rl(T x, T xinf, T tau, T dt)
{
    if constexpr (is_fp32<T>())
        return x + (xinf - x) * rl_frac(dt / tau);
    else
        return xinf + (x - xinf) * cexp(-dt / tau);
}

// Update policies for the models' right-hand sides.  A model writes each state once,
// as out.gate(i, xinf, tau) (dx/dt = (xinf - x)/tau), out.lin(i, a, b) (dx/dt = a - b x)
// or out.rate(i, r) (dx/dt = r), from rates evaluated at the state it was given.
//   RlStep: the Rush-Larsen / forward-Euler step of dt in place (gates exact for frozen
//           xinf, tau; the rest forward Euler).
//   Deriv:  the time derivative, the RATES array an ODE solver asks for
//           (TNNP.C:181-203, ORd.C:157-180, BuenoOrovio.C / AlievPanfilov.C derivatives).
//   RlStep with s: the FP32 tissue step on offset storage -- y (the full values) is only
//           read, and each state's increment is added to its stored offset s[i] instead
//           (see StateRef in models.hpp).
template <class T> struct RlStep {
    T* y;
    T dt;
    T* s = nullptr;
    bool precise = false;   // FP32: rl_frac_precise in place of rl_frac
    CG_HD void gate(int i, T xinf, T tau)
    {
        if constexpr (is_fp32<T>()) {
            const T f = precise ? rl_frac_precise(dt / tau) : rl_frac(dt / tau);
            if (s) s[i] += (xinf - y[i]) * f;
            else y[i] += (xinf - y[i]) * f;
            return;
        }
        y[i] = rl(y[i], xinf, tau, dt);
    }
    CG_HD void lin(int i, T a, T b) { gate(i, a / b, T(1) / b); }
    // the same step for a rate the model writes in its own form r = a - b x
    CG_HD void lin(int i, T a, T b, T) { gate(i, a / b, T(1) / b); }
    CG_HD void rate(int i, T r)
    {
        if constexpr (is_fp32<T>()) {
            if (s) {
                s[i] += dt * r;
                return;
            }
        }
        y[i] += dt * r;
    }
};
template <class T> struct Deriv {
    const T* y;
    T* d;
    CG_HD void gate(int i, T xinf, T tau) { d[i] = (xinf - y[i]) / tau; }
    CG_HD void lin(int i, T a, T b) { d[i] = a - y[i] * b; }
    CG_HD void lin(int i, T, T, T r) { d[i] = r; }
    CG_HD void rate(int i, T r) { d[i] = r; }
};

// Single-cell stimulus protocol of stimulusIO.C:238-286, evaluated in FP64 time.
struct Stimulus {
    double start, period_s1, duration, amplitude;
    int nstim1;
    double period_s2;
    int nstim2;
};

CG_HD double stimulus_current(double t, const Stimulus& s)
{
    double I = 0.0;
    if (s.period_s1 > 0.0 && s.nstim1 > 0) {
        double tp = t - s.start;
        if (tp >= 0 && tp <= s.period_s1 * s.nstim1) {
            double ph = tp - floor(tp / s.period_s1) * s.period_s1;
            if (ph >= 0 && ph <= s.duration) I = -s.amplitude;
        }
    }
    if (I == 0.0 && s.period_s2 > 0.0 && s.nstim2 > 0) {
        double t1 = s.start + s.period_s1 * s.nstim1;
        double tp = t - t1;
        if (tp >= 0 && tp <= s.period_s2 * s.nstim2) {
            double ph = tp - floor(tp / s.period_s2) * s.period_s2;
            if (ph >= 0 && ph <= s.duration) I = -s.amplitude;
        }
    }
    return I;
}

} // namespace cg
