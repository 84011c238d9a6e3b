/*
 * Oracle (test infrastructure): per-cell time integration.
 *
 *  - orc_rl_step: Rush-Larsen for states with a (xinf, tau) form, forward Euler for
 *    the rest, V frozen in tissue mode.  Restates BatchedKernelExecutor::advanceSubsteps
 *    with useExplicitEuler off / useHeun off (batchedKernelExecution.H:563-603,1036-1086).
 *  - orc_rkf45_solve: OpenFOAM (ESI v2312..v2512) ODESolver::solve + adaptiveSolver +
 *    RKF45 as the reference calls it through ionicModel::odeSolver()
 *    (ionicModel.H:235-242, BuenoOrovio.C:132).  Restated from the published
 *    Runge-Kutta-Fehlberg 4(5) method with OpenFOAM's step-size controller
 *    (safeScale 0.9, alphaInc 0.2, alphaDec 0.25, minScale 0.2, maxScale 10) and
 *    error norm max_i |err_i| / (absTol + relTol*max(|y0_i|,|y_i|)).
 *    In tissue mode the models leave dV/dt unwritten; it is taken as 0 here.
 *  - orc_ode_solve with method 2: OpenFOAM's `Euler` ODESolver (`solver Euler;`, the
 *    restitution tutorial's choice) under the same adaptiveSolver controller: one stage,
 *    y = y0 + dx dy/dx(x0), and the whole increment dx dy/dx(x0) as the error estimate.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

void orc_rl_step(int model, int tissue, double t, double dt, const double* c, double* y,
                 int solve_vm, const orc_stim_t* st)
{
    double r[ORC_MAX_STATES], xi[ORC_MAX_STATES], ta[ORC_MAX_STATES], iion, ist;
    int g[ORC_MAX_STATES];
    const int n = orc_model_nstates(model);
    orc_model_eval(model, tissue, t, c, y, solve_vm, st, r, g, xi, ta, &iion, &ist);
    for (int i = 0; i < n; ++i) {
        if (i == 0 && !solve_vm) continue;
        if (g[i]) y[i] = xi[i] + (y[i] - xi[i]) * exp(-dt / ta[i]);
        else y[i] += dt * r[i];
    }
}

/* ------------------------------------------------------------------ RKF45 */
typedef struct {
    int model, tissue, solve_vm, n;
    const double* c;
    const orc_stim_t* st;
} rhs_ctx;

static void rhs(const rhs_ctx* k, double t, const double* y, double* dydt)
{
    double xi[ORC_MAX_STATES], ta[ORC_MAX_STATES], iion, ist;
    int g[ORC_MAX_STATES];
    orc_model_eval(k->model, k->tissue, t, k->c, y, k->solve_vm, k->st, dydt, g, xi, ta,
                   &iion, &ist);
}

/* Fehlberg 4(5) tableau, 5th-order solution propagated */
static const double c2 = 1.0 / 4.0, c3 = 3.0 / 8.0, c4 = 12.0 / 13.0, c5 = 1.0, c6 = 1.0 / 2.0;
static const double a21 = 1.0 / 4.0;
static const double a31 = 3.0 / 32.0, a32 = 9.0 / 32.0;
static const double a41 = 1932.0 / 2197.0, a42 = -7200.0 / 2197.0, a43 = 7296.0 / 2197.0;
static const double a51 = 439.0 / 216.0, a52 = -8.0, a53 = 3680.0 / 513.0, a54 = -845.0 / 4104.0;
static const double a61 = -8.0 / 27.0, a62 = 2.0, a63 = -3544.0 / 2565.0, a64 = 1859.0 / 4104.0,
                    a65 = -11.0 / 40.0;
static const double b1 = 16.0 / 135.0, b3 = 6656.0 / 12825.0, b4 = 28561.0 / 56430.0,
                    b5 = -9.0 / 50.0, b6 = 2.0 / 55.0;
#define e1 (25.0 / 216.0 - b1)
#define e3 (1408.0 / 2565.0 - b3)
#define e4 (2197.0 / 4104.0 - b4)
#define e5 (-1.0 / 5.0 - b5)
#define e6 (-b6)

/* one embedded step from (x0,y0) of size dx; returns the normalised error */
static double rkf45_step(const rhs_ctx* k, double x0, const double* y0, const double* d0,
                         double dx, double* y, double at, double rt)
{
    const int n = k->n;
    double yt[ORC_MAX_STATES] = {0}, k2[ORC_MAX_STATES], k3[ORC_MAX_STATES],
           k4[ORC_MAX_STATES], k5[ORC_MAX_STATES], k6[ORC_MAX_STATES];
    for (int i = 0; i < n; ++i) yt[i] = y0[i] + a21 * dx * d0[i];
    rhs(k, x0 + c2 * dx, yt, k2);
    for (int i = 0; i < n; ++i) yt[i] = y0[i] + dx * (a31 * d0[i] + a32 * k2[i]);
    rhs(k, x0 + c3 * dx, yt, k3);
    for (int i = 0; i < n; ++i) yt[i] = y0[i] + dx * (a41 * d0[i] + a42 * k2[i] + a43 * k3[i]);
    rhs(k, x0 + c4 * dx, yt, k4);
    for (int i = 0; i < n; ++i)
        yt[i] = y0[i] + dx * (a51 * d0[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    rhs(k, x0 + c5 * dx, yt, k5);
    for (int i = 0; i < n; ++i)
        yt[i] = y0[i] + dx * (a61 * d0[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    rhs(k, x0 + c6 * dx, yt, k6);
    double maxErr = 0.0;
    for (int i = 0; i < n; ++i) {
        y[i] = y0[i] + dx * (b1 * d0[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
        double err = dx * (e1 * d0[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i]);
        double tol = at + rt * fmax(fabs(y0[i]), fabs(y[i]));
        double q = fabs(err) / tol;
        if (q > maxErr) maxErr = q;
    }
    return maxErr;
}

/* Euler::solve(x0, y0, dydx0, dx, y): err = dx dydx0, y = y0 + err */
static double euler_step(const rhs_ctx* k, double x0, const double* y0, const double* d0,
                         double dx, double* y, double at, double rt)
{
    (void)x0;
    double maxErr = 0.0;
    for (int i = 0; i < k->n; ++i) {
        double err = dx * d0[i];
        y[i] = y0[i] + err;
        double tol = at + rt * fmax(fabs(y0[i]), fabs(y[i]));
        double q = fabs(err) / tol;
        if (q > maxErr) maxErr = q;
    }
    return maxErr;
}

typedef double (*trial_fn)(const rhs_ctx*, double, const double*, const double*, double,
                           double*, double, double);

/* adaptiveSolver::solve(x, y, dxTry): retry until err <= 1, then grow */
static int adaptive_step(const rhs_ctx* k, trial_fn trial, double* x, double* y,
                         double* dxTry, double at, double rt)
{
    const double safeScale = 0.9, alphaInc = 0.2, alphaDec = 0.25, minScale = 0.2,
                 maxScale = 10.0;
    double d0[ORC_MAX_STATES], yt[ORC_MAX_STATES];
    double dx = *dxTry, err;
    rhs(k, *x, y, d0);
    do {
        err = trial(k, *x, y, d0, dx, yt, at, rt);
        if (err > 1) {
            double scale = fmax(safeScale * pow(err, -alphaDec), minScale);
            dx *= scale;
            if (dx < 1e-300) return -1;
        }
    } while (err > 1);
    *x += dx;
    memcpy(y, yt, sizeof(double) * k->n);
    if (err > pow(maxScale / safeScale, -1.0 / alphaInc)) {
        double scale = fmin(safeScale * pow(err, -alphaInc), maxScale);
        *dxTry = scale * dx;
    } else {
        *dxTry = safeScale * maxScale * dx;
    }
    return 0;
}

long orc_ode_solve(int method, int model, int tissue, const double* c, double* y,
                   int solve_vm, const orc_stim_t* st, double xStart, double xEnd,
                   double* dxTryIO, double at, double rt, long maxSteps)
{
    rhs_ctx k = {model, tissue, solve_vm, orc_model_nstates(model), c, st};
    const trial_fn trial = method == 2 ? euler_step : rkf45_step;
    double x = xStart;
    double dxTry = *dxTryIO;
    int last = 0;
    for (long nStep = 0; nStep < maxSteps; ++nStep) {
        double dxTry0 = dxTry;
        if ((x + dxTry - xEnd) * (x + dxTry - xStart) > 0) {
            last = 1;
            dxTry = xEnd - x;
        }
        if (adaptive_step(&k, trial, &x, y, &dxTry, at, rt) != 0) return -1;
        if ((x - xEnd) * (xEnd - xStart) >= 0) {
            if (nStep > 0 && last) dxTry = dxTry0;
            *dxTryIO = dxTry;
            return nStep + 1;
        }
    }
    return -1;
}

long orc_rkf45_solve(int model, int tissue, const double* c, double* y, int solve_vm,
                     const orc_stim_t* st, double xStart, double xEnd, double* dxTryIO,
                     double at, double rt, long maxSteps)
{
    return orc_ode_solve(1, model, tissue, c, y, solve_vm, st, xStart, xEnd, dxTryIO, at, rt,
                         maxSteps);
}

/* ------------------------------------------------------ tissue solveODE */
static void model_scaling(int model, double* vscale, double* voff, double* tscale,
                          double* iscale)
{
    /* model V = (1000*Vm + voff)/vscale; model time = t_s*tscale; Im = iscale*Iion */
    switch (model) {
    case ORC_BO: *vscale = 85.7; *voff = 84.0; *tscale = 1000.0; *iscale = 85.7; break;
    case ORC_AP: *vscale = 100.0; *voff = 80.0; *tscale = 1000.0 / 12.9; *iscale = 100.0; break;
    case ORC_FDA: *vscale = 1.0; *voff = 0.0; *tscale = 1.0; *iscale = 1.0; break;
    default: *vscale = 1.0; *voff = 0.0; *tscale = 1000.0; *iscale = 1.0; break;
    }
}

void orc_ionic_tissue_step(int model, int tissue, const double* consts, long c_stride,
                           double* states, long n, double t0_s, double dt_s,
                           const double* Vm, double* Im, int integrator, int n_sub,
                           double* step, double at, double rt, int nthreads)
{
    const int ns = orc_model_nstates(model);
    double vscale, voff, tscale, iscale;
    model_scaling(model, &vscale, &voff, &tscale, &iscale);
    /* AlievPanfilov.C:107-108 scales both ends; the others scale start and end; the
     * manufactured model runs in seconds, tEnd = tStart + deltaT
     * (monodomainFDAManufactured.C:107-108) */
    const int fda = model == ORC_FDA;
    const double tStart = fda ? t0_s : model == ORC_AP ? t0_s * 1000.0 / 12.9 : t0_s * 1000.0;
    const double tEnd = fda ? tStart + dt_s
                            : model == ORC_AP ? (t0_s + dt_s) * 1000.0 / 12.9 : (t0_s + dt_s) * 1000.0;
    const double dtModel = fda ? dt_s : model == ORC_AP ? dt_s * 1000.0 / 12.9 : dt_s * 1000.0;
    if (n_sub < 1) n_sub = 1;
    (void)tscale;

#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long cell = 0; cell < n; ++cell) {
        double y[ORC_MAX_STATES];
        const double* c = consts + (c_stride ? cell * c_stride : 0);
        for (int i = 0; i < ns; ++i) y[i] = states[(long)i * n + cell];
        y[0] = fda ? Vm[cell] : vscale == 1.0 ? Vm[cell] * 1000.0 : (Vm[cell] * 1000.0 + voff) / vscale;
        if (integrator == 0) {
            const double h = dtModel / n_sub;
            for (int s = 0; s < n_sub; ++s)
                orc_rl_step(model, tissue, tStart + s * h, h, c, y, 0, NULL);
        } else {
            double st = step[cell];
            if (st > dtModel) st = dtModel;
            orc_ode_solve(integrator, model, tissue, c, y, 0, NULL, tStart, tEnd, &st, at, rt,
                          1000000000L);
            step[cell] = st;
        }
        double r[ORC_MAX_STATES], xi[ORC_MAX_STATES], ta[ORC_MAX_STATES], iion, ist;
        int g[ORC_MAX_STATES];
        orc_model_eval(model, tissue, tEnd, c, y, 0, NULL, r, g, xi, ta, &iion, &ist);
        Im[cell] = fda ? iion / c[0] : iion * iscale;  /* FDA: Iion / Cm */
        for (int i = 0; i < ns; ++i) states[(long)i * n + cell] = y[i];
    }
}

int orc_single_cell_run(int model, int tissue, const double* c, double* y,
                        const orc_stim_t* st, double dt_s, long n_steps, int integrator,
                        double at, double rt, double* t_out, double* vm_out)
{
    /* singleCellSolver.C:179-229 with Time's accumulated t; ionicModel::step_ starts at
     * runTime.deltaT in seconds used as ms (ionicModel.C:38, singleCellSolver.C:48-57) */
    double t = 0.0;
    double step = dt_s;
    const int ap = model == ORC_AP;
    /* the manufactured model runs in seconds with tEnd = tStart + deltaT
     * (monodomainFDAManufactured.C:107-108); its dV/dt is 0, so a single cell moves u1, u2
     * at frozen V (tutorials/manufacturedSolutions/singleCellMonodomain) */
    const int fda = model == ORC_FDA;
    for (long k = 0; k < n_steps; ++k) {
        t = t + dt_s;
        const double t0 = t - dt_s;
        const double tStart = fda ? t0 : ap ? t0 * 1000.0 / 12.9 : t0 * 1000.0;
        const double tEnd = fda ? t0 + dt_s
                                : ap ? (t0 + dt_s) * 1000.0 / 12.9 : (t0 + dt_s) * 1000.0;
        const double dtModel = fda ? dt_s : ap ? dt_s * 1000.0 / 12.9 : dt_s * 1000.0;
        if (integrator == 0) {
            orc_rl_step(model, tissue, tStart, dtModel, c, y, 1, st);
        } else {
            if (step > dtModel) step = dtModel;
            if (orc_ode_solve(integrator, model, tissue, c, y, 1, st, tStart, tEnd, &step, at,
                              rt, 1000000000L) < 0)
                return -1;
        }
        if (t_out) t_out[k] = t;
        if (vm_out) vm_out[k] = orc_model_vm_mV(model, y);
    }
    return 0;
}

void orc_single_cell_trace(int model, int tissue, const double* c, double* y,
                           const orc_stim_t* st, double dt_model, long n_steps,
                           double* t_out, double* vm_out)
{
    for (long k = 0; k < n_steps; ++k) {
        orc_rl_step(model, tissue, (double)k * dt_model, dt_model, c, y, 1, st);
        if (t_out) t_out[k] = (double)(k + 1) * dt_model;
        if (vm_out) vm_out[k] = orc_model_vm_mV(model, y);
    }
}
