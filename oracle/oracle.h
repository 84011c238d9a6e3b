/*
 * CPU FP64 oracle for the monodomain tissue step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker that the HIP path is compared against.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product library never links or calls it.
 *
 * It restates, in plain C with OpenMP, the reference's per-cell ionic loop and
 * face-loop Laplacian:
 *   - ionic right-hand sides:   src/ionicModels/BuenoOrovio/BuenoOrovio_2008.H:116-225,
 *                               src/ionicModels/TNNP/TNNP_2004.H:309-603,
 *                               src/ionicModels/ORd/ORd_2011.H:493-989,
 *                               src/ionicModels/AlievPanfilov/AlievPanfilov_1996.H:72-120
 *   - unit scalings of solveODE: BuenoOrovio.C:107-152, TNNP.C:122-174, ORd.C:110-150,
 *                               AlievPanfilov.C:107-152
 *   - Rush-Larsen semantics:    src/ionicModels/ionicModel/batchedKernelExecution.H:563-653,1015-1103
 *   - adaptive RKF45 (OpenFOAM ODESolver/adaptiveSolver/RKF45, ESI v2312..v2512)
 *   - stimulus:                 src/genericWriter/stimulusIO.C:238-286 (single cell),
 *                               myocardiumDomain.C:307-342 (tissue boxes)
 *   - activation time:          myocardiumDomain.C:345-374
 *   - explicit dt:              myocardiumDomain.C:532-550
 *   - diffusion:                monodomainSolver.C:103-150 (fvc/fvm::laplacian, Gauss linear
 *                               corrected, tensor conductivity)
 *   - APD90:                    postProcessing_restCurves.py:115-200,270-295
 *
 * Parity status of each part is recorded next to its function below.
 */
#ifndef CARDIAC_ORACLE_H
#define CARDIAC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model ids (same numbering as the product C-ABI) */
enum { ORC_BO = 0, ORC_TNNP = 1, ORC_ORD = 2, ORC_AP = 3, ORC_FDA = 4, ORC_CRN = 5,
       ORC_PYS = 6, ORC_STW = 7, ORC_GRD = 8, ORC_TOR = 9, ORC_GAUR = 10 };
/* dimension flags of the manufactured model (its `dimension` word 1D/2D/3D) */
enum { ORC_DIM1 = 11, ORC_DIM2 = 12, ORC_DIM3 = 13 };
/* tissue flags, ionicSelector.C:60-66 */
enum { ORC_EPI = 1, ORC_MCELL = 2, ORC_ENDO = 3, ORC_MYOCYTE = 4 };

#define ORC_MAX_STATES 48
#define ORC_MAX_CONSTS 160

typedef struct {
    double start, period_s1, duration, amplitude;
    int nstim1;
    double period_s2;
    int nstim2;
} orc_stim_t;

/* ---------------------------------------------------------------- models */
int orc_model_nstates(int model);
/* fills constants and initial states; returns 0, or -1 if the tissue is not
 * supported by that model (BuenoOrovio/TNNP: epi/M/endo, ORd/AP: myocyte) */
int orc_model_init(int model, int tissue, double* consts, double* states);
/* index of a named conductance inside consts (g_Na, g_Kr, g_CaL, g_Ks for TNNP;
 * GNa, GKr, PCa, GKs for ORd); -1 if absent */
int orc_model_const_index(int model, const char* name);

/* One evaluation of the right-hand side at model time t (ms, or AP time units).
 * rates[i]   : dy_i/dt (rates[0]=0 unless solve_vm)
 * gate[i]    : 1 when state i is integrated by Rush-Larsen with (xinf[i], tau[i])
 * iion       : total ionic current in model units
 * istim      : single-cell stimulus current (stimulusIO.C:238-286)            */
void orc_model_eval(int model, int tissue, double t, const double* c, const double* y,
                    int solve_vm, const orc_stim_t* stim,
                    double* rates, int* gate, double* xinf, double* tau,
                    double* iion, double* istim);

/* stimulus protocol, stimulusIO.C:238-286 */
double orc_stimulus(double t, const orc_stim_t* s);

/* ALGEBRAIC variables in the reference's order (BuenoOrovio: 12, AlievPanfilov: 3) at
 * model time t; returns their count, or -1 for a model without a restated list. */
int orc_model_algebraic(int model, double t, const double* c, const double* y,
                        const orc_stim_t* stim, double* alg);

/* ----------------------------------------------------------- integrators */
/* One Rush-Larsen / forward-Euler step of dt (model units) on one cell. */
void orc_rl_step(int model, int tissue, double t, double dt, const double* c, double* y,
                 int solve_vm, const orc_stim_t* stim);

/* OpenFOAM ODESolver::solve(xStart, xEnd, y, dxTry) with adaptiveSolver and the stage
 * scheme `method`: 1 RKF45, 2 Euler.  Returns the number of accepted steps, or -1 on
 * maxSteps exceeded or step-size underflow. */
long orc_ode_solve(int method, int model, int tissue, const double* c, double* y,
                   int solve_vm, const orc_stim_t* stim, double x0, double x1, double* dx_try,
                   double abs_tol, double rel_tol, long max_steps);
long orc_rkf45_solve(int model, int tissue, const double* c, double* y, int solve_vm,
                     const orc_stim_t* stim, double x0, double x1, double* dx_try,
                     double abs_tol, double rel_tol, long max_steps);

/* Tissue ionic step of the reference's solveODE (ionicModel.H:250-256):
 * states SoA [nstates][n]; Vm in V; Im out in V/s.  integrator 0 = Rush-Larsen
 * (n_sub substeps), 1 = RKF45, 2 = OpenFOAM's adaptive Euler, with per-cell step[] (model
 * time units, persistent).
 * c_stride = 0: one constant set for all cells; else consts[cell*c_stride+k]. */
void orc_ionic_tissue_step(int model, int tissue, const double* consts, long c_stride,
                           double* states, long n, double t0_s, double dt_s,
                           const double* Vm, double* Im, int integrator, int n_sub,
                           double* step, double abs_tol, double rel_tol, int nthreads);

/* Reference single-cell driver (singleCellSolver.C:179-229) with Vm solved inside
 * the ODE.  Writes the trace (time s, Vm mV) for n_steps PDE steps of dt_s.
 * integrator as above; returns 0 on success. */
int orc_single_cell_run(int model, int tissue, const double* c, double* y,
                        const orc_stim_t* stim, double dt_s, long n_steps,
                        int integrator, double abs_tol, double rel_tol,
                        double* t_out, double* vm_out);

/* Batched-protocol single cell: n_steps Rush-Larsen steps of dt_model (model time units)
 * with the step start at k*dt_model (time as step count times dt, no accumulation) and
 * Vm inside the ODE; writes the trace t_out[k] = (k+1)*dt_model, vm_out[k] (mV) after
 * each step.  Parity: same RL step and stimulus as orc_single_cell_run. */
void orc_single_cell_trace(int model, int tissue, const double* c, double* y,
                           const orc_stim_t* stim, double dt_model, long n_steps,
                           double* t_out, double* vm_out);

/* Vm of the single cell in mV from a model state vector (ioVmTransform) */
double orc_model_vm_mV(int model, const double* y);

/* ------------------------------------------------------------ hex mesh */
typedef struct {
    int nx, ny, nz;
    double hx, hy, hz;
    long n_cells;
    long n_faces;        /* internal faces */
    int32_t* owner;      /* [n_faces] */
    int32_t* neighbour;  /* [n_faces] */
    int8_t* face_dir;    /* 0,1,2 = x,y,z normal */
} orc_mesh_t;

/* Pseudo-ECG of V at ne electrodes pos[3 ne] (pseudoECGSolver.C:36-83), sigma [9n]
 * per cell: out[ne], absout[ne] (sum of |terms|, may be NULL). */
void orc_pseudo_ecg(const orc_mesh_t* m, const double* sigma, const double* V, int ne,
                    const double* pos, double* out, double* absout, int nthreads);

/* Builds the OpenFOAM blockMesh ordering of a uniform nx*ny*nz hex block:
 * cell id = i + nx*(j + ny*k); internal faces in owner order, each owner's
 * faces sorted by neighbour (upper-triangular order). */
int orc_mesh_build(orc_mesh_t* m, int nx, int ny, int nz, double hx, double hy, double hz);
void orc_mesh_free(orc_mesh_t* m);

/* Explicit FV Laplacian div(sigma grad V) per unit volume (V/s * F/m^3 = A/m^3),
 * Gauss linear corrected with tensor sigma [9n] row-major per cell:
 *   flux_f = SfGammaSn * snGrad(V) + SfGammaCorr . interp(grad V)
 * with zeroGradient walls; grad is Gauss linear (grad_scheme 0) or least squares
 * (1, reduces to the same on a uniform hex).  out[n]. */
void orc_laplacian(const orc_mesh_t* m, const double* sigma, const double* V,
                   double* out, int nthreads);

/* The implicit operator: orthogonal part as an LDU matrix (diag[n], upper[F])
 * of  chiCm/dt * coefft * V - div(SfGammaSn snGrad V), and the explicit correction
 * source corr[n] = div(SfGammaCorr . interp(grad V)) per unit volume. */
void orc_assemble_implicit(const orc_mesh_t* m, const double* sigma, double chiCm,
                           double dt, double coefft, double* diag, double* upper);
void orc_correction_div(const orc_mesh_t* m, const double* sigma, const double* V,
                        double* out, int nthreads);

/* y = A x for an LDU matrix with lower == upper */
void orc_ldu_amul(const orc_mesh_t* m, const double* diag, const double* upper,
                  const double* x, double* y);
/* Jacobi-preconditioned CG on the LDU matrix until ||b - Ax||_2 <= tol*||b||_2.
 * Returns iterations, final relative residual in *res. */
int orc_pcg(const orc_mesh_t* m, const double* diag, const double* upper,
            const double* b, double* x, double tol, int max_iter, double* res);

/* ------------------------------------------------- tissue orchestration */
typedef struct {
    int n_boxes;
    const double* box_min;   /* [3*n_boxes] m */
    const double* box_max;
    const double* t_start;   /* s */
    const double* duration;  /* s */
    const double* intensity; /* A/m^3 */
    const double* centres;   /* [3*n_cells] cell centres (m) as the mesh's C(), or null for
                                the block's (i+1/2)h lattice */
} orc_boxes_t;

/* source[n] at t0 (myocardiumDomain.C:307-342): the boxes whose closed extent contains the
 * cell centre (boundBox::contains is inclusive) */
void orc_box_stimulus(const orc_mesh_t* m, const orc_boxes_t* b, double t0, double* src);
/* activation (myocardiumDomain.C:345-374); flag 1 = still to be computed */
void orc_activation(long n, const double* V, const double* Vold, uint8_t* flag,
                    double* tact, double t_old, double dt);
/* explicit dt = maxCo * min_f dx^2 / (n.sigma_f.n / chiCm), myocardiumDomain.C:532-550 */
double orc_explicit_dt(const orc_mesh_t* m, const double* sigma, double chiCm, double maxCo);

/* OpenFOAM PCG on the LDU system (diag, upper in face order); returns iterations,
 * initial and final normalised residuals */
int orc_pcg_of(const orc_mesh_t* m, const double* diag, const double* upper, const double* b,
               double* x, double tol, double rel_tol, int max_iter, int precond,
               double* init_res, double* final_res);

typedef struct {
    int model, tissue;
    int integrator;          /* 0 = Rush-Larsen (n_sub substeps), 1 = RKF45, 2 = Euler */
    int n_sub;
    double abs_tol, rel_tol; /* RKF45 */
    int ddt;                 /* 0 = Euler, 1 = backward (BDF2, Euler on step 1) */
    int implicit;            /* 0 = explicit fvc::laplacian, 1 = fvm::laplacian + PCG */
    double pcg_tol; int pcg_max_iter;
    double chi, cm;
    double dt;               /* s */
    int nthreads;
    /* implicit solver controls (fvSolution): norm 0 = ||r||2 <= pcg_tol ||b||2 (Jacobi
     * PCG), 1 = OpenFOAM's normalised residual with pcg_rel_tol and precond (0 diagonal,
     * 1 DIC); PIMPLE n_outer (nOuterCorrectors, <= 1: one solve) and residual_control
     * (initial-residual tolerance of the outer loop) */
    int norm;
    double pcg_rel_tol;
    int precond;
    int n_outer;
    double residual_control;
} orc_tissue_cfg_t;

typedef struct {
    orc_mesh_t mesh;
    orc_tissue_cfg_t cfg;
    int nstates;
    double consts[ORC_MAX_CONSTS];
    double* states;          /* SoA [nstates][n] */
    double* step;            /* RKF45 per-cell step (ms) */
    double* sigma;           /* [9n] */
    double* V;  double* V0;  double* V00;
    double* Iion; double* src; double* tact; uint8_t* flag;
    double* work;            /* scratch [4n] */
    double t;                /* accumulated time (s) */
    long step_index;
    long pcg_iters_total;
    long outer_total;        /* linear solves (outer correctors) over all steps */
    double last_init_res;    /* initial residual of the last solve */
} orc_tissue_t;

int orc_tissue_create(orc_tissue_t* T, int nx, int ny, int nz, double hx, double hy,
                      double hz, const orc_tissue_cfg_t* cfg, const double* sigma9);
void orc_tissue_free(orc_tissue_t* T);
/* one step: t += dt; t0 = t - dt; stimulus; ionic; diffusion; activation */
int orc_tissue_step(orc_tissue_t* T, const orc_boxes_t* boxes);

/* Tangent-linear run (tangent.cpp) of the BuenoOrovio tissue from rest: explicit, Euler
 * ddt, Rush-Larsen, sigma_c = kappa_c sig0[3] per cell, h-cube cells.  Loss = sum over
 * n_steps and probes of (Vm - target[step][probe])^2; dloss[k] its derivative along
 * direction k = (dkappa[k][n], dtheta[k][2]) of (kappa, tau_fi, tau_si), n_dir <= 8.
 * v_final [n] (may be NULL): Vm after the last step.  Returns 0, or < 0 on bad input. */
int orc_bo_tangent(int nx, int ny, int nz, double h, const double* sig0, int tissue,
                   double dt, double chi, double cm, const orc_boxes_t* boxes,
                   const double* kappa, const double* theta, int n_dir, const double* dkappa,
                   const double* dtheta, long n_steps, int n_probes, const long* probes,
                   const double* target, int nthreads, double* loss, double* dloss,
                   double* v_final);

/* ------------------------------------------------------------ APD90 */
typedef struct {
    double t_up, v_base, v_peak, v90, t_repol90;
    int repol_found;
} orc_beat_t;
/* detect_beats of postProcessing_restCurves.py:115-200 on a (time s, Vm mV) trace;
 * returns the number of beats written (<= max_beats). */
int orc_detect_beats(const double* t, const double* vm, long n, orc_beat_t* beats,
                     int max_beats, double dvdt_threshold, double baseline_s,
                     double peak_s);

#ifdef __cplusplus
}
#endif
#endif
