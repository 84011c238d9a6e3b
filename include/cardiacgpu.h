/*
 * cardiacgpu.h -- C-ABI of the MI355X (gfx950) monodomain tissue step.
 *
 * Drop-in boundary for cardiacFoam's MonoDomainSolver hot path.  An OpenFOAM-side
 * adapter (an ionicModel subclass and a myocardiumSolver subclass registered under the
 * existing runtime-selection keys) forwards to these functions; Python reaches the same
 * symbols through ctypes.  Only plain pointers, sizes and PODs cross the boundary.
 *
 * Units at the boundary follow the reference (src/electroModels/core/dimVoltage.H,
 * myocardiumDomain.C:196-251): time s, Vm V, Im/Iion V/s, sigma S/m, chi 1/m,
 * Cm F/m^2, stimulus A/m^3.  Model-internal units (ms, mV) never cross it.
 *
 * Every entry returns CG_OK (0) or an error code; cg_last_error() gives a thread-local
 * message the adapter turns into FatalErrorInFunction.  Each handle owns its device
 * memory and one HIP stream; calls on one handle must not overlap.  Handles may be
 * destroyed in any order (a destroy touches its own handle only, after a device-wide
 * synchronize); every other call needs the handles it was created from still alive.
 */
#ifndef CARDIACGPU_H
#define CARDIACGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
enum {
    CG_OK = 0,
    CG_ERR_ARG = 1,        /* invalid argument / size */
    CG_ERR_MODEL = 2,      /* unknown ionicModel (ionicModel.C:83-88) */
    CG_ERR_TISSUE = 3,     /* unsupported tissue (ionicSelector.C:32-52) */
    CG_ERR_HIP = 4,        /* HIP runtime failure */
    CG_ERR_NOMEM = 5,      /* device allocation failed */
    CG_ERR_NONFINITE = 6,  /* NaN/Inf detected in a checked output */
    CG_ERR_NOT_CONVERGED = 7, /* linear solver hit max iterations */
    CG_ERR_STIMULUS = 8,   /* malformed stimulus (stimulusIO.C:123-177) */
    CG_ERR_STATE = 9,      /* call not valid in the handle's configuration */
    CG_ERR_ODE = 10        /* ODE solver failure: step-size underflow or maxSteps exceeded */
};

/* model ids: the ionicModel runtime-selection keys */
enum { CG_MODEL_BUENO_OROVIO = 0, CG_MODEL_TNNP = 1, CG_MODEL_ORD = 2, CG_MODEL_ALIEV_PANFILOV = 3,
       CG_MODEL_FDA_MANUFACTURED = 4, /* monodomainFDAManufactured (verification model) */
       CG_MODEL_COURTEMANCHE = 5, CG_MODEL_PERIS_YAGUE_2022 = 6, CG_MODEL_STEWART = 7 };
/* tissue flags of ionicSelector::tissueFlag (ionicSelector.C:60-66); the manufactured
 * model takes its `dimension` word instead ("1D", "2D", "3D") */
enum { CG_TISSUE_EPI = 1, CG_TISSUE_MCELL = 2, CG_TISSUE_ENDO = 3, CG_TISSUE_MYOCYTE = 4,
       CG_DIMENSION_1D = 11, CG_DIMENSION_2D = 12, CG_DIMENSION_3D = 13 };
/* arithmetic: CG_FP64 = plain FP64 path; CG_MIXED = FP32 states/arithmetic, FP64 time,
 * reductions, residuals and activation times */
enum { CG_FP64 = 0, CG_MIXED = 1 };
/* flags */
enum { CG_DEVICE_PTRS = 1 };
/* per-cell ODE integrators of the ionic step */
enum { CG_ODE_RUSH_LARSEN = 0, CG_ODE_RKF45 = 1, CG_ODE_EULER = 2 };

const char* cg_last_error(void);
/* library version string */
const char* cg_version(void);

/* ------------------------------------------------------------- ionic model */
typedef struct cg_ionic cg_ionic;

/* Model lookup by ionicModel key ("BuenoOrovio", "TNNP", "ORd", "AlievPanfilov",
 * "monodomainFDAManufactured", "Courtemanche", "PerisYague_2022", "Stewart") and tissue word ("epicardialCells", "mCells",
 * "endocardialCells", "myocyte"; "1D", "2D", "3D" for the manufactured model).
 * Replaces ionicModel::New's table lookup (ionicModel.C:76-94) and
 * ionicSelector::tissueFlag (ionicSelector.C:60-66).  Return -1 when unknown. */
int cg_model_id(const char* ionic_model_key);
int cg_tissue_flag(const char* tissue_word);
/* number of states of a model (NUM_STATES of the generated headers), or -1 */
int cg_model_nstates(int model);
/* name of state i (the *_STATES_NAMES tables, e.g. BuenoOrovio_2008.H:24-30) */
int cg_model_state_name(int model, int i, char* buf, size_t len);
/* The model's units (host only; any pointer may be NULL).  solveODE feeds a tissue cell
 * the V state (Vm * vmul + voff) / vscale from Vm in V (BuenoOrovio.C:119-122: vmul 1000,
 * voff 84, vscale 85.7; AlievPanfilov 1000, 80, 100; the mV models 1000, 0, 1; the
 * manufactured model 1, 0, 1); ioVmTransform gives the trace's Vm in mV as
 * V * vscale - voff (BuenoOrovio.H:78-81, AlievPanfilov.H:78-81); model time is PDE time
 * in s times t_per_s (ms; AlievPanfilov 1000/12.9; s for the manufactured model,
 * BuenoOrovio.C:107-108, AlievPanfilov.C:107-108, monodomainFDAManufactured.C:98). */
int cg_model_units(int model, double* vmul, double* voff, double* vscale, double* t_per_s);

/* Replaces <Model>::<Model>(dict, nIntegrationPoints, initialDeltaT, solveVm)
 * (ionicModel.H:184-196; BuenoOrovio.C:45-81, TNNP.C:44-78, ORd.C:45-80,
 * AlievPanfilov.C:45-80, Courtemanche.C:44-80, PerisYague.C:44-80,
 * Stewart.C:44-81): allocates SoA states for n_cells integration points
 * initialised from the model's initConsts.  solve_vm = 1 is the single-cell mode
 * (singleCellSolver.C:48-57). */
int cg_ionic_create(int model, int tissue, int64_t n_cells, int solve_vm, int precision,
                    cg_ionic** out);
int cg_ionic_destroy(cg_ionic* h);
int cg_ionic_nstates(const cg_ionic* h);

/* Override model constant k for all cells.  k indexes the handle's parameter vector, in
 * the order of the parameter enum of the model's device header
 * (cardiacfoam_amd/csrc/model_*.hpp; e.g. Courtemanche: gNa gK1 gto gKur_base gKr gKs
 * gCaL gbCa gbNa INaK_max INaCa_max IpCa_max I_up_max K_rel sigma).  Replaces editing
 * CONSTANTS_ after initConsts. */
int cg_ionic_set_constant(cg_ionic* h, int k, double value);
/* Per-cell multipliers of four conductances, SoA [4][n_cells] (NULL clears):
 * TNNP g_Na g_Kr g_CaL g_Ks; ORd GNa GKr PCa GKs; Courtemanche, PerisYague_2022 and
 * Stewart gNa gKr gCaL gKs.
 * No reference counterpart
 * (constants are per model instance there, TNNP.H:61). */
int cg_ionic_set_conductance_scale(cg_ionic* h, const double* scale4xn);

/* Single-cell stimulus protocol (stimulusIO.C:142-194 keys, ms and model units).
 * Rejected for tissue handles, as myocardiumDomain.C:377-394 does. */
typedef struct {
    double stim_start, stim_period_s1, stim_duration, stim_amplitude;
    int nstim1;
    double stim_period_s2;
    int nstim2;
} cg_stimulus;
int cg_ionic_set_stimulus(cg_ionic* h, const cg_stimulus* s);

/* Selects the per-cell ODE integrator; replaces the `solver`, `absTol`, `relTol`,
 * `maxSteps` keys that ionicModel::odeSolver() hands to ODESolver::New
 * (ionicModel.H:233-248) and the per-cell step list step_ (ionicModel.H:244-248,
 * ionicModel.C:38).
 *   CG_ODE_RUSH_LARSEN: gates exponentially with frozen (xinf, tau), the rest forward
 *     Euler, one step per deltaT (batchedKernelExecution.H:563-653); tolerances unused.
 *   CG_ODE_RKF45: OpenFOAM's adaptive RKF45 over [tStart, tEnd] per cell with the error
 *     norm max |err| / (abs_tol + rel_tol max(|y0|, |y|)); each cell keeps its last trial
 *     step between calls and starts from min(step, deltaT) (BuenoOrovio.C:124-132).
 *     initial_step (model time units) seeds the kept steps; <= 0 seeds them with the
 *     first call's deltaT in seconds, as the reference's initialDeltaT does.  OpenFOAM's
 *     defaults when the keys are absent: abs_tol 1e-15, rel_tol 1e-4, max_steps 10000.
 *   CG_ODE_EULER: OpenFOAM's `Euler` ODESolver under the same controller and kept steps:
 *     y = y0 + dx f(x0, y0) with the increment itself as the error estimate.
 * A failing cell (step-size underflow, more than max_steps sub-steps) makes the next
 * synchronising call return CG_ERR_ODE.  Batched protocols (cg_batch_run) run all three
 * (cg_batch_autotune only Rush-Larsen), each adaptive protocol step [s dt, (s+1) dt] from
 * the cell's kept trial step; RKF45 on FP64 Gaur handles gets CG_ERR_STATE there.  Euler
 * takes sub-steps of at most abs_tol / |dV/dt| while V crosses 0 mV (the relative part of
 * its tolerance vanishes there): millions of them per crossing at abs_tol 1e-6, so
 * max_steps must allow for that.  CG_ERR_STATE for CG_ODE_RKF45 on an FP64 Gaur handle in tissue mode
 * (solve_vm 0): that kernel is switched off (its gfx950 build gives non-finite states
 * where the same source on the host does not); mixed precision and single cells run. */
int cg_ionic_set_integrator(cg_ionic* h, int kind, double abs_tol, double rel_tol,
                            int64_t max_steps, double initial_step);

/* Replaces ionicModel::solveODE(stepStartTime, deltaT, Vm, Im) (ionicModel.H:250-256)
 * for a tissue handle: Vm [n] in V, Im [n] out in V/s.  Pointers are host memory unless
 * flags has CG_DEVICE_PTRS.  Synchronous. */
int cg_ionic_step(cg_ionic* h, double t0_s, double dt_s, const double* Vm, double* Im,
                  int flags);

/* Single-cell handles: n_steps steps of dt_s from t0_s with Vm inside the ODE
 * (SingleCellSolver::evolve, singleCellSolver.C:179-229), all cells in one launch with
 * the states kept in registers; RKF45 evaluates the stimulus at every stage time.
 * Asynchronous; cg_ionic_get_states synchronises. */
int cg_ionic_run_cells(cg_ionic* h, double t0_s, double dt_s, int n_steps);
/* cg_ionic_run_cells that also records one cell after every step, the rows
 * SingleCellSolver::evolve writes to its trace file (singleCellSolver.C:218-221,
 * ionicModelIO.C:244-318): trace [n_steps][2][n_states] host doubles, the states and then
 * their rates (RATES, with the stimulus at the step's end time), in model units and state
 * order.  Synchronous; reports an RKF45 failure like cg_ionic_get_states. */
int cg_ionic_run_cells_trace(cg_ionic* h, double t0_s, double dt_s, int n_steps,
                             int64_t cell, double* trace);

/* ---------------------------------------------------- batched single cells
 * Thousands to millions of independent single cells (solve_vm = 1 handles), each with its
 * own stimulus protocol, run length and conductance scales, in one hipGraph launch.  No
 * reference counterpart as one call: the reference runs one cell per SingleCellSolver
 * (singleCellSolver.C:179-229) and post-processes the traces offline
 * (postProcessing_restCurves.py:115-200). */

/* Per-cell protocols, n_cells entries with the fields of cg_stimulus (model time units,
 * ms for BuenoOrovio/TNNP/ORd); NULL returns every cell to the handle's protocol. */
int cg_ionic_set_protocols(cg_ionic* h, const cg_stimulus* per_cell);
/* Run lengths in steps per cell (NULL: every cell runs n_steps_all). */
int cg_ionic_set_run_lengths(cg_ionic* h, const int64_t* steps, int64_t n_steps_all);
/* On-device beat detection during cg_batch_run, the detect_beats definition of
 * postProcessing_restCurves.py:115-200 on each cell's trace: upstroke where the centred
 * dVm/dt crosses dvdt_threshold (V/s; 5 V/s there), baseline = min Vm over baseline_s
 * before it, peak over peak_window_s (or up to the next upstroke), repolarisation at
 * 90 %.  Up to max_beats records per cell; 0 switches detection off. */
int cg_ionic_enable_beats(cg_ionic* h, int max_beats, double dvdt_threshold, double baseline_s,
                          double peak_window_s);
/* Runs every cell of every handle from time 0 for its run length with steps of dt_s
 * (states continue from their current values), chunk steps per launch, all launches of
 * all handles captured in one hipGraph (handles on their own streams run concurrently).
 * Synchronous. */
int cg_batch_run(cg_ionic* const* handles, int n_handles, double dt_s, int chunk);
/* Times each kernel shape (threads per block x register budget) on trial_steps steps of
 * every handle, restores the states, keeps the fastest shape per handle; a text report
 * goes to report (may be NULL). */
int cg_batch_autotune(cg_ionic* const* handles, int n_handles, double dt_s, int trial_steps,
                      char* report, size_t report_len);
/* Beats of the last cg_batch_run: counts[n_cells] (may exceed max_beats) and records
 * [n_cells][max_beats][5] = t_up (s), Vm base, peak, V90 (V), t_repol90 (s, NaN when the
 * trace ends first); unused records are NaN. */
int cg_ionic_get_beats(cg_ionic* h, int32_t* counts, double* records);
/* Early-stroke classification and APD90/DI of every cell's beats from the last batch run:
 * classify_early_strokes and compute_apd_di of the restitution post-processing
 * (tutorials/singleCellprotocols/restitutionCurves_s1s2Protocol/
 * setupRestitutionCurves_s1s2Protocol/postProcessing_restCurves.py:203-297), on the device.
 * A beat is early when the next upstroke comes before its 90 % repolarisation or it never
 * repolarised; the last two beats are the S1-S2 pair.  apd_di [n][2] (host): APD90 of S2
 * and DI = S2 upstroke - S1 90 % repolarisation in s, both NaN where the reference rejects
 * the pair (returns None), and where a cell detected more beats than max_beats.
 * flags [n][max_beats] (host, may be NULL): per beat 1 early, 2 valid APD, 4 valid DI, 8
 * the cell's beats overflowed max_beats.  Synchronous. */
int cg_ionic_get_restitution(cg_ionic* h, double* apd_di, uint8_t* flags);

/* States as host SoA doubles [n_vars][n_cells], starting at state first_var: the full
 * values (CG_MIXED handles keep their slow concentrations on the device as FP32 offsets
 * from the model's initial values; both calls convert).
 * Replaces the io*Ptr hooks / exportStates (ionicModel.H:115-168, 584-638). */
/* Derived variables of the handle's model (the reference's RATES_ and ALGEBRAIC_ lists,
 * ionicModel.H:115-168, as computeVariables leaves them at a step's end) evaluated on the
 * device in FP64 at n rows given by the caller: states [nstates][n] (host, model units, V
 * as state 0, e.g. the cells' states with their last Vm), t_model [n] the model time of
 * each row (single-cell handles: the stimulus of the handle's protocol at that time; may
 * be NULL for tissue handles, whose ODEs see no stimulus).  kind CG_RATES: dy/dt [nstates][n]
 * (dV/dt = 0 in tissue mode, as the reference's solver leaves it); CG_ALGEBRAIC: the
 * model's algebraic list [cg_model_nalgebraic][n] in the reference's order, evaluated for
 * BuenoOrovio (12) and AlievPanfilov (3); other models return CG_ERR_MODEL.  With n equal
 * to the handle's cells the per-cell conductance scales apply.  Synchronous. */
enum { CG_RATES = 0, CG_ALGEBRAIC = 1 };
int cg_ionic_eval(cg_ionic* h, int kind, int64_t n, const double* states,
                  const double* t_model, double* out);
/* Length of the model's algebraic list evaluated on the device (0: none; -1: unknown
 * model) and its names in the reference's order (BuenoOrovio_2008.H:30-43,
 * AlievPanfilov_1996.H:28-32). */
int cg_model_nalgebraic(int model);
int cg_model_algebraic_name(int model, int i, char* buf, size_t len);

int cg_ionic_get_states(cg_ionic* h, int first_var, int n_vars, double* host_soa);
int cg_ionic_set_states(cg_ionic* h, int first_var, int n_vars, const double* host_soa);

/* -------------------------------------------------------------- hex mesh */
typedef struct cg_mesh cg_mesh;

/* Uniform nx*ny*nz hex block in blockMesh cell order (c = i + nx*(j + ny*k)), spacing
 * h (m).  Replaces the fvMesh the myocardiumSolver receives (myocardiumSolver.H:34-106)
 * for single-block hex meshes. */
int cg_mesh_create_hex(int nx, int ny, int nz, double hx, double hy, double hz,
                       cg_mesh** out);
/* Any polyhedral mesh, as the fvMesh hands it over (the cg_mesh_create of the survey's
 * boundary; myocardiumSolver.H:34-106, monodomainSolver.C:103-150): n_faces internal faces
 * in lduAddressing order (owner < neighbour, mesh.owner()/mesh.neighbour()), their area
 * vectors Sf [n_faces][3] (pointing from owner to neighbour) and centres Cf [n_faces][3],
 * cell centres C [n_cells][3] and volumes V [n_cells], and the boundary faces (all
 * zeroGradient for Vm, as in every tissue case) by owner cell and outward Sf.  Host arrays,
 * copied.  Tissues on this mesh run fvc::laplacian / fvm::laplacian with a tensor
 * conductivity and Gauss linear corrected (non-orthogonal correction included) and solve
 * the implicit system with the lduMatrix PCG of cg_ldu_solve in OpenFOAM's residual norm;
 * cg_tissue_apply_operator and the adjoint cover the hex block only. */
int cg_mesh_create_general(int64_t n_cells, int64_t n_faces, const int32_t* owner,
                           const int32_t* neighbour, const double* Sf, const double* Cf,
                           const double* C, const double* V, int64_t n_boundary_faces,
                           const int32_t* boundary_owner, const double* boundary_Sf,
                           cg_mesh** out);
/* Checks that an lduAddressing (owner/neighbour, upper-triangular order) is the
 * blockMesh ordering of an nx*ny*nz block; 0 if it is. */
int cg_mesh_match_ldu(int nx, int ny, int nz, int64_t n_faces, const int32_t* owner,
                      const int32_t* neighbour);
int cg_mesh_destroy(cg_mesh* m);
int64_t cg_mesh_ncells(const cg_mesh* m);

/* Conductivity, row-major 3x3 per cell [9*n] (uniform == 1: one tensor for all cells).
 * Replaces MonodomainSolver::initialiseConductivity (monodomainSolver.C:48-100). */
int cg_mesh_set_conductivity(cg_mesh* m, const double* sigma9, int uniform);
/* Cell centres of a hex block as the fvMesh holds them (mesh().C().primitiveField(), m),
 * [3*n] interleaved x y z.  The box stimulus (myocardiumDomain.C:307-342) then tests these
 * centres with the reference's inclusive boundBox::contains instead of the (i+1/2)h
 * lattice; a box edge on a cell-centre plane selects cells by the centres' last bit.  The
 * membership of each cell is computed once per change of boxes or centres (1 byte per
 * cell on the device, read only while a box is active).  Hex blocks only. */
int cg_mesh_set_cell_centres(cg_mesh* m, const double* centres);
/* maxCo * min_f dx^2/(n.sigma_f.n/(chi Cm)) (myocardiumDomain.C:532-550) */
int cg_mesh_explicit_dt(const cg_mesh* m, double chi, double cm, double max_co, double* dt);

/* -------------------------------------------------------- tissue (fused step) */
typedef struct cg_tissue cg_tissue;

enum { CG_EXPLICIT = 0, CG_IMPLICIT = 1 };      /* solutionAlgorithm */
enum { CG_DDT_EULER = 0, CG_DDT_BACKWARD = 1 }; /* ddtSchemes */

typedef struct {
    int algorithm;     /* CG_EXPLICIT | CG_IMPLICIT */
    int ddt;           /* CG_DDT_EULER | CG_DDT_BACKWARD */
    double chi;        /* 1/m */
    double cm;         /* F/m^2 */
    double dt;         /* s */
    double vm0;        /* initial Vm (V), -0.084 in myocardiumDomain.C:207 */
    double tol;        /* implicit: relative L2 residual tolerance */
    int max_iter;      /* implicit: max inner iterations per solve */
} cg_tissue_config;

/* MyocardiumDomain + MonodomainSolver state on the device: Vm (and old levels),
 * Iion, activation time and flag, the ionic handle's states.  The ionic handle must be
 * a tissue handle with n_cells == mesh cells; it stays owned by the caller. */
int cg_tissue_create(cg_mesh* m, cg_ionic* ionic, const cg_tissue_config* cfg,
                     cg_tissue** out);
int cg_tissue_destroy(cg_tissue* t);

/* externalStimulus boxes (stimulusIO.C:305-390): [3n] min/max corners (m), start (s),
 * duration (s), intensity (A/m^3).  At most 8 boxes. */
int cg_tissue_set_stimulus_boxes(cg_tissue* t, int n, const double* box_min,
                                 const double* box_max, const double* t_start,
                                 const double* duration, const double* intensity);

/* n_steps of the reference's step: t += dt; t0 = t - dt; prepareTimeStep(t0) (box
 * stimulus); advance(t0, dt): solveODE -> solveDiffusion{Explicit,Implicit} ->
 * updateActivationTime (myocardiumDomain.C:487-529).  Asynchronous on the handle's
 * stream; cg_tissue_sync waits. */
int cg_tissue_step(cg_tissue* t, int n_steps);
int cg_tissue_sync(cg_tissue* t);

/* Activation threshold in V: a cell's activation time is the first step whose new Vm
 * exceeds v_threshold (+1e-15), interpolated linearly to v_threshold.  The default, 0, is
 * the reference's (myocardiumDomain.C:362-371).  No reference counterpart otherwise: other
 * thresholds measure how well each cell's activation time is conditioned (a crossing
 * that grazes 0 V moves far under a small Vm perturbation).  Applies to steps not yet
 * run; |v_threshold| < 1. */
int cg_tissue_set_activation_threshold(cg_tissue* t, double v_threshold);

enum { CG_FIELD_VM = 0, CG_FIELD_IION = 1, CG_FIELD_ACTIVATION = 2, CG_FIELD_VM_OLD = 3 };
/* copies a field [n] to host doubles (synchronous) */
int cg_tissue_get_field(cg_tissue* t, int which, double* host);
int cg_tissue_set_field(cg_tissue* t, int which, const double* host);
/* accumulated time (s) and step count */
int cg_tissue_time(const cg_tissue* t, double* time_s, int64_t* steps);
/* implicit path: total inner iterations and last relative residual */
int cg_tissue_solver_stats(const cg_tissue* t, int64_t* iterations, double* last_residual);

/* Implicit path: the fvSolution controls of the reference's Vm solve
 * (tutorials/regressionTests/NiedererEtAl2012/system/fvSolution, solvers{Vm} and PIMPLE):
 *   norm CG_NORM_L2: ||b - A x||_2 <= tolerance ||b||_2 (rel_tol unused);
 *   norm CG_NORM_OPENFOAM: OpenFOAM's lduMatrix residual sum|b - A x| / normFactor,
 *     normFactor = sum |A x0 - sumA xRef| + |b - sumA xRef| (xRef = mean of the initial
 *     guess x0, sumA = row sums), converged below tolerance or below rel_tol x initial;
 *   n_outer_correctors, residual_control: the PIMPLE loop around solveDiffusionImplicit
 *     (monodomainSolver.C:139-150).  Every corrector re-evaluates the explicit tensor
 *     correction from the latest Vm and solves again; after the first, an initial residual
 *     of the previous corrector below residual_control makes the current one the last.
 *     n_outer_correctors <= 1: one solve per step.
 * The preconditioner is Jacobi in both norms (the reference's DIC is a sequential sweep).
 * Replaces the solver-dictionary lookups of fvMatrix::solve and pimpleControl. */
enum { CG_NORM_L2 = 0, CG_NORM_OPENFOAM = 1 };
int cg_tissue_set_solver_controls(cg_tissue* t, int norm, double tolerance, double rel_tol,
                                  int max_iter, int n_outer_correctors,
                                  double residual_control);
/* last step of the implicit path: linear solves (outer correctors), initial residual of the
 * last corrector and final residual of the last solve, both in the controls' norm */
int cg_tissue_solver_report(const cg_tissue* t, int* solves, double* initial_residual,
                            double* final_residual);

/* Per-kernel HIP-event timing on the tissue's stream (the equivalent of the clockTime
 * phases staggeredElectrophysicsAdvanceScheme.C:54-74 stores in
 * electrophysicsAdvanceTimings).  Enabling resets the totals. */
int cg_tissue_set_profiling(cg_tissue* t, int on);
/* Up to max kernels: names (32 chars each), summed milliseconds and launch counts;
 * launches that exit early on a converged solver are not counted.  Synchronises. */
int cg_tissue_kernel_times(cg_tissue* t, int max, char* names, double* total_ms,
                           int64_t* launches, int* n_out);

/* ------------------------------------------- lduMatrix solve on any mesh
 * The Vm equation's lduMatrix of an arbitrary (non-hex) mesh: diag[n_cells] and one
 * coefficient per face (upper == lower, symmetric), addressed by the mesh's lduAddressing
 * (lowerAddr/upperAddr).  Replaces fvMatrix::solve -> lduMatrix::solve with PCG for Vm
 * (monodomainSolver.C:129-135; tolerance/relTol/maxIter of the fvSolution entry), for the
 * meshes the hex-block path does not cover.  The addressing is rebuilt once into ELL rows
 * on the device; the preconditioner is Jacobi (the reference's DIC is a sequential sweep). */
typedef struct cg_ldu cg_ldu;
int cg_ldu_create(int64_t n_cells, int64_t n_faces, const int32_t* lower_addr,
                  const int32_t* upper_addr, cg_ldu** out);
int cg_ldu_destroy(cg_ldu* L);
/* slots per ELL row (the largest number of neighbours of a cell) */
int cg_ldu_width(const cg_ldu* L);
/* lduMatrix::diag() and upper() (flags: CG_DEVICE_PTRS) */
int cg_ldu_set_coeffs(cg_ldu* L, const double* diag, const double* upper, int flags);
/* y = A x (lduMatrix::Amul) */
int cg_ldu_amul(cg_ldu* L, const double* x, double* y, int flags);
/* PCG on A psi = source from the given psi, stopping as lduMatrix::solver does: residual =
 * sum|source - A psi| / normFactor, normFactor = sum |A psi0 - sumA xRef| + |source -
 * sumA xRef| + 1e-20 (xRef = mean of psi0, sumA = row sums); converged below tolerance or
 * below rel_tol x the initial residual, tested before the first iteration; at most
 * max_iter iterations.  Reports the iterations and the initial and final residuals. */
int cg_ldu_solve(cg_ldu* L, double* psi, const double* source, double tolerance,
                 double rel_tol, int max_iter, int flags, int* iterations,
                 double* initial_residual, double* final_residual);
/* milliseconds of the last cg_ldu_amul kernel alone (device pointers), for rooflines */
int cg_ldu_amul_timed(cg_ldu* L, const double* x, double* y, float* kernel_ms);

/* ------------------------------------------------------- standalone operators */
/* Explicit FV Laplacian div(sigma grad V) per unit volume (A/m^3) of the "Gauss linear
 * corrected" scheme with the mesh's conductivity (fvc::laplacian, monodomainSolver.C:111),
 * on the hex block or a general mesh (non-orthogonal correction included there);
 * V and out are device pointers of n doubles.  Synchronous. */
int cg_mesh_laplacian(cg_mesh* m, const double* V_dev, double* out_dev);

/* ------------------------------------------------------------------ adjoint
 * Reverse-mode gradients of Vm at probe cells over n steps of an Euler-ddt, BuenoOrovio
 * FP64 tissue, explicit or implicit, with respect to a per-cell scale kappa of the
 * (uniform, diagonal) conductivity and the model constants tau_fi, tau_si.  No reference
 * counterpart (cardiacFoam has no derivatives); the forward map is the reference's step
 * (myocardiumDomain.C:487-529, monodomainSolver.C:103-136, BuenoOrovio.C:107-152) with
 * sigma_c = kappa_c sigma.  Implicit tissues solve M V^{n+1} = V^n - dt Iion + dt/(chi Cm)
 * src, M = 1 - dt/(chi Cm) L, by Jacobi PCG to a relative residual of 1e-14 in both
 * passes (the adjoint solves M lambda = dLoss/dV^{n+1}). */
typedef struct cg_adjoint cg_adjoint;
/* Snapshots the tissue's current Vm, ionic states, stimulus boxes and time as the start of
 * every forward run.  The forward run checkpoints the state every checkpoint_every steps
 * and the backward run re-runs each segment; when checkpoint_every >= n_steps and every
 * step's state (80 B per cell and step) fits in 64 GiB, all states are kept instead and
 * nothing is re-run. */
int cg_adjoint_create(cg_tissue* t, int n_probes, const int64_t* probe_cells,
                      int checkpoint_every, cg_adjoint** out);
int cg_adjoint_destroy(cg_adjoint* a);
/* From the snapshot: kappa_dev [n_cells] (device), theta = {tau_fi, tau_si} (host, NULL
 * keeps the model's values), Vm (V) after each step at the probes into probes_dev
 * [n_steps][n_probes] (device).  Synchronous.  The tissue handle itself is not advanced.
 * *generation (may be NULL) receives this run's number: every forward run of the handle
 * gets a new one, and only the latest can be differentiated. */
int cg_adjoint_forward(cg_adjoint* a, const double* kappa_dev, const double* theta, int n_steps,
                       double* probes_dev, int64_t* generation);
/* dLoss/dprobes (device [n_steps][n_probes], same layout) of the forward run `generation`
 * with n_steps steps -> dLoss/dkappa (device [n_cells]) and dLoss/d{tau_fi, tau_si} (host
 * [2]).  Re-runs each checkpoint segment, then sweeps it in reverse.  Synchronous.  Returns
 * CG_ERR_STATE, touching nothing, when that run is not the handle's latest forward run
 * (its checkpoints, face coefficients and constants have been replaced) or n_steps
 * differs from it. */
int cg_adjoint_backward(cg_adjoint* a, int64_t generation, int n_steps,
                        const double* dprobes_dev, double* dkappa_dev, double* dtheta);

/* y = A x on the device with the implicit tissue's matrix (the fvm::laplacian + fvm::ddt
 * assembly of monodomainSolver.C:129-135, orthogonal part):
 *   A = ddt_coeff*chi*Cm/dt I + L,  (L x)_P = sum_f sigma_nn,f/h_f^2 (x_P - x_N),
 * per unit volume; ddt_coeff is 1 for Euler and 1.5 for backward.  precision CG_FP64:
 * x, y are double; CG_MIXED: float (the operator of the inner solve, mixed tissues only).
 * Asynchronous on the tissue's stream (cg_tissue_sync waits); timed as kernel "apply"
 * when profiling is on. */
int cg_tissue_apply_operator(cg_tissue* t, double ddt_coeff, int precision, const void* x_dev,
                             void* y_dev);

/* The diffusion part of one step, MonodomainSolver::solveDiffusionExplicit /
 * solveDiffusionImplicit (monodomainSolver.C:103-150), with the ionic current supplied by
 * the caller: for an adapter that keeps its ionic model elsewhere.  Iion [n] in V/s,
 * source [n] in A/m^3 (NULL: the tissue's stimulus boxes at the step's start time), host
 * arrays unless flags has CG_DEVICE_PTRS.  Advances Vm, its old levels, the activation
 * times and the time by one step without calling the tissue's ionic handle.  Runs on the
 * general-mesh operators (on a hex tissue, the block described as a general mesh, built
 * on the first call); implicit solves use the lduMatrix PCG in OpenFOAM's norm with the
 * tolerances and PIMPLE controls of cg_tissue_set_solver_controls (CG_NORM_OPENFOAM) or the
 * tissue's tolerance.  Synchronous. */
int cg_tissue_diffusion_step(cg_tissue* t, const double* Iion, const double* source,
                             int flags);

/* Pseudo-ECG of the current Vm (PseudoECGSolver::compute, pseudoECGSolver.C:36-83, called
 * by ECGDomain::evolve, ecgDomain.C:167-190): for each electrode P,
 *   phi(P) = sum_c ((sigma_c . grad Vm_c) vol_c) . (C_c - P) / |C_c - P|^3
 * over cells with |C_c - P| > VSMALL, grad Vm by Gauss linear with zeroGradient walls.
 * positions [n][3] in m (host), values [n] out (host; the sum itself, no 1/(4 pi sigma)
 * factor, as the reference reports it);
 * 1 <= n <= 256.  Synchronous. */
int cg_tissue_pseudo_ecg(cg_tissue* t, int n_electrodes, const double* positions,
                         double* values);
/* cg_tissue_pseudo_ecg without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot call hipMalloc (serving loops, harnesses that forbid it).
 * Its scratch (electrode positions, per-block partial sums, results: 16 x 1028 doubles,
 * 131,584 B) is a fixed device array of the library, one per GPU, held until the process
 * ends; every call overwrites it entirely, so nothing carries over between calls.
 * 1 <= n <= 16 electrodes (CG_ERR_ARG beyond).  A tissue with per-cell conductivity needs
 * one cg_tissue_pseudo_ecg call first, which uploads that conductivity (CG_ERR_STATE
 * otherwise).  Calls that overlap on one GPU share the scratch and race: serialise them.
 * Results equal cg_tissue_pseudo_ecg's bit for bit.  Synchronous. */
int cg_tissue_pseudo_ecg_noalloc(cg_tissue* t, int n_electrodes, const double* positions,
                                 double* values);

/* ------------------------------------------------------ on-disk ASCII formats
 * Numbers of OpenFOAM's ASCII field files (the internalField List<scalar> that
 * volScalarField::write emits for Vm, activationTime, ionicCurrent and
 * externalStimulusCurrent, myocardiumDomain.C:581-634, and the volTensorField
 * conductivity that monodomainSolver.C:48-100 reads) and of the single-cell trace
 * (ionicModelIO.C:27-85, fixed with 7 decimals per singleCellSolver.C:117-118). */
/* values in the std::ostream form, general with `precision` significant digits (printf
 * %g) or, when fixed != 0, fixed with `precision` decimals; `row` values to a line,
 * separated by one space, each line ended by a newline (row 1: one value per line, the
 * List<scalar> body).  buf has cap bytes; *len gets the bytes written.  CG_ERR_ARG if cap
 * is too small.  Host memory; several threads format slices of a large list. */
int cg_format_scalars(const double* values, int64_t n, int precision, int fixed, int64_t row,
                      char* buf, int64_t cap, int64_t* len);
/* n numbers from text[0, len): whitespace and the brackets '(' ')' separate them, so a
 * List<scalar>, List<vector> or List<tensor> body parses in one call.  *consumed gets the
 * offset just past the n-th number.  CG_ERR_ARG on a malformed number or too few. */
int cg_parse_scalars(const char* text, int64_t len, double* out, int64_t n, int64_t* consumed);

#ifdef __cplusplus
}
#endif
#endif
