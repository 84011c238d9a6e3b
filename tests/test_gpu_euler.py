"""OpenFOAM's adaptive Euler ODE solver on the device (`solver Euler;`,
cg_ionic_set_integrator(CG_ODE_EULER)) against the oracle's restatement (oracle/ode.c,
method 2) on the same inputs: the restitution tutorial's single cell
(tutorials/singleCellprotocols/restitutionCurves_s1s2Protocol: BuenoOrovio epi, absTol
1e-6, relTol 1e-4, deltaT 1e-5 s), the tissue-mode ionic step of every model, a fused
explicit tissue run, and the maxSteps error path.  Euler's error estimate is the
increment itself (no cancellation), so the two take the same sub-steps: FP64 bounds are
1e-9 relative.  A mixed-precision (FP32) handle accepts and rejects at other sub-steps,
and two valid step sequences of a first-order method differ by its own error, which
absTol bounds per accepted step: states such as cytosolic calcium (~1e-4 mM, where absTol
1e-6 allows 1 % per step) move by up to a few tolerance units.  The mixed bound is 10
tolerance units, |got - ref| <= 10 (absTol + relTol |ref|), and 5e-3 on Im."""
import sys

import numpy as np
import pytest

from tests.gpu_cases import MODELS, physiological_states

pytestmark = pytest.mark.gpu

# restitutionCurves_s1s2Protocol/constant/electroProperties and system/controlDict
RESTITUTION = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells",
               "solutionAlgorithm": "explicit", "solver": "Euler", "initialODEStep": 1e-5,
               "maxSteps": 1000000000, "absTol": 1e-6, "relTol": 1e-4,
               "singleCellStimulus": {"stim_start": 20, "stim_duration": 1,
                                      "stim_amplitude": 0.4, "stim_period_S1": 2000,
                                      "nstim1": 10, "stim_period_S2": 250, "nstim2": 2},
               "writeAfterTime": 0.0, "outputVariables": {"ionic": {"export": ["Vm"]}}}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_restitution_single_cell_euler(gpu, orc, tmp_path):
    """The restitution tutorial's cell through the dictionary-driven SingleCellSolver
    (`solver Euler;`) over the first beat (0.5 s, 50,000 steps): the written trace equals
    the oracle's Euler run."""
    _say("euler restitution single cell")
    from cardiacfoam_amd.single_cell import SingleCellSolver
    from cardiacfoam_amd import foam_io
    dt, n = 1e-5, 50000
    solver = SingleCellSolver(RESTITUTION, dt, str(tmp_path), precision="fp64")
    assert solver.ionic.solver == "Euler"
    path = solver.run(n * dt, chunk=10000)
    _, data = foam_io.read_trace(path)
    s = RESTITUTION["singleCellStimulus"]
    st = orc.make_stim(start=s["stim_start"], period_s1=s["stim_period_S1"],
                       duration=s["stim_duration"], amplitude=s["stim_amplitude"],
                       nstim1=s["nstim1"], period_s2=s["stim_period_S2"], nstim2=s["nstim2"])
    t, v, y = orc.single_cell_run(orc.BO, orc.EPI, st, dt, n, integrator=orc.EULER,
                                  abs_tol=1e-6, rel_tol=1e-4)
    err = np.max(np.abs(data[:, 1] - v))
    _say("  peak %.4f mV, trace max |device - oracle| %.3e mV" % (data[:, 1].max(), err))
    assert data[:, 1].max() > 30.0                      # the first S1 beat fired
    assert err <= 1e-6                                  # 7-decimal trace rows
    got = solver.ionic.states()[:, 0]
    np.testing.assert_allclose(got, y, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name", ["BO_epi", "TNNP_epi", "ORd", "AP", "CRN", "PYS", "STW",
                                  "GRD", "TOR_M", "GAUR"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ionic_tissue_step_euler(gpu, orc, name, precision):
    """cg_ionic_step with Euler over 20 steps from states along an action potential; the
    per-cell trial steps persist between calls on both sides."""
    _say("euler ionic step", name, precision)
    m, tis, key, word, _ = MODELS[name]
    n = 2048
    c, states, Vm = physiological_states(orc, name, n, seed=5)
    ion = gpu.IonicModel(key, word, n, precision=precision)
    ion.set_integrator("Euler", 1e-6, 1e-4, 1000000)
    ion.set_states(states)
    ref = states.copy()
    dt = 2e-5
    step = np.full(n, dt)
    rng = np.random.default_rng(4)
    worst_im = 0.0
    for k in range(20):
        V = np.ascontiguousarray(Vm + rng.normal(0.0, 2e-4, n))
        Im = np.zeros(n)
        ion.solveODE(k * dt, dt, V, Im)
        Im_ref = orc.ionic_tissue_step(m, tis, c, ref, k * dt, dt, V, integrator=orc.EULER,
                                       step=step, abs_tol=1e-6, rel_tol=1e-4)
        worst_im = max(worst_im, np.max(np.abs(Im - Im_ref)) / np.max(np.abs(Im_ref)))
    got = ion.states()
    scale = np.abs(ref[1:]) + 1e-12 + 1e-6 * np.max(np.abs(ref[1:]), axis=1, keepdims=True)
    rel = np.max(np.abs(got[1:] - ref[1:]) / scale)
    units = np.max(np.abs(got[1:] - ref[1:]) / (1e-6 + 1e-4 * np.abs(ref[1:])))
    _say("  worst Im rel %.3e, worst state rel %.3e, %.2f tolerance units" %
         (worst_im, rel, units))
    if precision == "fp64":
        assert worst_im < 1e-9 and rel < 1e-9
    else:
        assert worst_im < 5e-3 and units <= 10.0


def test_tissue_euler_vs_oracle(gpu, orc):
    """cg_tissue_step (fused explicit kernel, backward ddt) with Euler ionic updates
    against the oracle tissue with integrator Euler."""
    _say("euler tissue")
    shape, h, dt, n_steps = (20, 12, 8), 1.5e-4, 5e-6, 1200
    box = {"stimulusLocationMin": [0.0, 0.0, 0.0], "stimulusLocationMax": [6e-4, 6e-4, 6e-4],
           "stimulusDuration": 2e-3, "stimulusIntensity": 5e4, "stimulusStartTime": 0.0}
    coeffs = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells", "chi": 1.4e5,
              "cm": 0.01, "solutionAlgorithm": "explicit",
              "conductivity": [0.17, 0, 0, 0, 0.019, 0, 0, 0, 0.019], "externalStimulus": box,
              "solver": "Euler", "absTol": 1e-6, "relTol": 1e-4, "maxSteps": 1000000000}
    dom = gpu.Monodomain(gpu.HexMesh(*shape, h), coeffs, dt=dt, ddt="backward")
    dom.step(n_steps)
    nx, ny, nz = shape
    T = orc.OracleTissue(nx, ny, nz, h, np.array(coeffs["conductivity"]), orc.BO, orc.EPI, dt,
                         1.4e5, 0.01, integrator=orc.EULER, ddt=1, abs_tol=1e-6, rel_tol=1e-4,
                         nthreads=8)
    bx = orc.make_boxes(box["stimulusLocationMin"], box["stimulusLocationMax"], [0.0], [2e-3],
                        [5e4])
    for _ in range(n_steps):
        T.step(bx)
    err = np.max(np.abs(dom.Vm - T.V)) / np.max(np.abs(T.V))
    ta, tr = dom.activationTime, T.tact
    _say("  Vm rel err %.3e, activated %d, max dt_act %.3e" %
         (err, (tr > 0).sum(), np.max(np.abs(ta - tr))))
    assert (tr > 0).sum() > 0 and err <= 1e-9
    assert np.array_equal(ta > 0, tr > 0) and np.max(np.abs(ta - tr)) <= 1e-10


def test_euler_max_steps_error(gpu, orc):
    """More sub-steps than maxSteps: the call fails with CG_ERR_ODE (ODESolver's fatal
    error) and the handle keeps working afterwards."""
    _say("euler maxSteps")
    from cardiacfoam_amd._capi import CgError
    ion = gpu.IonicModel("BuenoOrovio", "epicardialCells", 64)
    ion.set_integrator("Euler", 1e-9, 1e-9, 3, initial_step=1e-6)
    Vm = np.full(64, -0.03)
    Im = np.zeros(64)
    with pytest.raises(CgError, match="Euler"):
        ion.solveODE(0.0, 1e-4, Vm, Im)
    ion.set_integrator("Euler", 1e-6, 1e-4, 1000000)
    ion.solveODE(1e-4, 1e-4, Vm, Im)
    assert np.all(np.isfinite(Im))
