"""GPU ionic step (cg_ionic_step / cg_ionic_run_cells) against the oracle on the same
seeded inputs.  FP64: agreement to 1e-10 relative after 20 steps (transcendental
routines differ by ulps between the host and device libraries).  Mixed: FP32 states
and arithmetic, compared with the FP64 oracle at tolerances stated per quantity."""
import sys

import numpy as np
import pytest

from tests.gpu_cases import MODELS, physiological_states

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ionic_tissue_step_parity(gpu, orc, name, precision):
    _say("ionic parity", name, precision)
    m, tis, key, word, _ = MODELS[name]
    n = 4096
    c, states, Vm = physiological_states(orc, name, n, seed=11)
    ion = gpu.IonicModel(key, word, n, solve_vm=False, precision=precision)
    ion.set_states(states)
    ref_states = states.copy()
    rng = np.random.default_rng(5)
    dt = 1e-5 if m != 0 else 5e-6
    worst_im = 0.0
    for k in range(20):
        V = Vm + rng.normal(0.0, 2e-4, n)
        Im = np.zeros(n)
        ion.solveODE(k * dt, dt, np.ascontiguousarray(V), Im)
        Im_ref = orc.ionic_tissue_step(m, tis, c, ref_states, k * dt, dt, V, integrator=0)
        scale = max(np.max(np.abs(Im_ref)), 1e-30)
        worst_im = max(worst_im, np.max(np.abs(Im - Im_ref)) / scale)
    got = ion.states()
    ref = ref_states
    rel = np.max(np.abs(got[1:] - ref[1:]) / (np.abs(ref[1:]) + 1e-12 + 1e-6 * np.max(np.abs(ref[1:]), axis=1, keepdims=True)))
    _say("  worst Im rel", worst_im, "worst state rel", rel)
    if precision == "fp64":
        assert worst_im < 1e-10
        assert rel < 1e-10
    else:
        assert worst_im < 2e-3
        assert rel < 1e-3


@pytest.mark.parametrize("name", ["BO_endo", "TNNP_epi", "ORd", "AP", "CRN", "PYS", "STW",
                                  "GRD", "TOR_epi", "TOR_M", "TOR_endo", "GAUR"])
def test_single_cell_run_parity(gpu, orc, name):
    """cg_ionic_run_cells (Vm inside the ODE, stimulus protocol) vs the oracle's RL
    single cell, stimulus edges placed off the step grid."""
    _say("single-cell parity", name)
    m, tis, key, word, amp = MODELS[name]
    ion = gpu.IonicModel(key, word, 3, solve_vm=True, precision="fp64")
    stim = {"stim_start": 10.0025, "stim_period_S1": 500.0, "stim_duration": 1.0,
            "stim_amplitude": amp, "nstim1": 2}
    ion.set_stimulus_protocol(stim)
    dt = 1e-5
    n_steps = 60000 if m != 3 else 40000
    ion.run_cells(0.0, dt, n_steps)
    got = ion.states()[:, 0]
    c, y = orc.model_init(m, tis)
    st = orc.make_stim(start=10.0025, period_s1=500.0, duration=1.0, amplitude=amp, nstim1=2)
    yy = np.zeros(orc.MAX_STATES)
    yy[:len(y)] = y
    dtm = dt * 1000.0 / 12.9 if m == 3 else dt * 1000.0
    for k in range(n_steps):
        orc.lib().orc_rl_step(m, tis, k * dtm, dtm, c, yy, 1, st)
    ref = yy[:len(y)]
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("name,integ", [("ORd", "RL"), ("TNNP_epi", "RL"), ("CRN", "RL"),
                                        ("ORd", "RKF45")])
def test_mixed_resting_concentration_drift(gpu, orc, name, integ):
    """Resting cells (V held at the initial potential) for 2000 steps of 0.02 ms: Na_i and
    K_i move by ~1e-8 per step, far below half an FP32 ulp of their values.  The mixed
    path stores them as offsets from a reference and must follow the FP64 path's drift
    (a plain FP32 store leaves them where they started)."""
    _say("mixed resting drift", name, integ)
    m, tis, key, word, _ = MODELS[name]
    n, steps, dt = 256, 2000, 2e-5
    _, y0 = orc.model_init(m, tis)
    V = np.full(n, y0[0] * 1e-3)
    res = {}
    for prec in ("fp64", "mixed"):
        ion = gpu.IonicModel(key, word, n, solve_vm=False, precision=prec)
        if integ == "RKF45":
            ion.set_integrator("RKF45", 1e-6, 1e-4, 100000)
        Im = np.zeros(n)
        for k in range(steps):
            ion.solveODE(k * dt, dt, V, Im)
        res[prec] = ion.states()
    idx = {"ORd": (2, 4), "TNNP_epi": (2, 1), "CRN": (1, 2)}[name]   # Na_i, K_i
    for q in idx:
        change = res["fp64"][q] - y0[q]
        err = res["mixed"][q] - res["fp64"][q]
        _say("  state %d: fp64 change %.3e, mixed - fp64 %.3e" % (q, change[0], np.max(np.abs(err))))
        assert np.max(np.abs(change)) > 1e-6            # the drift is real
        assert np.max(np.abs(err)) <= 0.02 * np.max(np.abs(change)) + 1e-7


def test_mixed_state_round_trip(gpu, orc):
    """set_states/get_states on an FP32 handle go through the offset storage: values come
    back to FP32 precision of the offset, far finer than FP32 of the value itself."""
    _say("mixed state round trip")
    c, states, _ = physiological_states(orc, "ORd", 128, seed=2)
    ion = gpu.IonicModel("ORd", "myocyte", 128, precision="mixed")
    ion.set_states(states)
    got = ion.states()
    # K_i ~ 145 mM stored as an offset from 145: error ~ulp(offset), not ulp(145) = 1.5e-5
    assert np.max(np.abs(got[4] - states[4])) < 2e-6
    np.testing.assert_allclose(got[1:], states[1:], rtol=2e-7, atol=1e-12)
