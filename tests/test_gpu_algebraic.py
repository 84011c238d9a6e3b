"""The models' RATES and ALGEBRAIC lists on the device (cg_ionic_eval) against the oracle,
and the two places the reference writes them: the single-cell trace's export columns
(SingleCellSolver, ionicModelIO.C:170-225) and the tissue's exported ionic fields
(MyocardiumDomain::exportStates at output times, electroModel.C:321-330).

The device evaluates each row in FP64 with the model's own expressions, so rates and
algebraic values equal the oracle's to 1e-12 relative (the tanh and exp of two libraries
differ in the last bits).  Written files hold 6 significant digits (fields) or 7 decimals
(traces): they are compared to what the oracle's values print as, within one unit of the
last printed digit."""
import os
import sys

import numpy as np
import pytest

from tests.gpu_cases import MODELS, physiological_states

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_rows(orc, m, tis, c, S, t, solve_vm, stim):
    R, A = [], []
    for i in range(S.shape[1]):
        e = orc.model_eval(m, tis, t[i], c, S[:, i], solve_vm=solve_vm, stim=stim)
        R.append(e["rates"])
        a = orc.model_algebraic(m, tis, t[i], S[:, i], stim=stim, consts=c)
        A.append(a if a is not None else [])
    return np.array(R).T, np.array(A).T


@pytest.mark.parametrize("name", ["BO_epi", "BO_endo", "AP", "TNNP_M", "ORd", "CRN", "PYS",
                                  "STW", "GRD", "TOR_M", "GAUR"])
def test_rates_and_algebraic_single_cell(gpu, orc, name):
    """A single-cell handle (V inside the ODE, the handle's stimulus protocol at each row's
    model time): RATES of every model, ALGEBRAIC of BuenoOrovio and AlievPanfilov; the
    other models refuse ALGEBRAIC with CG_ERR_MODEL."""
    _say("eval single cell", name)
    m, tis, key, word, amp = MODELS[name]
    n = 512
    c, S, _ = physiological_states(orc, name, n, seed=11)
    ion = gpu.IonicModel(key, word, n, solve_vm=True)
    stim = {"stim_start": 2.0, "stim_period_S1": 50.0, "stim_duration": 1.0,
            "stim_amplitude": amp, "nstim1": 4}
    ion.set_stimulus_protocol(stim)
    t = np.random.default_rng(2).uniform(0.0, 200.0, n)
    t[:8] = [2.0, 2.5, 3.0, 52.0, 52.999, 152.5, 153.0, 202.5]   # window edges
    st = orc.make_stim(start=2.0, period_s1=50.0, duration=1.0, amplitude=amp, nstim1=4)
    R, A = _oracle_rows(orc, m, tis, c, S, t, True, st)
    got_r = ion.eval("rates", S, t)
    scale = np.max(np.abs(R), axis=1, keepdims=True) + 1e-300
    err_r = np.max(np.abs(got_r - R) / scale)
    _say("  rates: max error %.3e of each rate's range" % err_r)
    assert err_r <= 1e-12
    names = ion.algebraic_names()
    if A.size:
        got_a = ion.eval("algebraic", S, t)
        assert got_a.shape == A.shape == (len(names), n)
        sa = np.max(np.abs(A), axis=1, keepdims=True) + 1e-300
        err_a = np.max(np.abs(got_a - A) / sa)
        _say("  algebraic %s: max error %.3e" % (names, err_a))
        assert err_a <= 1e-12
        ist = got_a[names.index("Istim")]
        assert np.count_nonzero(ist) > 0 and np.count_nonzero(ist == 0.0) > 0
    else:
        assert names == []
        from cardiacfoam_amd._capi import CgError
        with pytest.raises(CgError, match="algebraic"):
            ion.eval("algebraic", S, t)


@pytest.mark.parametrize("name", ["BO_epi", "AP", "TNNP_epi"])
def test_rates_and_algebraic_tissue_handle(gpu, orc, name):
    """A tissue handle: no stimulus, dV/dt = 0 (the reference's solveVmWithinODESolver
    false), and the per-cell conductance scales apply when the rows are the cells."""
    _say("eval tissue", name)
    m, tis, key, word, _ = MODELS[name]
    n = 256
    c, S, _ = physiological_states(orc, name, n, seed=12)
    ion = gpu.IonicModel(key, word, n)
    R, A = _oracle_rows(orc, m, tis, c, S, np.zeros(n), False, None)
    got_r = ion.eval("rates", S)
    assert np.all(got_r[0] == 0.0)
    scale = np.max(np.abs(R), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(got_r - R) / scale) <= 1e-12
    if A.size:
        got_a = ion.eval("algebraic", S)
        names = ion.algebraic_names()
        assert np.all(got_a[names.index("Istim")] == 0.0)
        sa = np.max(np.abs(A), axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(got_a - A) / sa) <= 1e-12


BO_TRACE = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells",
            "solutionAlgorithm": "explicit", "solver": "RKF45", "initialODEStep": 1e-5,
            "maxSteps": 1000000000, "absTol": 1e-6, "relTol": 1e-4,
            "singleCellStimulus": {"stim_start": 20, "stim_duration": 1, "stim_amplitude": 0.4,
                                   "stim_period_S1": 1000, "nstim1": 1, "stim_period_S2": 0,
                                   "nstim2": 0},
            "writeAfterTime": 0,
            "outputVariables": {"ionic": {"export": ["Vm", "v", "w", "s", "Jsi", "Jion",
                                                     "Istim", "tauS", "RATES_w"]}}}


def test_single_cell_trace_algebraic_columns(gpu, orc, tmp_path):
    """The single-cell trace with algebraic columns (the NiedererEtAl2012 export names plus
    others): each row's Jsi, Jion, Istim and tauS equal the oracle's algebraic list at
    that row's states and model time, to the trace's 7 decimals (the states themselves
    printed at 7 decimals, so the comparison takes the printed states)."""
    _say("single-cell trace with algebraic columns")
    from cardiacfoam_amd.single_cell import SingleCellSolver
    from cardiacfoam_amd import foam_io
    dt, n = 5e-5, 600                                    # 30 ms: rest, stimulus, upstroke
    solver = SingleCellSolver(BO_TRACE, dt, str(tmp_path), precision="fp64")
    names, data = foam_io.read_trace(solver.run(n * dt, chunk=200))
    assert names == ["time", "Vm", "v", "w", "s", "Jsi", "Jion", "Istim", "tauS", "RATES_w"]
    c, _ = orc.model_init(orc.BO, orc.EPI)
    st = orc.make_stim(start=20.0, period_s1=1000.0, duration=1.0, amplitude=0.4, nstim1=1)
    worst = 0.0
    for row in data:
        y = np.array([(row[1] + 84.0) / 85.7, row[2], row[3], row[4]])
        a = orc.model_algebraic(orc.BO, orc.EPI, row[0] * 1000.0, y, stim=st, consts=c)
        e = orc.model_eval(orc.BO, orc.EPI, row[0] * 1000.0, c, y, solve_vm=True, stim=st)
        want = np.array([a[2], a[3], a[4], a[9], e["rates"][2]])
        # the printed states carry 5e-8 each; Jion's sensitivity to u is O(10)
        tol = np.array([1e-6, 2e-5, 5e-8, 1e-4, 1e-6]) + 1e-6 * np.abs(want)
        err = np.abs(row[5:] - want)
        if min(abs(row[0] * 1000.0 - 20.0), abs(row[0] * 1000.0 - 21.0)) < 1e-6:
            err[2] = 0.0        # a pulse edge: the printed time decides which side
        worst = max(worst, np.max(err / tol))
        assert np.all(err <= tol), (row[0], row[5:], want)
    stim_rows = data[(data[:, 0] > 0.0201) & (data[:, 0] < 0.0209)]
    _say("  %d rows, worst error %.2f of tolerance, Istim during the pulse %s" %
         (len(data), worst, np.unique(stim_rows[:, 7])))
    assert len(stim_rows) > 0 and np.all(stim_rows[:, 7] != 0.0)
    assert np.all(data[data[:, 0] < 0.0199, 7] == 0.0)


def test_tissue_exported_fields(gpu, orc, tmp_path):
    """Monodomain with outputVariables.ionic.export (Vm Jsi s RATES_v tauWMinus unknownName):
    written after the domain's own fields, dimensionless, each the raw model-unit value of
    the cells after the last step (V = (Vm_old*1000 + 84)/85.7, algebraic and rates at the
    states with no stimulus); the unknown name is dropped with a warning, and the file Vm
    ends up holding the V state as in the reference's write order."""
    _say("tissue exported fields")
    from cardiacfoam_amd import foam_io
    shape, h, dt = (12, 10, 6), 2e-4, 1e-5
    box = {"stimulusLocationMin": [0, 0, 0], "stimulusLocationMax": [6e-4, 6e-4, 6e-4],
           "stimulusDuration": 2e-3, "stimulusIntensity": 5e4, "stimulusStartTime": 0.0}
    coeffs = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells", "chi": 1.4e5,
              "cm": 0.01, "solutionAlgorithm": "explicit",
              "conductivity": [0.17, 0, 0, 0, 0.019, 0, 0, 0, 0.019], "externalStimulus": box,
              "outputVariables": {"ionic": {"export": ["Vm", "Jsi", "s", "RATES_v",
                                                       "tauWMinus", "unknownName"]}}}
    dom = gpu.Monodomain(gpu.HexMesh(*shape, h), coeffs, dt=dt, ddt="backward")
    dom.step(400)
    vm_old = dom.field(gpu._capi.FIELD_VM_OLD)
    vm = dom.Vm
    S = dom.ionic.states()
    with pytest.warns(UserWarning, match="unknownName"):
        d = dom.write_fields(str(tmp_path))
    c, _ = orc.model_init(orc.BO, orc.EPI)
    u = (vm_old * 1000.0 + 84.0) / 85.7
    Y = np.vstack([u, S[1:]])
    A = np.array([orc.model_algebraic(orc.BO, orc.EPI, 0.0, Y[:, i], consts=c)
                  for i in range(dom.mesh.n)]).T
    R = np.array([orc.model_eval(orc.BO, orc.EPI, 0.0, c, Y[:, i])["rates"]
                  for i in range(dom.mesh.n)]).T
    want = {"Vm": u, "Jsi": A[2], "s": S[3], "RATES_v": R[1], "tauWMinus": A[8]}
    assert not os.path.exists(os.path.join(d, "unknownName"))
    for name, ref in want.items():
        f = foam_io.read_field(os.path.join(d, name))
        got = f.expand(dom.mesh.n)
        printed = np.array([float("%.6g" % x) for x in ref])
        ulp6 = 10.0 ** (np.floor(np.log10(np.abs(ref) + 1e-300)) - 5)
        err = np.max(np.abs(got - printed) / ulp6)
        _say("  %-10s max |file - oracle| %.2f units of the 6th digit" % (name, err))
        assert err <= 1.0, name
        with open(os.path.join(d, name)) as fh:
            assert "dimensions      [0 0 0 0 0 0 0];" in fh.read()
    assert np.max(np.abs(vm - vm_old)) > 0.0           # the tissue moved over the last step
    assert (u > 0.5).sum() > 0                         # some cells are activated


def test_coupling_signals_and_io_states(gpu, orc):
    """ionicModel::hasSignal / signal (ionicModel.C:120-204) on the host mirror: VM in mV
    through the model's Vm transform of the V state the last step was fed, CAI from the
    state named like Ca_i (TNNP: Ca_i; none for BuenoOrovio); vmRest from the initial
    state; stateVariableValues of one integration point."""
    _say("coupling signals")
    n = 64
    Vm = np.linspace(-0.086, 0.02, n)
    Im = np.zeros(n)
    bo = gpu.IonicModel("BuenoOrovio", "epicardialCells", n)
    assert bo.vmRest() == pytest.approx(-0.084, abs=1e-15)      # u = 0: 85.7*0 - 84 mV
    bo.solveODE(0.0, 1e-5, Vm, Im)
    Vm_next = Vm + 1e-3                  # the caller's Vm moves on; the fed copy does not
    np.testing.assert_allclose(bo.signals("VM"), Vm * 1000.0, rtol=0, atol=1e-12)
    assert bo.signal(5, "VM") == pytest.approx(Vm[5] * 1000.0, abs=1e-12)
    assert bo.hasSignal("VM") and not bo.hasSignal("CAI")
    from cardiacfoam_amd import FatalIOError
    with pytest.raises(FatalIOError):
        bo.signals("CAI")
    with pytest.raises(FatalIOError):
        bo.signal(n, "VM")
    tn = gpu.IonicModel("TNNP", "epicardialCells", n)
    assert tn.hasSignal("CAI")
    c, y0 = orc.model_init(orc.TNNP, orc.EPI)
    assert tn.vmRest() == pytest.approx(y0[0] * 1e-3, rel=1e-15)
    tn.solveODE(0.0, 2e-5, Vm_next, Im)
    S = tn.states()
    names = tn.stateVariableNames()
    np.testing.assert_array_equal(tn.signals("CAI"), S[names.index("Ca_i")])
    np.testing.assert_allclose(tn.signals("VM"), Vm_next * 1000.0, rtol=1e-15)
    v = tn.stateVariableValues(7)
    assert v.shape == (len(names),) and v[0] == pytest.approx(Vm_next[7] * 1000.0, rel=1e-15)
    np.testing.assert_array_equal(v[1:], S[1:, 7])
