"""The reference's manufactured solution on the device (monodomainPseudoECG tutorial:
monodomainFDAManufactured, implicit, backward ddt, RKF45; closed form of
manufacturedFDAReference.H): the device matches the oracle step for step, and its error
against the closed form falls at second order as the cells halve with dt ~ h^2.  This pins
the implicit diffusion operator, BDF2 and the ionic coupling against mathematics rather
than against a restatement."""
import sys

import numpy as np
import pytest

from tests import manufactured as mf

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _coeffs(alg, solver):
    return {"ionicModel": "monodomainFDAManufactured", "dimension": "3D", "chi": mf.CHI,
            "cm": mf.CM, "solutionAlgorithm": alg,
            "conductivity": [mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]],
            "solver": solver, "absTol": 1e-12, "relTol": 1e-10, "maxSteps": 1000000}


def _run(gpu, n, t_end, alg="implicit", solver="RKF45", precision="fp64"):
    x, y, z, h = mf.centres(n)
    dt = 0.2 * h * h if alg == "implicit" else 0.05 * h * h
    steps = int(round(t_end / dt))
    dt = t_end / steps
    mesh = gpu.HexMesh(n, n, n, h)
    dom = gpu.Monodomain(mesh, _coeffs(alg, solver), dt=dt, ddt="backward", tol=1e-13,
                         max_iter=5000, precision=precision)
    V0, u1, u2, u3 = mf.fields(x, y, z, 0.0)
    dom.set_field(gpu._capi.FIELD_VM, V0)
    dom.set_field(gpu._capi.FIELD_VM_OLD, V0)
    dom.ionic.set_states(np.stack([V0, u1, u2, u3]))
    dom.step(steps)
    return dom, x, y, z, dt, steps


def test_manufactured_device_equals_oracle(gpu, orc):
    _say("manufactured vs oracle")
    n, t_end = 8, 0.05
    dom, x, y, z, dt, steps = _run(gpu, n, t_end)
    h = 1.0 / n
    sig = np.array([mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]])
    T = orc.OracleTissue(n, n, n, h, sig, orc.FDA, orc.DIM3, dt, mf.CHI, mf.CM, integrator=1,
                         ddt=1, implicit=1, pcg_tol=1e-13, abs_tol=1e-12, rel_tol=1e-10,
                         nthreads=8)
    V0, u1, u2, u3 = mf.fields(x, y, z, 0.0)
    T.set_V(V0)
    T.set_states(np.stack([V0, u1, u2, u3]))
    for _ in range(steps):
        T.step(None)
    err = np.max(np.abs(dom.Vm - T.V)) / np.max(np.abs(T.V))
    st = dom.ionic.states()
    serr = np.max(np.abs(st[1:3] - T.states[1:3]) / np.abs(T.states[1:3]))
    _say("  Vm rel %.3e, u1/u2 rel %.3e" % (err, serr))
    assert err <= 1e-9 and serr <= 1e-9


@pytest.mark.parametrize("alg,solver", [("implicit", "RKF45"), ("implicit", "RushLarsen"),
                                        ("explicit", "RKF45")])
def test_manufactured_second_order(gpu, orc, alg, solver):
    _say("manufactured order", alg, solver)
    t_end = 0.05
    errs = []
    for n in (16, 32):
        dom, x, y, z, dt, steps = _run(gpu, n, t_end, alg, solver)
        Ve = mf.fields(x, y, z, t_end)[0]
        errs.append(np.max(np.abs(dom.Vm - Ve)))
        _say("  n %d, %d steps, Vm Linf error %.3e" % (n, steps, errs[-1]))
    order = np.log2(errs[0] / errs[1])
    _say("  observed order %.3f" % order)
    assert 1.8 < order < 2.3
    assert errs[1] < 5e-4


def test_manufactured_mixed_precision(gpu, orc):
    """The mixed path on the same case stays within the FP64 discretisation error."""
    _say("manufactured mixed")
    t_end = 0.05
    d64, x, y, z, _, _ = _run(gpu, 32, t_end)
    d32, _, _, _, _, _ = _run(gpu, 32, t_end, precision="mixed")
    Ve = mf.fields(x, y, z, t_end)[0]
    e64 = np.max(np.abs(d64.Vm - Ve))
    e32 = np.max(np.abs(d32.Vm - Ve))
    d = np.max(np.abs(d32.Vm - d64.Vm))
    _say("  error fp64 %.3e mixed %.3e, mixed - fp64 %.3e" % (e64, e32, d))
    assert d < 0.05 * e64


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_manufactured_single_cells_closed_form(gpu, orc, precision):
    """tutorials/manufacturedSolutions/singleCellMonodomain through cg_ionic_run_cells
    (RKF45 at the tutorial's tolerances, dt 5e-5 s to 0.2 s): cells with V = 0 and
    G = 1 + x over x in (0, 1) stay on u1 = (1 + t) G, u2 = G^-1/2 / (1 + t), V = 0; FP64
    equals the oracle's single-cell RKF45; Rush-Larsen (forward Euler) converges at first
    order.  Mixed (FP32 states) stays within the rounding drift of 4000 FP32 updates."""
    _say("manufactured single cells", precision)
    n = 1000
    G = 1.0 + (np.arange(n) + 0.5) / n
    y0 = np.stack([np.zeros(n), G, G ** -0.5, np.zeros(n)])
    dt, t_end = 5e-5, 0.2
    steps = int(round(t_end / dt))
    u1e, u2e = mf.single_cell(G, t_end)
    stim = {"stim_start": 0.0, "stim_duration": 0.0, "stim_amplitude": 0.0,
            "stim_period_S1": 1000.0, "nstim1": 0, "stim_period_S2": 0.0, "nstim2": 0}
    coeffs = {"ionicModel": "monodomainFDAManufactured", "dimension": "1D", "solver": "RKF45",
              "maxSteps": 1000000000, "absTol": 1e-6, "relTol": 1e-4,
              "singleCellStimulus": stim}
    # initialDeltaT is runTime.deltaT, as singleCellSolver.C:48-57 passes it
    ion = gpu.IonicModel.New(coeffs, n, dt, True, precision=precision)
    ion.set_states(y0)
    ion.run_cells(0.0, dt, steps)
    y = ion.states()
    e1 = np.max(np.abs(y[1] - u1e) / u1e)
    e2 = np.max(np.abs(y[2] - u2e) / u2e)
    _say("  RKF45 u1 rel %.3e, u2 rel %.3e" % (e1, e2))
    assert np.all(y[0] == 0.0) and np.all(y[3] == 0.0)
    # FP32 states: 4000 rounded increments bound the drift by steps * 2^-23
    bound = 1e-10 if precision == "fp64" else steps * 2.0 ** -23
    assert e1 <= bound and e2 <= bound
    if precision == "mixed":
        return
    s = orc.make_stim(start=0.0, period_s1=1000.0, duration=0.0, amplitude=0.0, nstim1=0)
    for c in (0, 499, 999):
        _, _, yo = orc.single_cell_run(orc.FDA, orc.DIM1, s, dt, steps, integrator=1,
                                       abs_tol=1e-6, rel_tol=1e-4, y0=y0[:, c])
        np.testing.assert_allclose(y[:, c], yo, rtol=1e-12, atol=0)
    errs = []
    for k in (1, 2):
        ion = gpu.IonicModel.New(dict(coeffs, solver="RushLarsen"), n, dt, True)
        ion.set_states(y0)
        ion.run_cells(0.0, 8 * dt / k, steps // 8 * k)
        errs.append(np.max(np.abs(ion.states()[2] - u2e)))
    order = np.log2(errs[0] / errs[1])
    _say("  Rush-Larsen u2 errors %.3e %.3e, order %.3f" % (errs[0], errs[1], order))
    assert 0.9 < order < 1.1
