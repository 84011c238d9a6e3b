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
