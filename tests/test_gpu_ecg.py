"""Pseudo-ECG on the device (cg_tissue_pseudo_ecg, PseudoECGSolver::compute of
pseudoECGSolver.C:36-83): equal to the oracle's dipole sum on the same Vm (uniform and
per-cell rotating-fibre conductivity, electrodes outside and inside the block), and, on the
manufactured V of tutorials/manufacturedSolutions/monodomainPseudoECG, second-order
convergence to the Gauss-Legendre quadrature of the continuous integral."""
import sys

import numpy as np
import pytest

from tests import manufactured as mf
from tests.gpu_cases import rotating_fibre_sigma

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tissue(gpu, shape, h, sigma):
    coeffs = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells", "chi": 1.4e5,
              "cm": 0.01, "solutionAlgorithm": "implicit", "conductivity": sigma}
    return gpu.Monodomain(gpu.HexMesh(*shape, h), coeffs, dt=2e-5)


@pytest.mark.parametrize("rotating", [False, True])
def test_pseudo_ecg_equals_oracle(gpu, orc, rotating):
    _say("pseudo-ECG vs oracle, rotating", rotating)
    shape, h = (40, 24, 18), 2e-4
    sig = rotating_fibre_sigma(*shape) if rotating else [0.17, 0, 0, 0, 0.019, 0, 0, 0, 0.019]
    dom = _tissue(gpu, shape, h, sig)
    n = int(np.prod(shape))
    V = -0.084 + 0.1 * np.random.default_rng(4).random(n)
    dom.set_field(gpu._capi.FIELD_VM, V)
    ext = np.array(shape) * h
    rng = np.random.default_rng(5)
    # 20 electrodes: two kernel passes of 16 and 4
    E = np.concatenate([rng.uniform(-1, 2, size=(16, 3)) * ext,           # around the block
                        [[0.5 * h, 0.5 * h, 0.5 * h]],                    # on a cell centre
                        rng.uniform(0.1, 0.9, size=(3, 3)) * ext])        # inside
    got = dom.pseudo_ecg(E)
    M = orc.OracleMesh(*shape, h, h, h)
    ref, ab = M.pseudo_ecg(np.asarray(sig), V, E)
    err = np.abs(got - ref) / ab
    _say("  %d electrodes, worst |diff| / sum|terms| %.3e" % (len(E), err.max()))
    assert np.all(err <= 1e-13)
    # the same Vm once more: identical bits (fixed reduction order)
    assert np.array_equal(dom.pseudo_ecg(E), got)


def test_pseudo_ecg_manufactured_convergence(gpu, orc):
    _say("pseudo-ECG manufactured")
    E = np.array(mf.ELECTRODES)
    ref = np.array([mf.pseudo_ecg_reference(P) for P in E])
    sig = [mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]]
    errs = []
    for n in (16, 32, 64):
        x, y, z, h = mf.centres(n)
        dom = _tissue(gpu, (n, n, n), h, sig)
        dom.set_field(gpu._capi.FIELD_VM, mf.fields(x, y, z, 0.0)[0])
        v = dom.pseudo_ecg(E)
        errs.append(np.max(np.abs(v - ref)))
        _say("  n %d: %s, max error %.3e" % (n, np.array2string(v, precision=6), errs[-1]))
        assert np.max(np.abs(v[:2])) < 1e-12
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    _say("  orders", orders)
    assert np.all((orders > 1.9) & (orders < 2.1)) and errs[-1] < 1e-5
