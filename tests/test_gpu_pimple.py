"""Implicit path with the reference's fvSolution semantics (cg_tissue_set_solver_controls):
OpenFOAM's normalised residual and the PIMPLE outer correctors that re-evaluate the
explicit tensor correction from the latest Vm (monodomainSolver.C:139-150), against the
oracle's restatement on the same seeded inputs (rotating fibres, so the correction is
non-zero).  FP64: Vm within 1e-9 relative and the same number of linear solves per step;
mixed: within the 0.5 mV / 0.05 ms gate of the FP64 run."""
import sys

import numpy as np
import pytest

from tests.gpu_cases import MODELS, rotating_fibre_sigma

pytestmark = pytest.mark.gpu

BOX = {"stimulusLocationMin": [0.0, 0.0, 0.0], "stimulusLocationMax": [6e-4, 6e-4, 6e-4],
       "stimulusDuration": 2e-3, "stimulusIntensity": 5e4, "stimulusStartTime": 0.0}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _coeffs(key, word, sig):
    return {"ionicModel": key, "tissue": word, "chi": 1.4e5, "cm": 0.01,
            "solutionAlgorithm": "implicit", "conductivity": sig, "externalStimulus": BOX}


def _oracle(orc, shape, h, sig, m, tis, dt, n_steps, tol, n_outer, rc, rel=0.0):
    T = orc.OracleTissue(*shape, h, sig, m, tis, dt, 1.4e5, 0.01, integrator=0, implicit=1,
                         norm=1, pcg_tol=tol, pcg_rel_tol=rel, pcg_max_iter=1000, precond=0,
                         n_outer=n_outer, residual_control=rc)
    bx = orc.make_boxes(BOX["stimulusLocationMin"], BOX["stimulusLocationMax"],
                        [BOX["stimulusStartTime"]], [BOX["stimulusDuration"]],
                        [BOX["stimulusIntensity"]])
    solves = []
    for _ in range(n_steps):
        before = T.solver_counts[1]
        T.step(bx)
        solves.append(T.solver_counts[1] - before)
    return T, solves


@pytest.mark.parametrize("name,shape", [("BO_epi", (20, 14, 10)), ("ORd", (19, 14, 10))])
def test_pimple_openfoam_norm_fp64_vs_oracle(gpu, orc, name, shape):
    """nOuterCorrectors 1000, residualControl 1e-6 (the reference's PIMPLE dictionary) and
    the Vm solver in OpenFOAM's norm; quads (nx = 20) and single cells (nx = 19)."""
    _say("pimple fp64", name, shape)
    m, tis, key, word, _ = MODELS[name]
    h, dt, n_steps, tol = 2e-4, 2e-5, 150, 1e-13
    sig = rotating_fibre_sigma(*shape)
    mesh = gpu.HexMesh(*shape, h)
    dom = gpu.Monodomain(mesh, _coeffs(key, word, sig), dt=dt)
    dom.set_solver_controls("OpenFOAM", tol, 0.0, 1000, 1000, 1e-6)
    solves = []
    for _ in range(n_steps):
        dom.step(1)
        solves.append(dom.solver_report()[0])
    T, solves_ref = _oracle(orc, shape, h, sig, m, tis, dt, n_steps, tol, 1000, 1e-6)
    err = np.max(np.abs(dom.Vm - T.V)) / np.max(np.abs(T.V))
    ta, tr = dom.activationTime, T.tact
    _say("  Vm rel err %.3e, solves per step device %s oracle %s, activated %d" %
         (err, sorted(set(solves)), sorted(set(solves_ref)), (tr > 0).sum()))
    assert solves == solves_ref
    assert max(solves) > 3                       # the correction needed extra correctors
    assert err <= 1e-9
    assert np.array_equal(ta > 0, tr > 0)
    assert np.max(np.abs(ta - tr)) <= 1e-10
    n, r0, r1 = dom.solver_report()
    assert r1 < tol and r0 < 1e-6


def test_pimple_diagonal_sigma_three_solves(gpu, orc):
    """Diagonal conductivity: no correction, so the second corrector starts converged and
    the third is the last (the same count as the oracle); the result equals one plain
    solve."""
    _say("pimple diagonal")
    shape, h, dt = (16, 12, 8), 2e-4, 2e-5
    sig = [0.17, 0, 0, 0, 0.019, 0, 0, 0, 0.019]
    mesh = gpu.HexMesh(*shape, h)
    a = gpu.Monodomain(mesh, _coeffs("BuenoOrovio", "epicardialCells", sig), dt=dt)
    a.set_solver_controls("OpenFOAM", 1e-13, 0.0, 1000, 1000, 1e-6)
    b = gpu.Monodomain(mesh, _coeffs("BuenoOrovio", "epicardialCells", sig), dt=dt)
    b.set_solver_controls("OpenFOAM", 1e-13, 0.0, 1000, 1, 0.0)
    for _ in range(40):
        a.step(1)
        b.step(1)
        assert a.solver_report()[0] == 3 and b.solver_report()[0] == 1
    assert np.max(np.abs(a.Vm - b.Vm)) <= 1e-12 * np.max(np.abs(b.Vm))


def test_openfoam_norm_single_corrector_vs_oracle(gpu, orc):
    """One corrector per step in OpenFOAM's norm with relTol: the same solution as the
    oracle's lduMatrix PCG."""
    _say("openfoam norm relTol")
    m, tis, key, word, _ = MODELS["BO_epi"]
    shape, h, dt, n_steps = (20, 14, 10), 2e-4, 2e-5, 100
    sig = rotating_fibre_sigma(*shape)
    mesh = gpu.HexMesh(*shape, h)
    dom = gpu.Monodomain(mesh, _coeffs(key, word, sig), dt=dt)
    dom.set_solver_controls("OpenFOAM", 1e-14, 1e-13, 1000, 1, 0.0)
    dom.step(n_steps)
    T, _ = _oracle(orc, shape, h, sig, m, tis, dt, n_steps, 1e-14, 1, 0.0, rel=1e-13)
    err = np.max(np.abs(dom.Vm - T.V)) / np.max(np.abs(T.V))
    _say("  Vm rel err %.3e" % err)
    assert err <= 1e-9


def test_pimple_mixed_vs_fp64(gpu, orc):
    """Mixed path with the same controls: within 0.5 mV and 0.05 ms of the FP64 path."""
    _say("pimple mixed")
    name, shape = "ORd", (20, 14, 10)
    m, tis, key, word, _ = MODELS[name]
    h, dt, n_steps = 2e-4, 2e-5, 200
    sig = rotating_fibre_sigma(*shape)
    mesh = gpu.HexMesh(*shape, h)
    out = {}
    for prec in ("fp64", "mixed"):
        dom = gpu.Monodomain(mesh, _coeffs(key, word, sig), dt=dt, precision=prec)
        dom.set_solver_controls("OpenFOAM", 1e-11, 0.0, 1000, 1000, 1e-6)
        dom.step(n_steps)
        out[prec] = (dom.Vm, dom.activationTime, dom.solver_report())
    dv = np.max(np.abs(out["mixed"][0] - out["fp64"][0])) * 1e3
    act = (out["fp64"][1] > 0) & (out["mixed"][1] > 0)
    da = np.max(np.abs(out["mixed"][1][act] - out["fp64"][1][act])) * 1e3 if act.any() else 0.0
    _say("  Vm Linf %.4f mV, activation %.2e ms, reports %s %s" %
         (dv, da, out["fp64"][2], out["mixed"][2]))
    assert dv <= 0.5 and da <= 0.05
    assert np.array_equal(out["fp64"][1] > 0, out["mixed"][1] > 0)
