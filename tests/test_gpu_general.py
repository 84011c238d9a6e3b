"""The tissue step on a general polyhedral mesh (cg_mesh_create_general): the mesh as the
fvMesh hands it over, fvc::laplacian / fvm::laplacian of a tensor conductivity with Gauss
linear corrected (non-orthogonal correction included), the implicit system solved by the
lduMatrix PCG in OpenFOAM's norm inside the PIMPLE loop.  Checked against the hex path on
an undeformed block, against the oracle's face-by-face restatement (oracle/general.py) on
warped blocks, and by second-order convergence to the manufactured solution on a warped
unit cube (volume-weighted L2 error; dropping the non-orthogonal correction stops the
convergence, which the oracle's test shows)."""
import ctypes
import sys

import numpy as np
import pytest

from tests import manufactured as mf
from tests import poly_mesh as pm
from tests.gpu_cases import rotating_fibre_sigma

pytestmark = pytest.mark.gpu

BOX = {"stimulusLocationMin": [0.0, 0.0, 0.0], "stimulusLocationMax": [6e-4, 6e-4, 6e-4],
       "stimulusDuration": 2e-3, "stimulusIntensity": 5e4, "stimulusStartTime": 0.0}
BOXES = [((0.0, 0.0, 0.0), (6e-4, 6e-4, 6e-4), 0.0, 2e-3, 5e4)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _coeffs(alg, sig):
    return {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells", "chi": 1.4e5,
            "cm": 0.01, "solutionAlgorithm": alg, "conductivity": sig, "externalStimulus": BOX}


def _poly(gpu, m):
    return gpu.PolyMesh(*pm.mesh_args(m))


def test_general_laplacian_vs_oracle(gpu, orc):
    import torch
    from oracle import general as og
    _say("general laplacian")
    shape, L = (10, 9, 8), (2e-3, 1.8e-3, 1.6e-3)
    m = pm.block(*shape, lengths=L, warp=pm.smooth_warp(0.05, L))
    sig = rotating_fibre_sigma(*shape)
    mesh = _poly(gpu, m)
    mesh.set_conductivity(sig)
    V = np.random.default_rng(1).random(m["n"]) * 0.1 - 0.08
    out = torch.empty(m["n"], dtype=torch.float64, device="cuda")
    Vd = torch.tensor(V, device="cuda")
    gpu._capi.check(gpu._capi.lib().cg_mesh_laplacian(mesh.m, ctypes.c_void_p(Vd.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr())))
    ref = og.laplacian(m, sig, V)
    err = np.max(np.abs(out.cpu().numpy() - ref)) / np.max(np.abs(ref))
    _say("  rel err %.3e" % err)
    assert err < 1e-12


@pytest.mark.parametrize("alg", ["explicit", "implicit"])
def test_general_block_equals_hex(gpu, orc, alg):
    """An undeformed block through the general path equals the hex path (rotating fibres,
    box stimulus, backward ddt)."""
    _say("general block vs hex", alg)
    shape, h, dt = (16, 12, 10), 2e-4, 1e-5
    n_steps = 300 if alg == "explicit" else 150
    L = tuple(s * h for s in shape)
    sig = rotating_fibre_sigma(*shape)
    gen = gpu.Monodomain(_poly(gpu, pm.block(*shape, lengths=L)), _coeffs(alg, sig), dt=dt,
                         ddt="backward", tol=1e-13, max_iter=1000)
    hexd = gpu.Monodomain(gpu.HexMesh(*shape, h), _coeffs(alg, sig), dt=dt, ddt="backward",
                          tol=1e-13, max_iter=1000)
    if alg == "implicit":
        for d in (gen, hexd):
            d.set_solver_controls("OpenFOAM", 1e-13, 0.0, 1000, 1, 0.0)
    gen.step(n_steps)
    hexd.step(n_steps)
    err = np.max(np.abs(gen.Vm - hexd.Vm)) / np.max(np.abs(hexd.Vm))
    ta, tb = gen.activationTime, hexd.activationTime
    _say("  Vm rel %.3e, activated %d / %d" % (err, (ta > 0).sum(), (tb > 0).sum()))
    assert err <= (1e-12 if alg == "explicit" else 1e-9)
    assert (tb > 0).sum() > 0 and np.array_equal(ta > 0, tb > 0)
    assert np.max(np.abs(ta - tb)) <= 1e-10


@pytest.mark.parametrize("alg,n_outer", [("explicit", 1), ("implicit", 1), ("implicit", 1000)])
def test_general_warped_vs_oracle(gpu, orc, alg, n_outer):
    """A warped block (faces tilted up to ~15 degrees) with rotating fibres: the device
    against the oracle's face-by-face restatement, the same solves per step."""
    from oracle import general as og
    _say("general warped vs oracle", alg, n_outer)
    shape, dt = (12, 10, 8), 1e-5
    L = (2.4e-3, 2e-3, 1.6e-3)
    n_steps = 200 if alg == "explicit" else 260
    m = pm.block(*shape, lengths=L, warp=pm.smooth_warp(0.05, L))
    sig = rotating_fibre_sigma(*shape)
    dom = gpu.Monodomain(_poly(gpu, m), _coeffs(alg, sig), dt=dt, ddt="backward", tol=1e-13,
                         max_iter=1000)
    rc = 1e-6 if n_outer > 1 else 0.0
    if alg == "implicit":
        dom.set_solver_controls("OpenFOAM", 1e-13, 0.0, 1000, n_outer, rc)
    solves = []
    for _ in range(n_steps):
        dom.step(1)
        if alg == "implicit":
            solves.append(dom.solver_report()[0])
    ref = og.run_tissue(orc, m, sig, 0, 1, dt, 1.4e5, 0.01, n_steps, implicit=alg == "implicit",
                        backward=True, boxes=BOXES, tol=1e-13, n_outer=n_outer,
                        residual_control=rc)
    err = np.max(np.abs(dom.Vm - ref["V"])) / np.max(np.abs(ref["V"]))
    ta = dom.activationTime
    _say("  Vm rel %.3e, activated %d, solves %s / %s" %
         (err, (ref["tact"] > 0).sum(), sorted(set(solves)), sorted(set(ref["solves"]))))
    assert (ref["tact"] > 0).sum() > 0
    assert err <= 1e-9
    assert np.array_equal(ta > 0, ref["tact"] > 0)
    assert np.max(np.abs(ta - ref["tact"])) <= 1e-10
    if alg == "implicit":
        assert solves == ref["solves"]
        if n_outer > 1:
            assert max(solves) > 2


@pytest.mark.parametrize("alg", ["explicit", "implicit"])
def test_general_manufactured_convergence(gpu, orc, alg):
    """monodomainFDAManufactured on a warped unit cube (eps 0.03): the volume-weighted L2
    error against the closed form at t = 0.05 s falls at second order (dt ~ h^2)."""
    _say("general manufactured", alg)
    t_end = 0.05
    errs = []
    for N in (8, 16, 24):
        m = pm.block(N, N, N, warp=pm.smooth_warp(0.03))
        x, y, z = m["C"].T
        h = 1.0 / N
        dt = (0.2 if alg == "implicit" else 0.05) * h * h
        steps = int(round(t_end / dt))
        dt = t_end / steps
        coeffs = {"ionicModel": "monodomainFDAManufactured", "dimension": "3D", "chi": mf.CHI,
                  "cm": mf.CM, "solutionAlgorithm": alg,
                  "conductivity": [mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]],
                  "solver": "RKF45", "absTol": 1e-12, "relTol": 1e-10, "maxSteps": 1000000}
        dom = gpu.Monodomain(_poly(gpu, m), coeffs, dt=dt, ddt="backward", tol=1e-13,
                             max_iter=2000)
        if alg == "implicit":
            dom.set_solver_controls("OpenFOAM", 1e-13, 0.0, 2000, 100, 1e-10)
        V0, u1, u2, u3 = mf.fields(x, y, z, 0.0)
        dom.set_field(gpu._capi.FIELD_VM, V0)
        dom.set_field(gpu._capi.FIELD_VM_OLD, V0)
        dom.ionic.set_states(np.stack([V0, u1, u2, u3]))
        dom.step(steps)
        Ve = mf.fields(x, y, z, t_end)[0]
        e2 = np.sqrt(np.sum((dom.Vm - Ve) ** 2 * m["V"]))
        errs.append(e2)
        _say("  N %d, %d steps, L2 %.3e, Linf %.3e" % (N, steps, e2,
                                                       np.max(np.abs(dom.Vm - Ve))))
    orders = [np.log(errs[i] / errs[i + 1]) / np.log((16, 24)[i] / (8, 16)[i])
              for i in range(2)]
    _say("  orders", orders)
    assert min(orders) > 1.7 and errs[-1] < 3e-4


def test_general_pseudo_ecg(gpu, orc):
    """The pseudo-ECG on a general mesh: equal to the hex path on an undeformed block, and
    to the dipole sum over the oracle's Gauss gradient, cell centres and volumes on a
    warped block with rotating fibres (pseudoECGSolver.C:36-83)."""
    from oracle import general as og
    _say("general pseudo-ECG")
    shape, h = (16, 12, 10), 2e-4
    L = tuple(s * h for s in shape)
    sig = rotating_fibre_sigma(*shape)
    rng = np.random.default_rng(3)
    # electrodes around and inside the block (none on a cell centre: there the hex path's
    # centre is exact and the general path's is a rounding away from it)
    E = rng.uniform(-1, 2, size=(19, 3)) * np.array(L)
    V = -0.084 + 0.1 * rng.random(int(np.prod(shape)))
    gen = gpu.Monodomain(_poly(gpu, pm.block(*shape, lengths=L)), _coeffs("explicit", sig),
                         dt=1e-5)
    hexd = gpu.Monodomain(gpu.HexMesh(*shape, h), _coeffs("explicit", sig), dt=1e-5)
    for d in (gen, hexd):
        d.set_field(gpu._capi.FIELD_VM, V)
    a, b = gen.pseudo_ecg(E), hexd.pseudo_ecg(E)
    scale = orc.OracleMesh(*shape, h, h, h).pseudo_ecg(sig, V, E)[1]
    _say("  block vs hex %.3e" % np.max(np.abs(a - b) / scale))
    assert np.all(np.abs(a - b) <= 1e-13 * scale)
    m = pm.block(*shape, lengths=L, warp=pm.smooth_warp(0.05, L))
    w = gpu.Monodomain(_poly(gpu, m), _coeffs("explicit", sig), dt=1e-5)
    w.set_field(gpu._capi.FIELD_VM, V)
    got = w.pseudo_ecg(E)
    g = og.grad(m, V)
    S = sig.reshape(-1, 3, 3)
    dip = np.einsum("cij,cj->ci", S, g) * m["V"][:, None]
    ref = np.zeros(len(E))
    ab = np.zeros(len(E))
    for e, P in enumerate(E):
        r = m["C"] - P
        rn = np.linalg.norm(r, axis=1)
        ok = rn > 1e-300
        term = np.einsum("ci,ci->c", dip[ok], r[ok]) / rn[ok] ** 3
        ref[e], ab[e] = term.sum(), np.abs(term).sum()
    err = np.abs(got - ref) / ab
    _say("  warped vs oracle %.3e" % err.max())
    assert np.all(err <= 1e-13)


@pytest.mark.parametrize("kind,alg", [("hex", "explicit"), ("hex", "implicit"),
                                      ("warped", "explicit"), ("warped", "implicit")])
def test_diffusion_step_with_caller_ionic(gpu, orc, kind, alg):
    """cg_tissue_diffusion_step: the caller runs the ionic model (solveODE on the host) and
    hands Iion over; the result equals the fused step (cg_tissue_step) of a twin tissue."""
    _say("diffusion entry", kind, alg)
    shape, h, dt = (16, 12, 10), 2e-4, 1e-5
    n_steps = 200 if alg == "explicit" else 220
    L = tuple(s * h for s in shape)
    sig = rotating_fibre_sigma(*shape)

    def mesh():
        if kind == "hex":
            return gpu.HexMesh(*shape, h)
        return _poly(gpu, pm.block(*shape, lengths=L, warp=pm.smooth_warp(0.05, L)))

    a = gpu.Monodomain(mesh(), _coeffs(alg, sig), dt=dt, ddt="backward", tol=1e-13)
    b = gpu.Monodomain(mesh(), _coeffs(alg, sig), dt=dt, ddt="backward", tol=1e-13)
    for d in (a, b):
        if alg == "implicit":
            d.set_solver_controls("OpenFOAM", 1e-13, 0.0, 1000, 1, 0.0)
    n = b.mesh.n
    Im = np.zeros(n)
    for _ in range(n_steps):
        t0 = b.time[0]
        b.ionic.solveODE(t0, dt, b.Vm, Im)
        b.diffusion_step(Im)
    a.step(n_steps)
    err = np.max(np.abs(a.Vm - b.Vm)) / np.max(np.abs(a.Vm))
    ta, tb = a.activationTime, b.activationTime
    _say("  Vm rel %.3e, activated %d / %d" % (err, (ta > 0).sum(), (tb > 0).sum()))
    assert (ta > 0).sum() > 0
    assert err <= (1e-12 if alg == "explicit" else 1e-9)
    assert np.array_equal(ta > 0, tb > 0) and np.max(np.abs(ta - tb)) <= 1e-10
