"""lduMatrix solve on arbitrary addressing (cg_ldu_*): ELL rows rebuilt from
lowerAddr/upperAddr on the device, y = A x against scipy, and PCG with OpenFOAM's
normalised residual against the oracle's lduMatrix PCG (same Jacobi preconditioner, so the
same iterates up to rounding: iteration counts within one) and scipy's direct solve."""
import sys

import numpy as np
import pytest

from tests import ldu_cases as lc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(kind):
    if kind == "triangles":
        n, lo, up = lc.triangle_mesh(3000, seed=11)
    elif kind == "renumbered_hex":
        n, lo, up, _ = lc.renumbered_hex(24, 16, 12, seed=12)
    else:  # a hex block in its own order, 7-point rows
        n, lo, up, _ = lc.renumbered_hex(20, 12, 10, seed=None)
    diag, u = lc.diffusion_coeffs(n, lo, up, seed=13)
    return n, lo, up, diag, u


@pytest.mark.parametrize("kind", ["triangles", "renumbered_hex"])
def test_ldu_amul_matches_scipy(gpu, orc, kind):
    _say("ldu amul", kind)
    n, lo, up, diag, u = _system(kind)
    L = gpu.LduMatrix(lo, up, n)
    L.set_coeffs(diag, u)
    x = np.random.default_rng(1).normal(size=n)
    y = L.Amul(x)
    ref = lc.csr(n, lo, up, diag, u) @ x
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    _say("  width %d, rel err %.3e" % (L.width, err))
    assert err < 1e-14


@pytest.mark.parametrize("kind", ["triangles", "renumbered_hex"])
def test_ldu_solve_vs_oracle_and_scipy(gpu, orc, kind):
    _say("ldu solve", kind)
    import scipy.sparse.linalg as spla
    n, lo, up, diag, u = _system(kind)
    rng = np.random.default_rng(2)
    b = rng.normal(size=n) * 10.0 + 3.0
    x0 = rng.normal(size=n) * 0.1
    L = gpu.LduMatrix(lo, up, n)
    L.set_coeffs(diag, u)
    psi = x0.copy()
    perf = L.solve(psi, b, tolerance=1e-13, relTol=0.0, maxIter=1000)
    xo, it_o, r0_o, r1_o = orc.ldu_pcg_of(n, lo, up, diag, u, b, x0, tol=1e-13, precond=0)
    xs = spla.spsolve(lc.csr(n, lo, up, diag, u).tocsc(), b)
    _say("  device %s, oracle it %d r0 %.6e r1 %.3e" % (perf, it_o, r0_o, r1_o))
    assert abs(perf["initialResidual"] - r0_o) <= 1e-12 * r0_o
    assert abs(perf["nIterations"] - it_o) <= 1
    assert perf["finalResidual"] < 1e-13
    assert np.max(np.abs(psi - xs)) <= 1e-10 * np.max(np.abs(xs))
    assert np.max(np.abs(psi - xo)) <= 1e-10 * np.max(np.abs(xs))


def test_ldu_solve_device_tensors_and_reltol(gpu, orc):
    """Device pointers, relTol stopping, a zero-iteration solve from the exact answer, and
    the hex block through lduAddressing equal to the same system solved by scipy."""
    import torch
    import scipy.sparse.linalg as spla
    _say("ldu device tensors")
    n, lo, up, diag, u = _system("hex")
    L = gpu.LduMatrix(lo, up, n)
    L.set_coeffs(torch.tensor(diag, device="cuda"), torch.tensor(u, device="cuda"))
    assert L.width == 6
    b = np.random.default_rng(3).normal(size=n)
    psi = torch.zeros(n, dtype=torch.float64, device="cuda")
    perf = L.solve(psi, torch.tensor(b, device="cuda"), tolerance=0.0, relTol=1e-6, maxIter=500)
    _say("  relTol", perf)
    assert perf["finalResidual"] < 1e-6 * perf["initialResidual"] and perf["nIterations"] > 0
    xs = spla.spsolve(lc.csr(n, lo, up, diag, u).tocsc(), b)
    psi2 = xs.copy()
    perf2 = L.solve(psi2, b, tolerance=1e-10, maxIter=500)
    assert perf2["nIterations"] == 0
    # maxIter caps the iterations
    psi3 = np.zeros(n)
    perf3 = L.solve(psi3, b, tolerance=1e-30, maxIter=3)
    assert perf3["nIterations"] == 3
