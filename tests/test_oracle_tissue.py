"""Pins of the oracle's finite-volume operators by what the mathematics fixes.

- Laplacian of a constant is zero for any conductivity tensor.
- Isotropic/diagonal sigma on the uniform hex = the textbook 7-point finite-difference
  Laplacian with zero-flux walls.
- Off-diagonal sigma: for V = x y the interior divergence is exactly 2 sigma_xy; for a
  linear V it is exactly 0 (Gauss linear corrected with a tensor, gaussLaplacianScheme).
- Implicit matrix: symmetric, row sum = mass term; PCG agrees with scipy's sparse LU.
- Explicit dt: maxCo * h^2 chiCm / sigma_max (myocardiumDomain.C:532-550).
- Box stimulus inclusive on the box faces; activation interpolation.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests.gpu_cases import rotating_fibre_sigma


def _fd7(V, shape, h, s):
    nx, ny, nz = shape
    A = V.reshape(nz, ny, nx)
    out = np.zeros_like(A)
    P = np.pad(A, 1, mode="edge")  # zero-flux walls: ghost = cell
    c = P[1:-1, 1:-1, 1:-1]
    out += s[0] * (P[1:-1, 1:-1, 2:] - 2 * c + P[1:-1, 1:-1, :-2]) / h[0] ** 2
    out += s[1] * (P[1:-1, 2:, 1:-1] - 2 * c + P[1:-1, :-2, 1:-1]) / h[1] ** 2
    out += s[2] * (P[2:, 1:-1, 1:-1] - 2 * c + P[:-2, 1:-1, 1:-1]) / h[2] ** 2
    return out.ravel()


def _centres(shape, h):
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i.ravel() + 0.5) * h[0], (j.ravel() + 0.5) * h[1], (k.ravel() + 0.5) * h[2])


def _tensor(theta, sl=0.17, st=0.019):
    f = np.array([np.cos(theta), np.sin(theta), 0.0])
    return (st * np.eye(3) + (sl - st) * np.outer(f, f)).ravel()


def test_laplacian_constant_is_zero(orc):
    shape, h = (7, 5, 4), (2e-4, 3e-4, 2.5e-4)
    m = orc.OracleMesh(*shape, *h)
    rng = np.random.default_rng(1)
    sig = np.stack([_tensor(t) for t in rng.uniform(-1, 1, m.n)])
    lap = m.laplacian(sig, np.full(m.n, -0.084))
    assert np.max(np.abs(lap)) < 1e-9


def test_diagonal_sigma_is_seven_point_fd(orc):
    shape, h = (9, 6, 5), (2e-4, 3e-4, 2.5e-4)
    s = (0.17, 0.03, 0.019)
    m = orc.OracleMesh(*shape, *h)
    rng = np.random.default_rng(2)
    V = rng.normal(size=m.n)
    sig = np.tile(np.diag(s).ravel(), m.n)
    np.testing.assert_allclose(m.laplacian(sig, V), _fd7(V, shape, h, s), rtol=1e-12, atol=1e-3)


def test_offdiagonal_sigma_closed_forms(orc):
    shape, h = (8, 7, 6), (2e-4, 2e-4, 2e-4)
    m = orc.OracleMesh(*shape, *h)
    sig9 = _tensor(0.7)
    sig = np.tile(sig9, m.n)
    x, y, z = _centres(shape, h)
    nx, ny, nz = shape
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.ravel()
    # V = x y: div(sigma grad V) = 2 sigma_xy
    lap = m.laplacian(sig, x * y)
    np.testing.assert_allclose(lap[inner], 2 * sig9[1], rtol=1e-9)
    # linear V: zero in the interior
    lap = m.laplacian(sig, 3.0 * x - 2.0 * y + 0.5 * z)
    assert np.max(np.abs(lap[inner])) < 1e-6 * np.max(np.abs(sig9)) / h[0] ** 2 * 1e-4
    # the correction alone carries the off-diagonal part (orthogonal part of x y is 0)
    corr = m.correction_div(sig, x * y)
    np.testing.assert_allclose(corr[inner], 2 * sig9[1], rtol=1e-9)


def test_implicit_matrix_and_pcg(orc):
    shape, h = (6, 5, 4), (2e-4, 2e-4, 2e-4)
    m = orc.OracleMesh(*shape, *h)
    rng = np.random.default_rng(3)
    sig = np.stack([_tensor(t) for t in rng.uniform(-1, 1, m.n)])
    chiCm, dt = 1400.0, 2e-5
    diag, upper = m.assemble_implicit(sig, chiCm, dt)
    n = m.n
    A = sp.coo_matrix((np.concatenate([diag, upper, upper]),
                       (np.concatenate([np.arange(n), m.owner, m.neighbour]),
                        np.concatenate([np.arange(n), m.neighbour, m.owner]))), shape=(n, n)).tocsr()
    assert abs(A - A.T).max() == 0
    np.testing.assert_allclose(A @ np.ones(n), chiCm / dt, rtol=1e-12)
    x0 = rng.normal(size=n)
    np.testing.assert_allclose(m.amul(diag, upper, x0), A @ x0, rtol=1e-12)
    b = rng.normal(size=n) * 1e4
    x, it, res = m.pcg(diag, upper, b, np.zeros(n), tol=1e-12)
    assert res <= 1e-12 and it > 0
    np.testing.assert_allclose(x, spla.spsolve(A.tocsc(), b), rtol=1e-9, atol=1e-12)


def test_explicit_dt(orc):
    shape, h = (5, 4, 3), (2e-4, 2e-4, 2e-4)
    m = orc.OracleMesh(*shape, *h)
    sig = np.tile(np.diag([0.1334177215, 0.01760617761, 0.01760617761]).ravel(), m.n)
    dt = m.explicit_dt(sig, 1400.0, 0.1)
    assert dt == pytest.approx(0.1 * h[0] ** 2 * 1400.0 / 0.1334177215, rel=1e-14)


def test_box_stimulus_inclusive(orc):
    shape, h = (10, 10, 10), (1.5e-4, 1.5e-4, 1.5e-4)
    m = orc.OracleMesh(*shape, *h)
    src = np.zeros(m.n)
    bx = orc.make_boxes([0, 0, 0], [1.5e-3, 1.5e-3, 1.5e-3], [0.0], [2e-3], [5e4])
    for t0, on in ((0.0, True), (2e-3, True), (2.0000001e-3, False), (-1e-9, False)):
        orc.lib().orc_box_stimulus(ctypes.byref(m.m), ctypes.byref(bx), t0, src)
        assert (src.sum() > 0) == on
        if on:
            # centres 0.075 .. 1.425 mm: all 1000 cells inside the 1.5 mm cube
            assert np.count_nonzero(src) == 1000 and np.all(src[src > 0] == 5e4)


def test_activation_interpolation(orc):
    n = 4
    V0 = np.array([-0.01, -0.02, 0.0, 0.01])
    V1 = np.array([0.01, -0.01, 1e-16, 0.02])
    flag = np.ones(n, np.uint8)
    flag[3] = 0
    tact = np.zeros(n)
    orc.lib().orc_activation(n, V1, V0, flag, tact, 1.0, 0.1)
    assert tact[0] == pytest.approx(1.05)
    assert tact[1] == 0.0 and flag[1] == 1          # not crossed
    assert tact[2] == 0.0 and flag[2] == 1          # 1e-16 is not > SMALL (1e-15)
    assert tact[3] == 0.0                           # already activated


def _ldu_csr(m, diag, upper):
    n = m.n
    return sp.coo_matrix((np.concatenate([diag, upper, upper]),
                          (np.concatenate([np.arange(n), m.owner, m.neighbour]),
                           np.concatenate([np.arange(n), m.neighbour, m.owner]))),
                         shape=(n, n)).tocsr()


def test_dic_is_exact_cholesky_in_one_dimension(orc):
    """On a chain of cells the lduMatrix is tridiagonal, incomplete Cholesky with zero
    fill is the exact factorisation, and DIC-preconditioned CG converges in one
    iteration; the diagonal preconditioner needs many."""
    m = orc.OracleMesh(40, 1, 1, 2e-4, 2e-4, 2e-4)
    sig = np.tile(np.array([0.17, 0, 0, 0, 0.019, 0, 0, 0, 0.019]), m.n)
    diag, upper = m.assemble_implicit(sig, 1400.0, 2e-5)
    b = np.random.default_rng(2).normal(size=m.n)
    x, it, r0, r1 = m.pcg_of(diag, upper, b, np.zeros(m.n), tol=1e-14, precond=1)
    assert it == 1 and r1 < 1e-14
    np.testing.assert_allclose(x, spla.spsolve(_ldu_csr(m, diag, upper).tocsc(), b), rtol=1e-12)
    _, it_j, _, _ = m.pcg_of(diag, upper, b, np.zeros(m.n), tol=1e-14, precond=0)
    assert it_j > 5


def test_openfoam_normalised_residual(orc):
    """The normalised residual is invariant to scaling the system, zero at the exact
    solution, independent of the preconditioner at the start, and DIC-PCG meets it."""
    shape = (6, 5, 4)
    m = orc.OracleMesh(*shape, 2e-4, 2e-4, 2e-4)
    rng = np.random.default_rng(4)
    sig = np.stack([_tensor(t) for t in rng.uniform(-1, 1, m.n)])
    diag, upper = m.assemble_implicit(sig, 1400.0, 2e-5)
    b = rng.normal(size=m.n) * 1e3 - 80.0
    x0 = rng.normal(size=m.n) * 0.01 - 0.08
    _, _, r0, _ = m.pcg_of(diag, upper, b, x0, tol=1.0, precond=1)
    _, _, r0s, _ = m.pcg_of(7.5 * diag, 7.5 * upper, 7.5 * b, x0, tol=1.0, precond=0)
    assert abs(r0 - r0s) <= 1e-13 * r0
    xs = spla.spsolve(_ldu_csr(m, diag, upper).tocsc(), b)
    _, it, rz, _ = m.pcg_of(diag, upper, b, xs, tol=1e-11, precond=1)
    assert it == 0 and rz < 1e-13
    for pc in (0, 1):
        x, it, ri, rf = m.pcg_of(diag, upper, b, x0, tol=1e-14, precond=pc)
        assert ri == r0 and rf < 1e-14 and it > 0
        assert np.max(np.abs(x - xs)) <= 1e-9 * np.max(np.abs(xs))
    # relTol: stops at relTol x initial
    _, it_r, ri, rf = m.pcg_of(diag, upper, b, x0, tol=0.0, rel_tol=1e-3, precond=1)
    assert rf < 1e-3 * ri and it_r > 0


def test_pimple_outer_correctors_reach_implicit_correction(orc):
    """With outer correctors iterated to convergence, one implicit step solves
    (A - C) V = rhs0 where C is the tensor-correction operator (linear in V, assembled here
    column by column from the oracle's correction_div): the explicit correction becomes
    implicit.  One corrector leaves a difference of the size of C (V - V^n)."""
    shape, h = (6, 5, 4), 2e-4
    nx, ny, nz = shape
    sig = rotating_fibre_sigma(*shape)
    dt, chi, cm = 2e-5, 1.4e5, 0.01
    box = orc.make_boxes([0, 0, 0], [6e-4, 6e-4, 6e-4], [0.0], [2e-3], [5e4])
    runs = {}
    for n_outer in (1, 400):
        T = orc.OracleTissue(nx, ny, nz, h, sig, orc.BO, orc.EPI, dt, chi, cm, integrator=0,
                             implicit=1, norm=1, pcg_tol=1e-15, pcg_max_iter=1000, precond=1,
                             n_outer=n_outer, residual_control=1e-13)
        T.step(box)
        runs[n_outer] = (T.V, T.Iion, T.solver_counts)
    m = orc.OracleMesh(nx, ny, nz, h, h, h)
    n = m.n
    diag, upper = m.assemble_implicit(sig, chi * cm, dt)
    A = _ldu_csr(m, diag, upper).toarray()
    Cm = np.column_stack([m.correction_div(sig, np.eye(n)[j]) for j in range(n)])
    src = np.zeros(n)
    lib = orc.lib()
    lib.orc_box_stimulus(ctypes.byref(m.m), ctypes.byref(box), 0.0, src)
    V0 = np.full(n, -0.084)
    rhs0 = chi * cm / dt * V0 - chi * cm * runs[400][1] + src
    V_exact = np.linalg.solve(A - Cm, rhs0)
    V1, Vk = runs[1][0], runs[400][0]
    e_k = np.max(np.abs(Vk - V_exact)) / np.max(np.abs(V_exact))
    e_1 = np.max(np.abs(V1 - V_exact)) / np.max(np.abs(V_exact))
    assert e_k < 1e-12, e_k
    assert e_1 > 1e3 * e_k
    assert runs[400][2][1] > 2 and runs[1][2][1] == 1


@pytest.mark.parametrize("kind", ["triangles", "renumbered_hex"])
def test_openfoam_pcg_on_unstructured_addressing(orc, kind):
    """The oracle's lduMatrix PCG on addressing that is not a hex block (a Delaunay
    triangulation; a randomly renumbered hex block) against scipy's direct solve, with
    the diagonal and the DIC preconditioner."""
    from tests import ldu_cases as lc
    if kind == "triangles":
        n, lo, up = lc.triangle_mesh(400, seed=1)
    else:
        n, lo, up, _ = lc.renumbered_hex(8, 6, 5, seed=2)
    diag, u = lc.diffusion_coeffs(n, lo, up, seed=3)
    b = np.random.default_rng(4).normal(size=n) * 10.0 + 3.0
    xs = spla.spsolve(lc.csr(n, lo, up, diag, u).tocsc(), b)
    for pc in (0, 1):
        x, it, r0, r1 = orc.ldu_pcg_of(n, lo, up, diag, u, b, np.zeros(n), tol=1e-14,
                                       precond=pc)
        assert r1 < 1e-14 and it > 0
        assert np.max(np.abs(x - xs)) <= 1e-10 * np.max(np.abs(xs))


def test_manufactured_solution_converges_second_order(orc):
    """The oracle's implicit tissue step (backward ddt, RKF45, the manufactured model) on
    the reference's manufactured solution: the error in V at t = 0.05 s falls by about 4
    when the cells halve (dt ~ h^2), and the states stay on their closed form."""
    from tests import manufactured as mf
    errs = []
    for n in (8, 16):
        x, y, z, h = mf.centres(n)
        dt = 0.2 * h * h
        steps = int(round(0.05 / dt))
        dt = 0.05 / steps
        sig = np.array([mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]])
        T = orc.OracleTissue(n, n, n, h, sig, orc.FDA, orc.DIM3, dt, mf.CHI, mf.CM,
                             integrator=1, ddt=1, implicit=1, pcg_tol=1e-13,
                             abs_tol=1e-12, rel_tol=1e-10, nthreads=8)
        V0, u1, u2, u3 = mf.fields(x, y, z, 0.0)
        T.set_V(V0)
        T.set_states(np.stack([V0, u1, u2, u3]))
        for _ in range(steps):
            T.step(None)
        Ve = mf.fields(x, y, z, 0.05)[0]
        errs.append(np.max(np.abs(T.V - Ve)))
    order = np.log2(errs[0] / errs[1])
    assert errs[1] < 2e-3 and 1.7 < order < 2.3, (errs, order)


def test_manufactured_single_cell_closed_form(orc):
    """tutorials/manufacturedSolutions/singleCellMonodomain: one 1D cell at x = 0.5, where
    F = cos(pi x) = 0, so V = 0 and the model (dV/dt = 0) leaves u1 = (1 + t) G and
    u2 = G^-1/2 / (1 + t) with G = 1 + x.  RKF45 at the tutorial's tolerances stays on the
    closed form; Rush-Larsen (forward Euler for both states) converges at first order."""
    from tests import manufactured as mf
    dt, t_end = 5e-5, 0.2                     # the tutorial's controlDict
    steps = int(round(t_end / dt))
    stim = orc.make_stim(start=0.0, period_s1=1000.0, duration=0.0, amplitude=0.0, nstim1=0)
    errs = {}
    for G in (1.5, 1.05, 1.95):
        y0 = [0.0, G, G ** -0.5, 0.0]
        u1e, u2e = mf.single_cell(G, t_end)
        _, v, y = orc.single_cell_run(orc.FDA, orc.DIM1, stim, dt, steps, integrator=1,
                                      abs_tol=1e-6, rel_tol=1e-4, y0=y0)
        assert y[0] == 0.0 and np.all(v == 0.0) and y[3] == 0.0
        assert abs(y[1] - u1e) <= 1e-4 * u1e and abs(y[2] - u2e) <= 1e-4 * u2e, (y, u1e, u2e)
        _, _, y = orc.single_cell_run(orc.FDA, orc.DIM1, stim, dt, steps, integrator=1,
                                      abs_tol=1e-12, rel_tol=1e-12, y0=y0)
        assert abs(y[1] - u1e) <= 1e-10 * u1e and abs(y[2] - u2e) <= 1e-10 * u2e
        for k in (1, 2):
            _, _, y = orc.single_cell_run(orc.FDA, orc.DIM1, stim, dt * 8 / k,
                                          steps // 8 * k, integrator=0, y0=y0)
            errs[G, k] = abs(y[2] - u2e)
        order = np.log2(errs[G, 1] / errs[G, 2])
        assert 0.9 < order < 1.1, (G, errs, order)


def test_adaptive_euler_closed_forms(orc):
    """OpenFOAM's adaptive Euler (`solver Euler;`) on the manufactured single cell: the
    error estimate is the whole increment, so on u' ~ u / (1 + t) every accepted step has
    err = dx / (relTol (1 + t)) and the controller's fixed point 0.9 err^-1/5 = 1 gives
    dx = 0.9^5 relTol (1 + t): ln(1 + T) / (0.9^5 relTol) accepted steps over [0, T]
    (a wrong growth exponent or error estimate moves the count by >= 10 %); the error of a
    first-order step taken at dx ~ relTol falls linearly with relTol."""
    from tests import manufactured as mf
    G, T = 1.5, 0.2
    u1e, u2e = mf.single_cell(G, T)
    errs = []
    for rt in (1e-5, 1e-6):
        y, _, n = orc.ode_solve(orc.EULER, orc.FDA, orc.DIM1, [0.0, G, G ** -0.5, 0.0], 0.0, T,
                                1e-9, 1e-4 * rt, rt)
        pred = np.log(1.0 + T) / (0.9 ** 5 * rt)
        assert abs(n / pred - 1.0) < 2e-3, (rt, n, pred)
        assert y[0] == 0.0 and y[3] == 0.0
        errs.append((abs(y[1] - u1e), abs(y[2] - u2e)))
    for q in (0, 1):
        assert 9.5 < errs[0][q] / errs[1][q] < 10.5, errs
    # RKF45 under the same controller: its fifth-order error estimate lets it take
    # steps orders of magnitude longer for the same tolerance
    y, _, n45 = orc.ode_solve(1, orc.FDA, orc.DIM1, [0.0, G, G ** -0.5, 0.0], 0.0, T, 1e-9,
                              1e-10, 1e-6)
    assert n45 < 1e-3 * n and abs(y[2] - u2e) < 1e-6 * u2e


def test_pseudo_ecg_converges_to_the_continuous_integral(orc):
    """The oracle's pseudo-ECG (pseudoECGSolver.C:36-83) of the manufactured V at the
    monodomainPseudoECG tutorial's electrodes: second-order convergence to the
    Gauss-Legendre quadrature (order 96) of the continuous dipole integral; the two
    electrodes on the line y = z = 1/2 see zero by the symmetry of cos(2 pi y) cos(3 pi z);
    the sum is linear in V and scales with sqrt(1 + t)."""
    from tests import manufactured as mf
    E = np.array(mf.ELECTRODES)
    ref = np.array([mf.pseudo_ecg_reference(P) for P in E])
    assert np.max(np.abs(ref[:2])) < 1e-15 and np.min(np.abs(ref[2:])) > 1e-3
    sig = np.array([mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]])
    errs = []
    for n in (12, 24):
        x, y, z, h = mf.centres(n)
        V = mf.fields(x, y, z, 0.0)[0]
        M = orc.OracleMesh(n, n, n, h, h, h)
        v, ab = M.pseudo_ecg(sig, V, E)
        assert np.max(np.abs(v[:2])) < 1e-13 * ab.max()
        errs.append(np.max(np.abs(v - ref)))
        v3, _ = M.pseudo_ecg(sig, np.sqrt(1.3) * V + 2.0, E)   # a constant adds nothing
        np.testing.assert_allclose(v3, np.sqrt(1.3) * v, rtol=1e-12, atol=1e-15)
    order = np.log2(errs[0] / errs[1])
    assert 1.85 < order < 2.15 and errs[1] < 1e-4, (errs, order)
