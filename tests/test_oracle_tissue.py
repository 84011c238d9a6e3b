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
