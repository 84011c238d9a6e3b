"""Pins of the oracle's general-mesh restatement (oracle/general.py): it reduces to the
hex oracle on an undeformed block, annihilates constants on warped meshes, and on a warped
unit cube its implicit manufactured run converges at second order in the volume-weighted L2
norm; with the non-orthogonal correction planted out, the same run does not converge (so
the convergence test sees that term).  Also the test mesh generator's closure."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests import manufactured as mf
from tests import poly_mesh as pm
from tests.gpu_cases import rotating_fibre_sigma


@pytest.fixture(scope="module")
def og():
    from oracle import general
    return general


def test_mesh_generator_closed_cells():
    m = pm.block(6, 5, 4, warp=pm.smooth_warp(0.04))
    S = np.zeros((m["n"], 3))
    np.add.at(S, m["owner"], m["Sf"])
    np.add.at(S, m["neighbour"], -m["Sf"])
    np.add.at(S, m["b_owner"], m["b_Sf"])
    assert np.abs(S).max() < 1e-15
    assert abs(m["V"].sum() - 1.0) < 1e-14 and m["V"].min() > 0
    d = m["C"][m["neighbour"]] - m["C"][m["owner"]]
    cos = np.einsum("ij,ij->i", d, m["Sf"]) / np.linalg.norm(d, axis=1) / np.linalg.norm(
        m["Sf"], axis=1)
    assert np.degrees(np.arccos(cos.min())) > 10.0      # really non-orthogonal


def test_general_reduces_to_hex(orc, og):
    nx, ny, nz, h = 7, 6, 5, 0.1
    m = pm.block(nx, ny, nz, lengths=(nx * h, ny * h, nz * h))
    sig = rotating_fibre_sigma(nx, ny, nz)
    V = np.random.default_rng(0).random(m["n"])
    ref = orc.OracleMesh(nx, ny, nz, h, h, h).laplacian(sig, V)
    assert np.max(np.abs(og.laplacian(m, sig, V) - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_constant_field_has_zero_laplacian(og):
    m = pm.block(8, 8, 8, warp=pm.smooth_warp(0.04))
    lap = og.laplacian(m, rotating_fibre_sigma(8, 8, 8), np.full(m["n"], 3.0))
    assert np.max(np.abs(lap)) < 1e-12


def _manufactured_l2(orc, og, N, eps, t_end=0.05):
    m = pm.block(N, N, N, warp=pm.smooth_warp(eps))
    x, y, z = m["C"].T
    sig = np.array([mf.SIGMA[0], 0, 0, 0, mf.SIGMA[1], 0, 0, 0, mf.SIGMA[2]])
    h = 1.0 / N
    dt = 0.2 * h * h
    steps = int(round(t_end / dt))
    dt = t_end / steps
    chiCm = mf.CHI * mf.CM
    V0, u1, u2, u3 = mf.fields(x, y, z, 0.0)
    st = np.ascontiguousarray(np.stack([V0, u1, u2, u3]))
    c, _ = orc.model_init(orc.FDA, orc.DIM3)
    V, Vold = V0.copy(), V0.copy()
    step = np.full(m["n"], dt)
    solve = {}
    for s in range(steps):
        Im = orc.ionic_tissue_step(orc.FDA, orc.DIM3, c, st, s * dt, dt, V, integrator=1,
                                   step=step, abs_tol=1e-12, rel_tol=1e-10)
        cc, c0, c00 = (1.5, 2.0, 0.5) if s >= 1 else (1.0, 1.0, 0.0)
        if cc not in solve:
            d, up = og.matrix(m, sig, chiCm, cc, dt)
            idx = np.arange(m["n"])
            A = sp.coo_matrix((np.concatenate([d, up, up]),
                               (np.concatenate([idx, m["owner"], m["neighbour"]]),
                                np.concatenate([idx, m["neighbour"], m["owner"]])))).tocsc()
            solve[cc] = spla.factorized(A)
        b0 = m["V"] * (chiCm / dt * (c0 * V - c00 * Vold) - chiCm * Im)
        psi = V.copy()
        for _ in range(40):     # the corrector loop to its fixed point
            f, bf = og.face_fluxes(m, sig, psi, orth=False)
            new = solve[cc](b0 + og.divergence(m, f, bf))
            done = np.max(np.abs(new - psi)) < 1e-13
            psi = new
            if done:
                break
        Vold, V = V, psi
    Ve = mf.fields(x, y, z, t_end)[0]
    return np.sqrt(np.sum((V - Ve) ** 2 * m["V"]))


def test_manufactured_convergence_on_a_warped_cube(orc, og):
    e8 = _manufactured_l2(orc, og, 8, 0.03)
    e16 = _manufactured_l2(orc, og, 16, 0.03)
    order = np.log2(e8 / e16)
    assert 1.8 < order < 2.2 and e16 < 5e-4, (e8, e16, order)


def test_dropping_the_nonorthogonal_correction_breaks_convergence(orc, og, monkeypatch):
    """A planted mistake in the oracle (correction vectors k = 0): the warped-cube error
    no longer falls with h, so the convergence pin above does see that term."""
    real = og.face_fluxes

    def no_k(m, sigma9, V, orth=True):
        mm = dict(m)
        f, b = real(mm, sigma9, V, orth)
        # remove SfGammaSn k . interp(grad V) from every internal face flux
        o, nb = m["owner"], m["neighbour"]
        S = og._sigma_cells(m, sigma9).reshape(-1, 3, 3)
        w = og.weights(m)
        n = m["Sf"] / np.linalg.norm(m["Sf"], axis=1)[:, None]
        d = m["C"][nb] - m["C"][o]
        D = 1.0 / np.maximum(np.einsum("ij,ij->i", n, d), 0.05 * np.linalg.norm(d, axis=1))
        k = n - D[:, None] * d
        sig_f = w[:, None, None] * S[o] + (1.0 - w)[:, None, None] * S[nb]
        gsn = np.einsum("fi,fij,fj->f", m["Sf"], sig_f, n)
        g = og.grad(m, V)
        gf = w[:, None] * g[o] + (1.0 - w)[:, None] * g[nb]
        return f - gsn * np.einsum("fi,fi->f", k, gf), b

    monkeypatch.setattr(og, "face_fluxes", no_k)
    e8 = _manufactured_l2(orc, og, 8, 0.03)
    e16 = _manufactured_l2(orc, og, 16, 0.03)
    assert e16 > 0.9 * e8, (e8, e16)
