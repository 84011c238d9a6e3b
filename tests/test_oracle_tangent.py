"""The oracle's tangent-linear BuenoOrovio tissue run (oracle/tangent.cpp), the reference
gradient of the adjoint over a whole horizon.

- Its value trajectory is orc_tissue_step's (explicit, Euler ddt, Rush-Larsen) bit for bit.
- Its directional derivatives equal central differences of the oracle's own loss while no
  threshold crossing moves (a short horizon), for kappa directions and for tau_fi, tau_si.
- Directions ride in lanes of one pass; a direction's derivative does not depend on the
  lane or on the other directions beside it.
"""
import numpy as np
import pytest

SIG0 = np.array([0.17, 0.019, 0.019])
DT = 1e-5
BOX = ([0.0, 0.0, 0.0], [6e-4, 6e-4, 6e-4], [0.0], [2e-3], [5e4])


def _case(seed=7, shape=(12, 10, 6), steps=300, n_probes=24):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    cells = rng.choice(n, n_probes, replace=False)
    kappa = rng.uniform(0.8, 1.25, n)
    target = -0.084 + 0.1 * rng.random((steps, n_probes))
    return rng, shape, n, cells, kappa, target


def _oracle_loss(orc, shape, kappa, theta, cells, target, steps):
    n = kappa.size
    sig = np.zeros((n, 9))
    sig[:, 0], sig[:, 4], sig[:, 8] = kappa * SIG0[0], kappa * SIG0[1], kappa * SIG0[2]
    T = orc.OracleTissue(*shape, 2e-4, sig, orc.BO, orc.EPI, DT, 1.4e5, 0.01, integrator=0,
                         ddt=0, nthreads=4)
    T.set_const(orc.const_index(orc.BO, "tauFi"), theta[0])
    T.set_const(orc.const_index(orc.BO, "tauSi"), theta[1])
    bx = orc.make_boxes(*BOX)
    L = 0.0
    for s in range(steps):
        T.step(bx)
        L += float(np.sum((T.V[cells] - target[s]) ** 2))
    return L, T.V.copy()


def test_tangent_values_are_the_oracle_trajectory(orc):
    rng, shape, n, cells, kappa, target = _case()
    theta = np.array([0.11, 1.8875])
    L, _, v = orc.bo_tangent(shape, 2e-4, SIG0, orc.EPI, DT, 1.4e5, 0.01, orc.make_boxes(*BOX),
                             kappa, theta, np.zeros((1, n)), np.zeros((1, 2)), 300, cells,
                             target, nthreads=4)
    Lo, vo = _oracle_loss(orc, shape, kappa, theta, cells, target, 300)
    np.testing.assert_array_equal(v, vo)
    assert L == pytest.approx(Lo, rel=1e-13)


def test_tangent_equals_central_differences(orc):
    """300 steps (3 ms, the stimulus and the first spread) at eps = 1e-6, where no switch
    moves: agreement to the differences' own truncation and rounding (~1e-7)."""
    rng, shape, n, cells, kappa, target = _case()
    theta = np.array([0.11, 1.8875])
    dk = np.zeros((4, n))
    dk[:2] = rng.standard_normal((2, n))
    dth = np.zeros((4, 2))
    dth[2, 0], dth[3, 1] = 0.11, 1.8875
    _, d, _ = orc.bo_tangent(shape, 2e-4, SIG0, orc.EPI, DT, 1.4e5, 0.01, orc.make_boxes(*BOX),
                             kappa, theta, dk, dth, 300, cells, target, nthreads=4)
    eps = 1e-6
    for k in range(4):
        fp = _oracle_loss(orc, shape, kappa + eps * dk[k], theta + eps * dth[k], cells, target, 300)[0]
        fm = _oracle_loss(orc, shape, kappa - eps * dk[k], theta - eps * dth[k], cells, target, 300)[0]
        fd = (fp - fm) / (2 * eps)
        assert d[k] == pytest.approx(fd, rel=2e-6), (k, d[k], fd)


def test_tangent_directions_are_independent_lanes(orc):
    """Nine directions take two passes; each equals its own single-direction run."""
    rng, shape, n, cells, kappa, target = _case(seed=3, steps=120)
    theta = np.array([0.11, 1.8875])
    dk = rng.standard_normal((9, n))
    dth = rng.standard_normal((9, 2)) * [0.01, 0.1]
    bx = orc.make_boxes(*BOX)
    _, d, _ = orc.bo_tangent(shape, 2e-4, SIG0, orc.EPI, DT, 1.4e5, 0.01, bx, kappa, theta, dk,
                             dth, 120, cells, target, nthreads=4)
    for k in (0, 5, 8):
        _, d1, _ = orc.bo_tangent(shape, 2e-4, SIG0, orc.EPI, DT, 1.4e5, 0.01, bx, kappa, theta,
                                  dk[k:k + 1], dth[k:k + 1], 120, cells, target, nthreads=4)
        assert d1[0] == d[k]
    # linear in the direction
    _, ds, _ = orc.bo_tangent(shape, 2e-4, SIG0, orc.EPI, DT, 1.4e5, 0.01, bx, kappa, theta,
                              (dk[0] + dk[1])[None], (dth[0] + dth[1])[None], 120, cells, target,
                              nthreads=4)
    assert ds[0] == pytest.approx(d[0] + d[1], rel=1e-12)
