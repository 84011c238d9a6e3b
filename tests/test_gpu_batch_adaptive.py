"""Batched single cells (cg_batch_run) under RKF45 and OpenFOAM's Euler per protocol step,
each cell keeping its trial step, against the oracle's single-cell run with the same
integrator, tolerances and protocol; FP64 Gaur RKF45 batches are refused.

The device's protocol steps start at s*dt (model units) and the oracle's at the
accumulated time ((t + dt) - dt in seconds); the ulp-level difference moves the accepted
sub-steps next to a stimulus edge, after which two correct implementations agree to about
relTol (the single-cell RKF45 tests measure the same), so the bound is 10 (relTol |y| +
absTol): RKF45 at absTol 1e-10, relTol 1e-8; Euler at absTol 1e-3, relTol 1e-4 (Euler's
error estimate is the whole increment, so while V crosses 0 mV its sub-steps are at most
absTol / |dV/dt|)."""
import sys

import numpy as np
import pytest

from tests.gpu_cases import MODELS

pytestmark = pytest.mark.gpu

DT_MS = 0.01
N_STEPS = 40000          # 400 ms: one paced beat and most of its repolarisation


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


TOL = {"RKF45": (1e-10, 1e-8), "Euler": (1e-3, 1e-4)}


def _case(gpu, name, precision, solver, n, seed, tol):
    m, tis, key, word, amp = MODELS[name]
    rng = np.random.default_rng(seed)
    starts = 5.0025 + rng.uniform(0.0, 5.0, n)
    prots = [{"stim_start": float(s), "stim_period_S1": 1000.0, "stim_duration": 1.0,
              "stim_amplitude": amp, "nstim1": 1} for s in starts]
    lens = N_STEPS - rng.integers(0, 2000, n)
    ion = gpu.IonicModel(key, word, n, solve_vm=True, precision=precision)
    ion.set_protocols(prots)
    ion.set_run_lengths(np.asarray(lens, dtype=np.int64))
    ion.set_integrator(solver, tol[0], tol[1], 1000000)
    ion.enable_beats(max_beats=4, dvdt_threshold=5.0, baseline_s=2e-3, peak_window_s=2.0)
    return ion, prots, lens


@pytest.mark.parametrize("solver", ["RKF45", "Euler"])
@pytest.mark.parametrize("name", ["TNNP_epi", "ORd", "CRN"])
def test_batch_adaptive_fp64_vs_oracle(gpu, orc, name, solver):
    """FP64: each cell's final states equal the oracle's single-cell run with the same
    integrator to 10 (relTol |y| + absTol); one beat per cell."""
    _say("batch adaptive fp64", name, solver)
    m, tis = MODELS[name][0], MODELS[name][1]
    n = 3
    at, rt = TOL[solver]
    ion, prots, lens = _case(gpu, name, "fp64", solver, n, seed=11, tol=(at, rt))
    gpu.batch_run([ion], DT_MS * 1e-3, chunk=7919)    # chunk edges fall mid-beat
    cnt, _ = ion.beats()
    got = ion.states()
    integ = {"RKF45": 1, "Euler": 2}[solver]
    for i in range(n):
        p = prots[i]
        st = orc.make_stim(start=p["stim_start"], period_s1=p["stim_period_S1"],
                           duration=p["stim_duration"], amplitude=p["stim_amplitude"],
                           nstim1=p["nstim1"])
        _, _, y = orc.single_cell_run(m, tis, st, DT_MS * 1e-3, int(lens[i]), integrator=integ,
                                      abs_tol=at, rel_tol=rt)
        rel = np.max(np.abs(got[:, i] - y) / (np.abs(y) + 1e-12))
        excess = np.max(np.abs(got[:, i] - y) / (rt * np.abs(y) + at))
        _say("  cell %d: %d steps, %d beats, Vm %.9f (oracle %.9f), worst state rel %.3e"
             % (i, lens[i], cnt[i], got[0, i], y[0], rel))
        assert cnt[i] == 1
        assert excess <= 10.0, excess


@pytest.mark.parametrize("name", ["TNNP_epi", "ORd", "CRN"])
def test_batch_rkf45_mixed_vs_fp64(gpu, orc, name):
    """Mixed precision under RKF45 (absTol 1e-6, relTol 1e-4, tolerances FP32 can meet)
    against the FP64 batch at the same tolerances: same beat count, APD90 within 0.5 ms,
    final Vm within 0.5 mV."""
    _say("batch rkf45 mixed", name)
    n = 64
    h64, _, _ = _case(gpu, name, "fp64", "RKF45", n, seed=12, tol=(1e-6, 1e-4))
    h32, _, _ = _case(gpu, name, "mixed", "RKF45", n, seed=12, tol=(1e-6, 1e-4))
    gpu.batch_run([h64, h32], DT_MS * 1e-3, chunk=10000)
    (c64, r64), (c32, r32) = h64.beats(), h32.beats()
    np.testing.assert_array_equal(c64, c32)
    apd = lambda r: (r[:, 0, 4] - r[:, 0, 0]) * 1e3
    d_apd = np.nanmax(np.abs(apd(r64) - apd(r32)))
    u = MODELS[name]
    vm = lambda h: h.states()[0] * (85.7 if u[0] == 0 else 1.0)
    d_vm = np.max(np.abs(vm(h64) - vm(h32)))
    _say("  APD90 max diff %.4f ms, final Vm max diff %.4f mV" % (d_apd, d_vm))
    assert d_apd <= 0.5 and d_vm <= 0.5


def test_batch_autotune_refuses_adaptive(gpu):
    ion = gpu.IonicModel("TNNP", "epicardialCells", 4, solve_vm=True)
    ion.set_protocols([{"stim_start": 5.0, "stim_period_S1": 1000.0, "stim_duration": 1.0,
                        "stim_amplitude": 52.0, "nstim1": 1}] * 4)
    ion.set_run_lengths(np.full(4, 100, dtype=np.int64))
    ion.set_integrator("RKF45", 1e-6, 1e-4, 100000)
    with pytest.raises(gpu._capi.CgError, match="autotuning"):
        gpu.batch_autotune([ion], DT_MS * 1e-3, trial_steps=10)


def test_batch_refuses_fp64_gaur_rkf45(gpu):
    for key, word, solver, prec, msg in (("Gaur", "myocyte", "RKF45", "fp64", "Gaur"),):
        ion = gpu.IonicModel(key, word, 4, solve_vm=True, precision=prec)
        ion.set_protocols([{"stim_start": 5.0, "stim_period_S1": 1000.0,
                            "stim_duration": 1.0, "stim_amplitude": 40.0, "nstim1": 1}] * 4)
        ion.set_run_lengths(np.full(4, 100, dtype=np.int64))
        ion.set_integrator(solver, 1e-6, 1e-4, 100000)
        with pytest.raises(gpu._capi.CgError, match=msg):
            gpu.batch_run([ion], DT_MS * 1e-3, chunk=50)
