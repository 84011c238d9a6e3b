"""Batched single cells (cg_batch_run): per-cell S1/S2 protocols, uneven run lengths,
per-cell conductance scales and on-device beat detection, against the oracle's
single-cell Rush-Larsen trace and its detect_beats restatement
(postProcessing_restCurves.py:115-200) on the same seeded inputs."""
import sys

import numpy as np
import pytest

from tests.gpu_cases import MODELS

pytestmark = pytest.mark.gpu

DT_MS = 0.01
THR, BASE_MS, PEAK_MS = 5.0, 2.0, 2000.0   # 5000 mV/s, 2 ms, 2 s


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _protocols(rng, n, amp, s2=False):
    """Seeded S1 (and S2) trains with stimulus edges off the step grid."""
    out = []
    for _ in range(n):
        cl = float(rng.uniform(300.0, 700.0))
        d = {"stim_start": 5.0025 + float(rng.uniform(0, 5)), "stim_period_S1": cl,
             "stim_duration": 1.0, "stim_amplitude": amp, "nstim1": int(rng.integers(2, 4))}
        if s2:
            d["stim_period_S2"] = float(rng.uniform(250.0, 450.0))
            d["nstim2"] = 1
        out.append(d)
    return out


def _run_len(p, rng):
    """Steps to the end of the last stimulus train plus a seeded tail (uneven lengths)."""
    end = p["stim_start"] + p["stim_period_S1"] * p["nstim1"]
    end += p.get("stim_period_S2", 0.0) * p.get("nstim2", 0)
    return int((end + rng.uniform(0.0, 300.0)) / DT_MS)


def _oracle_cell(orc, m, tis, p, n_steps, scale=None):
    c, y = orc.model_init(m, tis)
    if scale is not None:
        names = {1: ["g_Na", "g_Kr", "g_CaL", "g_Ks"], 2: ["GNa", "GKr", "PCa", "GKs"],
                 5: ["gNa", "gKr", "gCaL", "gKs"], 6: ["gNa", "gKr", "gCaL", "gKs"],
                 7: ["g_Na", "g_Kr", "g_CaL", "g_Ks"]}[m]
        for k, nm in enumerate(names):
            c[orc.const_index(m, nm)] *= scale[k]
    st = orc.make_stim(start=p["stim_start"], period_s1=p["stim_period_S1"],
                       duration=p["stim_duration"], amplitude=p["stim_amplitude"],
                       nstim1=p["nstim1"], period_s2=p.get("stim_period_S2", 0.0),
                       nstim2=p.get("nstim2", 0))
    t, v, yf = orc.single_cell_trace(m, tis, st, DT_MS, n_steps, consts=c)
    beats = orc.detect_beats(t, v, dvdt_threshold=THR, baseline_s=BASE_MS, peak_s=PEAK_MS)
    return beats, yf


def _ref_records(beats):
    """oracle beats (ms, mV) -> records in s and V as the C-ABI returns them"""
    rec = []
    for b in beats:
        rec.append([b["t_up"] * 1e-3, b["v_base"] * 1e-3, b["v_peak"] * 1e-3, b["v90"] * 1e-3,
                    b["t_repol90"] * 1e-3 if b["repol_found"] else np.nan])
    return np.array(rec)


def _setup(gpu, name, n, precision, seed, s2=False, scales=True):
    m, tis, key, word, amp = MODELS[name]
    rng = np.random.default_rng(seed)
    prots = _protocols(rng, n, amp, s2)
    lens = np.array([_run_len(p, rng) for p in prots], dtype=np.int64)
    sc = rng.uniform(0.8, 1.2, (4, n)) if scales else None
    ion = gpu.IonicModel(key, word, n, solve_vm=True, precision=precision)
    ion.set_protocols(prots)
    ion.set_run_lengths(lens)
    if sc is not None:
        ion.set_conductance_scale(sc)
    ion.enable_beats(max_beats=16, dvdt_threshold=THR, baseline_s=BASE_MS * 1e-3,
                     peak_window_s=PEAK_MS * 1e-3)
    return ion, prots, lens, sc


@pytest.mark.parametrize("name,n,s2", [("TNNP_epi", 6, False), ("ORd", 3, True),
                                       ("CRN", 3, True), ("PYS", 3, False),
                                       ("STW", 3, False)])
def test_batch_beats_fp64_vs_oracle(gpu, orc, name, n, s2):
    """FP64 batch: every beat record and the final states equal the oracle's trace and its
    offline beat detection on the same protocol."""
    _say("batch fp64", name)
    m, tis = MODELS[name][0], MODELS[name][1]
    ion, prots, lens, sc = _setup(gpu, name, n, "fp64", seed=5, s2=s2)
    gpu.batch_run([ion], DT_MS * 1e-3, chunk=7919)   # chunk edges fall mid-beat
    cnt, rec = ion.beats()
    states = ion.states()
    for i in range(n):
        beats, yf = _oracle_cell(orc, m, tis, prots[i], int(lens[i]), sc[:, i])
        ref = _ref_records(beats)
        _say("  cell %d: %d steps, %d beats (oracle %d)" % (i, lens[i], cnt[i], len(ref)))
        assert cnt[i] == len(ref)
        got = rec[i, :cnt[i]]
        np.testing.assert_array_equal(np.isnan(got[:, 4]), np.isnan(ref[:, 4]))
        ok = ~np.isnan(ref)
        np.testing.assert_allclose(got[ok], ref[ok], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(states[:, i], yf, rtol=1e-9, atol=1e-12)


def test_batch_mixed_models_one_graph(gpu, orc):
    """TNNP and ORd handles in one cg_batch_run (two concurrent graph branches) give the
    same records as separate runs; mixed precision APD90 within 1 ms of FP64."""
    _say("batch mixed models")
    res = {}
    for prec in ("fp64", "mixed"):
        a, _, _, _ = _setup(gpu, "TNNP_epi", 64, prec, seed=9)
        b, _, _, _ = _setup(gpu, "ORd", 64, prec, seed=10, s2=True)
        gpu.batch_run([a, b], DT_MS * 1e-3, chunk=20000)
        res[prec] = (a.beats(), b.beats())
    # the same TNNP cells alone
    a1, _, _, _ = _setup(gpu, "TNNP_epi", 64, "fp64", seed=9)
    gpu.batch_run([a1], DT_MS * 1e-3, chunk=20000)
    c1, r1 = a1.beats()
    c0, r0 = res["fp64"][0]
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(np.nan_to_num(r0, nan=-1), np.nan_to_num(r1, nan=-1))
    for k in range(2):
        (c64, r64), (c32, r32) = res["fp64"][k], res["mixed"][k]
        np.testing.assert_array_equal(c64, c32)
        apd64 = r64[:, :, 4] - r64[:, :, 0]
        apd32 = r32[:, :, 4] - r32[:, :, 0]
        both = ~np.isnan(apd64) & ~np.isnan(apd32)
        err = np.max(np.abs(apd64[both] - apd32[both])) * 1e3
        _say("  model %d: %d beats, APD90 mixed-fp64 max %.4f ms" % (k, both.sum(), err))
        assert both.sum() > 64
        assert err < 1.0


def test_batch_autotune_keeps_results(gpu, orc):
    """The autotuner restores the states it used for its trials and the chosen kernel shape
    computes the same beats as the default one."""
    _say("batch autotune")
    a, _, _, _ = _setup(gpu, "TNNP_epi", 128, "fp64", seed=3)
    before = a.states()
    rep = gpu.batch_autotune([a], DT_MS * 1e-3, trial_steps=200)
    _say(rep)
    assert "chosen" in rep
    np.testing.assert_array_equal(a.states(), before)
    gpu.batch_run([a], DT_MS * 1e-3)
    c_t, r_t = a.beats()
    b, _, _, _ = _setup(gpu, "TNNP_epi", 128, "fp64", seed=3)
    gpu.batch_run([b], DT_MS * 1e-3)
    c_d, r_d = b.beats()
    np.testing.assert_array_equal(c_t, c_d)
    ok = ~np.isnan(r_d)
    np.testing.assert_allclose(r_t[ok], r_d[ok], rtol=1e-12, atol=0)


def test_batch_mixed_crossing_zero_mv(gpu, orc):
    """Regression: a TNNP cell whose FP32 Vm passes within 1e-7 mV of 0 mV (seed 20240601,
    cell 340 of the 1,048,576-cell batch) used to hit exp(2 V F/RT) - 1 == 0 in the ICaL
    GHK quotient and turn every state to inf/NaN; it now stays finite and tracks FP64."""
    _say("batch fp32 GHK quotient at 0 mV")
    n_all = 1048576
    rng = np.random.default_rng(20240601)
    cl = rng.uniform(300.0, 1000.0, n_all)
    nb = rng.integers(5, 21, n_all)
    sc = rng.uniform(0.8, 1.2, (4, n_all))
    c = 340
    steps = 169500
    out = {}
    for prec in ("fp64", "mixed"):
        ion = gpu.IonicModel("TNNP", "epicardialCells", 1, solve_vm=True, precision=prec)
        ion.set_protocol_arrays(10.0, cl[c:c + 1], 1.0, 60.0, nb[c:c + 1].astype(np.int32))
        ion.set_run_lengths(None, steps)
        ion.set_conductance_scale(sc[:, c:c + 1])
        ion.enable_beats(max_beats=32)
        gpu.batch_run([ion], DT_MS * 1e-3)
        out[prec] = (ion.states()[:, 0], ion.beats())
    s32, s64 = out["mixed"][0], out["fp64"][0]
    assert np.all(np.isfinite(s32))
    _say("  Vm fp32 %.4f fp64 %.4f mV" % (s32[0], s64[0]))
    assert abs(s32[0] - s64[0]) < 1.0
    (c64, r64), (c32, r32) = out["fp64"][1], out["mixed"][1]
    assert c64[0] == c32[0]
    ok = ~np.isnan(r64[0, :c64[0], 4])
    assert np.max(np.abs(r64[0, :c64[0], 4][ok] - r32[0, :c32[0], 4][ok])) < 1e-3


def test_restitution_classification_vs_reference_fixture(gpu, orc):
    """On-device beats, early-stroke classification and APD90/DI (cg_ionic_get_restitution)
    of the S1-S2 cells of tests/golden/restitution_beats.json, whose expected values the
    reference's own detect_beats / classify_early_strokes / compute_apd_di produced on the
    oracle's Rush-Larsen traces: one batched run per model, every case a cell."""
    import json
    import os
    from tests.golden import cases as gc
    _say("restitution classification vs reference fixture")
    with open(os.path.join(os.path.dirname(gc.__file__), "restitution_beats.json")) as f:
        fx = [c for c in json.load(f)["cases"] if "model" in c]
    keys = {0: ("BuenoOrovio", "epicardialCells"), 1: ("TNNP", "epicardialCells"),
            2: ("ORd", "myocyte")}
    for m in sorted({c["model"] for c in fx}):
        cs = [c for c in fx if c["model"] == m]
        ion = gpu.IonicModel(*keys[m], len(cs), solve_vm=True, precision="fp64")
        ion.set_protocols([{"stim_start": c["stimulus"]["start"],
                            "stim_period_S1": c["stimulus"]["period_s1"],
                            "stim_duration": c["stimulus"]["duration"],
                            "stim_amplitude": c["stimulus"]["amplitude"],
                            "nstim1": c["stimulus"]["nstim1"],
                            "stim_period_S2": c["stimulus"]["period_s2"],
                            "nstim2": c["stimulus"]["nstim2"]} for c in cs])
        ion.set_run_lengths(np.array([c["n_steps"] for c in cs], dtype=np.int64))
        ion.enable_beats(max_beats=16)
        gpu.batch_run([ion], cs[0]["dt_ms"] * 1e-3)
        cnt, rec = ion.beats()
        apd, di, flags = ion.restitution()
        for i, c in enumerate(cs):
            want = c["beats"]
            assert cnt[i] == len(want), (c["name"], cnt[i], len(want))
            for k, w in enumerate(want):
                assert abs(rec[i, k, 0] - w["t_up"]) <= 1e-9, (c["name"], k)
                if w["t_repol90"] is None:
                    assert np.isnan(rec[i, k, 4])
                else:
                    assert abs(rec[i, k, 4] - w["t_repol90"]) <= 1e-9, (c["name"], k)
                fw = (1 if w["early"] else 0) | (2 if w["valid_apd"] else 0) | \
                    (4 if w["valid_di"] else 0)
                assert flags[i, k] == fw, (c["name"], k, flags[i, k], fw)
            for got, ref in ((apd[i], c["apd"]), (di[i], c["di"])):
                if ref is None:
                    assert np.isnan(got), (c["name"], got)
                else:
                    assert abs(got - ref) <= 1e-9, (c["name"], got, ref)
            _say("  %s: %d beats, early %s, APD90 %s DI %s" % (c["name"], cnt[i], c["early"],
                                                             apd[i], di[i]))
