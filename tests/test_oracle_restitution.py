"""The oracle's beat detection, early-stroke classification and APD90/DI against the
reference's own post-processing (postProcessing_restCurves.py: detect_beats,
classify_early_strokes, compute_apd_di), whose outputs on seeded S1-S2 traces and on
piecewise-linear traces are the fixture tests/golden/restitution_beats.json."""
import json
import os

import numpy as np
import pytest

from tests.golden.make_restitution_fixtures import synthetic_trace

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                   "restitution_beats.json")


def _cases():
    with open(FIX) as f:
        return json.load(f)["cases"]


def trace_of(orc, c):
    """the case's input trace (time s, Vm mV), checked against the stored checksum"""
    if "knots_t" in c:
        return synthetic_trace(c["knots_t"], c["knots_v"], c["t_end"])
    st = orc.make_stim(**c["stimulus"])
    t, v, _ = orc.single_cell_trace(c["model"], c["tissue"], st, c["dt_ms"], c["n_steps"])
    assert float(np.sum(v)) == c["checksum"]["sum"] and float(v[-1]) == c["checksum"]["last"]
    return t * 1e-3, v


def same(a, b, tol=1e-12):
    if b is None:
        return a is None or not np.isfinite(a)
    return a is not None and abs(a - b) <= tol * max(1.0, abs(b))


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["name"])
def test_beats_classification_apd_di(orc, case):
    t, v = trace_of(orc, case)
    beats = orc.detect_beats(t, v)
    want = case["beats"]
    assert len(beats) == len(want)
    for b, w in zip(beats, want):
        for k in ("t_up", "v_base", "v_peak", "v90", "t_repol90"):
            assert same(b[k], w[k]), (k, b[k], w[k])
        assert b["repol_found"] == w["repol_found"]
    assert orc.classify_beats(beats) == case["early"]
    for b, w in zip(beats, want):
        assert (b["early"], b["valid_apd"], b["valid_di"]) == \
            (w["early"], w["valid_apd"], w["valid_di"])
    apd, di = orc.apd_di(beats)
    assert same(apd, case["apd"]) and same(di, case["di"]), (apd, di, case["apd"], case["di"])


def test_fixture_covers_every_branch():
    """interrupted repolarisation, next upstroke before APD90, and accepted pairs"""
    cases = _cases()
    assert any(c["early"] for c in cases) and any(c["apd"] for c in cases)
    assert any(not b["repol_found"] for c in cases for b in c["beats"])
    assert any(c["apd"] is None and c["early"] for c in cases)
