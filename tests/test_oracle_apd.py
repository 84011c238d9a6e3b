"""APD90 detection of the oracle on traces with closed-form answers
(postProcessing_restCurves.py:115-200 semantics)."""
import numpy as np
import pytest


def _trace(dt=1e-5, beats=((0.1, 0.25), (0.6, 0.2)), base=-85.0, peak=35.0, tend=1.0):
    """Piecewise-linear APs: 1 ms linear upstroke to peak, linear repolarisation
    back to base over `dur` seconds."""
    t = np.arange(0.0, tend, dt)
    v = np.full_like(t, base)
    for t_on, dur in beats:
        up = (t >= t_on) & (t < t_on + 1e-3)
        v_start = v[np.searchsorted(t, t_on) - 1]  # upstroke starts from the current V
        v[up] = v_start + (peak - v_start) * (t[up] - t_on) / 1e-3
        down = (t >= t_on + 1e-3) & (t < t_on + 1e-3 + dur)
        v[down] = peak - (peak - base) * (t[down] - t_on - 1e-3) / dur
    return t, v


def test_apd90_closed_form(orc):
    base, peak = -85.0, 35.0
    beats = ((0.1, 0.25), (0.6, 0.2))
    t, v = _trace(beats=beats, base=base, peak=peak)
    out = orc.detect_beats(t, v)
    assert len(out) == 2
    for b, (t_on, dur) in zip(out, beats):
        # dV/dt threshold 5000 mV/s crossed one sample into the upstroke (centred
        # differences straddle the kink): t_up is within one sample of t_on
        assert abs(b["t_up"] - t_on) <= 1e-5 + 1e-12
        assert b["v_base"] == pytest.approx(base)
        assert b["v_peak"] == pytest.approx(peak, abs=(peak - base) * 1e-5 / dur * 2)
        v90 = base + 0.1 * (b["v_peak"] - base)
        t_rep = t_on + 1e-3 + (b["v_peak"] - v90) / (peak - base) * dur
        assert b["t_repol90"] == pytest.approx(t_rep, abs=2e-5)


def test_interrupted_beat_has_no_repolarisation(orc):
    # second upstroke arrives before the first beat repolarises to V90
    t, v = _trace(beats=((0.1, 0.3), (0.3, 0.2)))
    out = orc.detect_beats(t, v)
    assert len(out) == 2
    assert not out[0]["repol_found"]
    assert out[1]["repol_found"]
