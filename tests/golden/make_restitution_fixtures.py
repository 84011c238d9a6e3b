"""Writes tests/golden/restitution_beats.json: the reference's own beat detection,
early-stroke classification and APD90/DI (detect_beats, classify_early_strokes and
compute_apd_di of tutorials/singleCellprotocols/restitutionCurves_s1s2Protocol/
setupRestitutionCurves_s1s2Protocol/postProcessing_restCurves.py) on seeded inputs:
Rush-Larsen single-cell S1-S2 traces of the oracle (inputs named by model, protocol and
step, with a checksum of the trace) and piecewise-linear traces given point by point.

The reference module is imported from the reference tree in the build container only;
run there:  python tests/golden/make_restitution_fixtures.py /root/reference
"""
import contextlib
import importlib.util
import io
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SCRIPT = ("tutorials/singleCellprotocols/restitutionCurves_s1s2Protocol/"
          "setupRestitutionCurves_s1s2Protocol/postProcessing_restCurves.py")

DT_MS = 0.01
# (case key, oracle model, tissue, S1 period ms, n S1, S2 interval ms): the S2 train's first
# pulse falls one S1 period after the last S1 (stimulusIO.C:268-283), the second S2 ms later
TRACES = [("BO_epi", 0, 1, 400.0, 3, s2) for s2 in (170.0, 230.0, 260.0, 300.0, 500.0)] + \
         [("TNNP_epi", 1, 1, 500.0, 2, s2) for s2 in (250.0, 290.0, 330.0, 450.0)] + \
         [("ORd", 2, 4, 500.0, 2, s2) for s2 in (240.0, 290.0, 400.0)]
AMP = {0: 0.4, 1: 52.0, 2: 80.0}


def synthetic_trace(knots_t, knots_v, t_end, dt=1e-4):
    """(time s, Vm mV): the piecewise-linear trace through the knots, sampled every dt"""
    t = np.round(np.arange(0.0, t_end + 1e-9, dt), 10)
    return t, np.interp(t, np.asarray(knots_t), np.asarray(knots_v))


def synthetic():
    """two beats where the second upstroke comes before, just after, and long after the
    first one's 90 % repolarisation: (name, knots_t, knots_v, t_end)"""
    out = []
    for name, t2 in (("interrupted", 0.150), ("just_after", 0.242), ("normal", 0.400)):
        first_t = [0.0, 0.010, 0.012, 0.150, 0.240]
        first_v = [-85.0, -85.0, 30.0, 10.0, -85.0]
        v1 = float(np.interp(t2, first_t, first_v))
        keep = [i for i in range(5) if first_t[i] <= t2]
        kt = [first_t[i] for i in keep] + [t2 + 1e-9, t2 + 0.002, t2 + 0.140, t2 + 0.230]
        kv = [first_v[i] for i in keep] + [v1, 30.0, 10.0, -85.0]
        out.append((name, kt, kv, t2 + 0.5))
    return out


def _clean(x):
    if isinstance(x, (bool, np.bool_)):
        return bool(x)
    if isinstance(x, (float, np.floating)):
        return None if not math.isfinite(float(x)) else float(x)
    return x


def main(ref):
    spec = importlib.util.spec_from_file_location("rest", os.path.join(ref, SCRIPT))
    import matplotlib
    matplotlib.use("Agg")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    import oracle as orc

    def analyse(t_s, v_mv):
        with contextlib.redirect_stdout(io.StringIO()):
            beats = mod.detect_beats(t_s, v_mv)
            early = mod.classify_early_strokes(beats)
            apd, di = mod.compute_apd_di(beats)
        return ([{k: _clean(v) for k, v in b.items()} for b in beats],
                [beats.index(b) for b in early], _clean(apd), _clean(di))

    cases = []
    for key, m, tis, s1, n1, s2 in TRACES:
        stim = dict(start=5.0025, period_s1=s1, duration=1.0, amplitude=AMP[m], nstim1=n1,
                    period_s2=s2, nstim2=2)
        end = stim["start"] + s1 * n1 + s2 + 600.0
        n_steps = int(end / DT_MS)
        t, v, _ = orc.single_cell_trace(m, tis, orc.make_stim(**stim), DT_MS, n_steps)
        beats, early, apd, di = analyse(t * 1e-3, v)
        cases.append({"name": "%s_S2_%g" % (key, s2), "model": m, "tissue": tis,
                      "stimulus": stim, "dt_ms": DT_MS, "n_steps": n_steps,
                      "checksum": {"sum": float(np.sum(v)), "last": float(v[-1])},
                      "beats": beats, "early": early, "apd": apd, "di": di})
    for name, kt, kv, t_end in synthetic():
        t, v = synthetic_trace(kt, kv, t_end)
        beats, early, apd, di = analyse(t, v)
        cases.append({"name": "synthetic_" + name, "knots_t": kt, "knots_v": kv,
                      "t_end": t_end, "beats": beats, "early": early, "apd": apd, "di": di})
    with open(os.path.join(HERE, "restitution_beats.json"), "w") as f:
        json.dump({"source": SCRIPT, "thresholds": {"dvdt": mod.DVDT_THRESHOLD,
                                                    "baseline": mod.BASELINE,
                                                    "peak_search": mod.PEAK_SEARCH},
                   "cases": cases}, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
