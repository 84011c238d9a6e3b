/*
 * Oracle (test infrastructure): beat detection and APD90 of a single-cell trace,
 * restating detect_beats (the second definition, which shadows the first) of
 * tutorials/singleCellprotocols/restitutionCurves_s1s2Protocol/
 * setupRestitutionCurves_s1s2Protocol/postProcessing_restCurves.py:115-200:
 *   - dV/dt by centred differences, one-sided copies at both ends
 *   - upstroke where dV/dt crosses the threshold upward (linear interpolation)
 *   - baseline = min V over [t_up - baseline, t_up)
 *   - peak = max V over [t_up, min(t_up + peak_window, next t_up)] (first max)
 *   - V90 = base + 0.1 (peak - base); t_repol90 = first downward crossing of V90
 *     after the peak (linear interpolation), abandoned once the next upstroke passes.
 * Pinned by closed-form piecewise-linear traces in tests/test_oracle_apd.py.
 */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

int orc_detect_beats(const double* t, const double* vm, long n, orc_beat_t* beats,
                     int max_beats, double thr, double baseline_s, double peak_s)
{
    if (n < 3) return 0;
    double* dv = malloc(sizeof(double) * n);
    for (long i = 1; i < n - 1; ++i) dv[i] = (vm[i + 1] - vm[i - 1]) / (t[i + 1] - t[i - 1]);
    dv[0] = dv[1];
    dv[n - 1] = dv[n - 2];
    /* candidates */
    long cap = 64, nc = 0;
    double* cand = malloc(sizeof(double) * cap);
    for (long i = 1; i < n; ++i) {
        if (dv[i - 1] < thr && thr <= dv[i]) {
            double frac = (thr - dv[i - 1]) / (dv[i] - dv[i - 1]);
            if (frac >= 0.0 && frac <= 1.0) {
                if (nc == cap) { cap *= 2; cand = realloc(cand, sizeof(double) * cap); }
                cand[nc++] = t[i - 1] + frac * (t[i] - t[i - 1]);
            }
        }
    }
    int nb = 0;
    for (long k = 0; k < nc && nb < max_beats; ++k) {
        const double tu = cand[k];
        double vbase = INFINITY;
        int any = 0;
        for (long i = 0; i < n; ++i)
            if (t[i] >= tu - baseline_s && t[i] < tu) { if (vm[i] < vbase) vbase = vm[i]; any = 1; }
        if (!any) continue;
        double tend = tu + peak_s;
        if (k < nc - 1 && cand[k + 1] < tend) tend = cand[k + 1];
        long first = -1, last = -1;
        for (long i = 0; i < n; ++i)
            if (t[i] >= tu && t[i] <= tend) { if (first < 0) first = i; last = i; }
        if (first < 0) continue;
        double vpeak = -INFINITY;
        long jp = first;
        for (long i = first; i <= last; ++i)
            if (vm[i] > vpeak) { vpeak = vm[i]; jp = i; }
        double v90 = vbase + 0.1 * (vpeak - vbase);
        double trep = NAN;
        for (long j = jp + 1; j <= last; ++j) {
            if (vm[j - 1] > v90 && v90 >= vm[j]) {
                double frac = (v90 - vm[j - 1]) / (vm[j] - vm[j - 1]);
                trep = t[j - 1] + frac * (t[j] - t[j - 1]);
                break;
            }
            if (k < nc - 1 && t[j] >= cand[k + 1]) break;
        }
        orc_beat_t* b = &beats[nb++];
        b->t_up = tu; b->v_base = vbase; b->v_peak = vpeak; b->v90 = v90;
        b->t_repol90 = trep; b->repol_found = !isnan(trep);
    }
    free(dv);
    free(cand);
    return nb;
}
