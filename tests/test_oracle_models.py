"""Pins of the oracle's ionic models beyond the BuenoOrovio goldens.

TNNP and ORd have no golden in the reference.  They are pinned by properties the
published models fix and a transcription slip would break:
  - charge conservation: the concentration balances (with rapid buffering) carry
    exactly the charge of the summed ionic current, F sum z V dc_total = -Cm Iion;
  - published action-potential shape at 1 Hz (resting potential, peak, APD90);
  - Rush-Larsen exactness for a gate with constant (xinf, tau).
Courtemanche: the same charge balance, the published action potential from the
published initial state, and the limits its generated code substitutes at the removable
singularities of the rate expressions.
PerisYague_2022: the charge balance with chloride, its guard values and its unguarded
quotients at their singular voltages.  Its gate kinetics have no published trace to pin
against (the reference's h-gate beta differs from Courtemanche's): parity unpinned there
beyond the device-vs-oracle agreement.
Grandi: the Na+/Ca2+ charge balance over its three compartments and their buffers (K_i
and Cl_i are fixed in that model), and first-order convergence of the Rush-Larsen split
to the adaptive RKF45 on one paced beat.
ToRORd_dynCl: the charge balance of Na, K, Cl and Ca with its rapid buffers (three tissues),
and the Rush-Larsen split against RKF45 on the published action potential.
AlievPanfilov: the cubic's roots 0, a, 1 and the recovery nullcline.
Stimulus: stimulusIO.C:238-286 cases, including the S2 placement at the end of S1.
"""
import numpy as np
import pytest


def _states_along_ap(orc, model, tissue, amp, n=12):
    st = orc.make_stim(start=5.0, period_s1=1000.0, duration=1.0, amplitude=amp, nstim1=1)
    out = []
    c, y = orc.model_init(model, tissue)
    dt = 0.01
    for k in range(int(400 / dt)):
        yy = _pad(y)
        orc.lib().orc_rl_step(model, tissue, k * dt, dt, c, yy, 1, st)
        y = yy[:len(y)]
        if k % int(400 / dt / n) == 0:
            out.append(y.copy())
    return c, out


def orc_max_states():
    import oracle
    return oracle.MAX_STATES


def _pad(y):
    a = np.zeros(orc_max_states())
    a[:len(y)] = y
    return a


def test_tnnp_charge_conservation(orc):
    # TNNP_2004.H constants: F, V_c, V_sr, Cm, Buf_c, K_buf_c, Buf_sr, K_buf_sr
    F, Vc, Vsr, Cm = 96.485, 16404.0, 1094.0, 185.0
    Bufc, Kbufc, Bufsr, Kbufsr = 0.15, 0.001, 10.0, 0.3
    for tissue in (orc.EPI, orc.MCELL, orc.ENDO):
        c, states = _states_along_ap(orc, orc.TNNP, tissue, 52.0)
        st = orc.make_stim(start=0.0, period_s1=1000.0, duration=2.0, amplitude=30.0, nstim1=1)
        for y in states:
            for t, stim in ((500.0, None), (1.0, st)):
                r = orc.model_eval(orc.TNNP, tissue, t, c, y, solve_vm=True, stim=stim)
                d = r["rates"]
                Nai, Ki, Cai, CaSR = 2, 1, 3, 15
                dCai_t = d[Cai] * (1 + Bufc * Kbufc / (y[Cai] + Kbufc) ** 2)
                dCaSR_t = d[CaSR] * (1 + Bufsr * Kbufsr / (y[CaSR] + Kbufsr) ** 2)
                lhs = F * (Vc * (d[Nai] + d[Ki] + 2 * dCai_t) + 2 * Vsr * dCaSR_t)
                rhs = -Cm * (r["iion"] + r["istim"])
                assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-9 * Cm)
                # dV/dt carries the same current
                assert d[0] == pytest.approx(-(r["iion"] + r["istim"]), rel=1e-14)


def test_ord_charge_conservation(orc):
    # ORd_2011.H:497-692 geometry and buffer constants
    L, rad = 0.01, 0.0011
    Ageo = 2 * 3.14 * rad * rad + 2 * 3.14 * rad * L
    vcell = 1000 * 3.14 * rad * rad * L
    Acap, vmyo, vss, vnsr, vjsr = 2 * Ageo, 0.68 * vcell, 0.02 * vcell, 0.0552 * vcell, 0.0048 * vcell
    F = 96485.0
    cmdnmax, kmcmdn, trpnmax, kmtrpn = 0.05, 0.00238, 0.07, 0.0005
    BSRmax, KmBSR, BSLmax, KmBSL = 0.047, 0.00087, 1.124, 0.0087
    csqnmax, kmcsqn = 10.0, 0.8
    c, states = _states_along_ap(orc, orc.ORD, orc.MYOCYTE, 80.0)
    st = orc.make_stim(start=0.0, period_s1=1000.0, duration=2.0, amplitude=40.0, nstim1=1)
    I = dict(V=0, Nai=2, Nass=3, Ki=4, Kss=5, Cass=6, Cansr=7, Cajsr=8, Cai=9)
    for y in states:
        for t, stim in ((500.0, None), (1.0, st)):
            r = orc.model_eval(orc.ORD, orc.MYOCYTE, t, c, y, solve_vm=True, stim=stim)
            d = r["rates"]
            Cai, Cass, Cajsr = y[I["Cai"]], y[I["Cass"]], y[I["Cajsr"]]
            dCai_t = d[I["Cai"]] * (1 + cmdnmax * kmcmdn / (kmcmdn + Cai) ** 2 + trpnmax * kmtrpn / (kmtrpn + Cai) ** 2)
            dCass_t = d[I["Cass"]] * (1 + BSRmax * KmBSR / (KmBSR + Cass) ** 2 + BSLmax * KmBSL / (KmBSL + Cass) ** 2)
            dCajsr_t = d[I["Cajsr"]] * (1 + csqnmax * kmcsqn / (kmcsqn + Cajsr) ** 2)
            lhs = F * (vmyo * (d[I["Nai"]] + d[I["Ki"]]) + vss * (d[I["Nass"]] + d[I["Kss"]]) +
                       2 * (vmyo * dCai_t + vss * dCass_t + vnsr * d[I["Cansr"]] + vjsr * dCajsr_t))
            # the K_i balance omits the stimulus (ORd_2011.H:962 reads Istim before :982)
            rhs = -Acap * r["iion"]
            assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-12)
            assert d[0] == pytest.approx(-(r["iion"] + r["istim"]), rel=1e-14)


@pytest.mark.parametrize("model,tissue,amp,rest,peak,apd", [
    # published ranges at 1 Hz pacing after 5 beats
    ("TNNP", "EPI", 52.0, (-87.5, -85.0), (25.0, 45.0), (250.0, 310.0)),
    ("TNNP", "ENDO", 52.0, (-87.5, -85.0), (25.0, 45.0), (250.0, 310.0)),
    ("TNNP", "MCELL", 52.0, (-87.5, -85.0), (25.0, 45.0), (300.0, 420.0)),
    ("ORD", "MYOCYTE", 80.0, (-89.0, -87.0), (35.0, 55.0), (240.0, 290.0)),
])
def test_published_action_potential(orc, model, tissue, amp, rest, peak, apd):
    m, t_ = getattr(orc, model), getattr(orc, tissue)
    st = orc.make_stim(start=10.0, period_s1=1000.0, duration=1.0, amplitude=amp, nstim1=5)
    t, v, _ = orc.single_cell_run(m, t_, st, 1e-5, 500000, integrator=0)
    beats = orc.detect_beats(t, v)
    assert len(beats) == 5
    b = beats[-1]
    assert rest[0] < b["v_base"] < rest[1]
    assert peak[0] < b["v_peak"] < peak[1]
    assert apd[0] < (b["t_repol90"] - b["t_up"]) * 1e3 < apd[1]
    # RL and adaptive RKF45 agree on the last beat's APD90 within 1 ms
    t2, v2, _ = orc.single_cell_run(m, t_, st, 1e-5, 500000, integrator=1)
    b2 = orc.detect_beats(t2, v2)[-1]
    assert abs((b2["t_repol90"] - b2["t_up"]) - (b["t_repol90"] - b["t_up"])) < 1e-3


def test_rush_larsen_exact_for_frozen_voltage(orc):
    """BuenoOrovio gates have (xinf, tau) depending on u only: with u frozen (tissue
    mode) one RL step of dt equals the exact solution, i.e. two RL steps of dt/2."""
    c, y0 = orc.model_init(orc.BO, orc.EPI)
    rng = np.random.default_rng(0)
    for _ in range(20):
        y = np.array([rng.uniform(-0.1, 1.4), rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1)])
        a = _pad(y)
        orc.lib().orc_rl_step(orc.BO, orc.EPI, 0.0, 0.5, c, a, 0, None)
        b = _pad(y)
        orc.lib().orc_rl_step(orc.BO, orc.EPI, 0.0, 0.25, c, b, 0, None)
        orc.lib().orc_rl_step(orc.BO, orc.EPI, 0.25, 0.25, c, b, 0, None)
        np.testing.assert_allclose(a[:4], b[:4], rtol=1e-13, atol=1e-15)


def test_aliev_panfilov_cubic(orc):
    c, y = orc.model_init(orc.AP, orc.MYOCYTE)
    a = 0.1  # (Vth - Vr)/(Vp - Vr) = 10/100
    for u in (0.0, a, 1.0):
        r = orc.model_eval(orc.AP, orc.MYOCYTE, 0.0, c, [u, 0.0], solve_vm=True)
        assert abs(r["iion"]) < 1e-15
    # recovery nullcline: dr/dt = 0 on r = -k u (u - a - 1)
    for u in (0.2, 0.5, 0.9):
        rr = -8.0 * u * (u - a - 1.0)
        r = orc.model_eval(orc.AP, orc.MYOCYTE, 0.0, c, [u, rr], solve_vm=True)
        assert abs(r["rates"][1]) < 1e-12
    r = orc.model_eval(orc.AP, orc.MYOCYTE, 0.0, c, [0.05, 0.0], solve_vm=True)
    assert r["iion"] == pytest.approx(8 * 0.05 * (0.05 - a) * (0.05 - 1.0), rel=1e-14)


def test_tissue_support_matches_reference(orc):
    # BuenoOrovio.C:92-96, TNNP.C:90-94, ORd.C:93-96, AlievPanfilov.C:92-95,
    # Courtemanche.C:92-95, PerisYague.C:92-95, Stewart.C:93-96, Grandi.C:93-96,
    # ToRORd_dynCl.C:93-96
    for tissue in (1, 2, 3):
        orc.model_init(orc.BO, tissue)
        orc.model_init(orc.TNNP, tissue)
        with pytest.raises(ValueError):
            orc.model_init(orc.ORD, tissue)
        with pytest.raises(ValueError):
            orc.model_init(orc.CRN, tissue)
        with pytest.raises(ValueError):
            orc.model_init(orc.PYS, tissue)
        with pytest.raises(ValueError):
            orc.model_init(orc.STW, tissue)
        with pytest.raises(ValueError):
            orc.model_init(orc.GRD, tissue)
        orc.model_init(orc.TOR, tissue)
    orc.model_init(orc.ORD, 4)
    orc.model_init(orc.CRN, 4)
    orc.model_init(orc.PYS, 4)
    orc.model_init(orc.STW, 4)
    orc.model_init(orc.GRD, 4)
    with pytest.raises(ValueError):
        orc.model_init(orc.TOR, 4)
    orc.model_init(orc.AP, 4)
    with pytest.raises(ValueError):
        orc.model_init(orc.BO, 4)


@pytest.mark.parametrize("t,expected", [
    (19.999, 0.0), (20.0, -0.4), (20.5, -0.4), (21.0, -0.4), (21.0001, 0.0),
    (1020.0, -0.4), (2020.5, -0.4), (3020.0, -0.4),   # tp == S1*n1 is inside (<=)
    (3020.5, 0.0),
])
def test_stimulus_s1(orc, t, expected):
    st = orc.make_stim(start=20.0, period_s1=1000.0, duration=1.0, amplitude=0.4, nstim1=3)
    assert orc.lib().orc_stimulus(t, st) == expected


def test_stimulus_s2_placement(orc):
    # first S2 pulse at tS1End = start + S1*n1 (already covered by S1's closed end),
    # then S2 later; the tutorials therefore use nstim2 = 2
    st = orc.make_stim(start=10.0, period_s1=1000.0, duration=1.0, amplitude=1.0, nstim1=2,
                       period_s2=300.0, nstim2=2)
    f = lambda t: orc.lib().orc_stimulus(t, st)
    assert f(2010.5) == -1.0          # S2 pulse at tS1End (=2010)
    assert f(2310.5) == -1.0          # next S2 pulse
    assert f(2610.0) == -1.0          # tp == S2*n2 is still inside (<=)
    assert f(2610.5) == 0.0           # tp > S2*n2: the train is over
    assert f(2911.0) == 0.0
    assert f(2200.0) == 0.0
    # defaults: nstim1 = 1 when a positive S1 period is given (stimulusIO.C:179-183)
    assert orc.make_stim(period_s1=10.0).nstim1 == 1
    assert orc.make_stim().nstim1 == 0


def test_crn_charge_conservation(orc):
    # Courtemanche_1998.H:341-431 geometry and buffer constants
    F, Cm, Vcell = 96.4867, 100.0, 20100.0
    Vi, Vup, Vrel = 0.68 * Vcell, 0.0552 * Vcell, 0.0048 * Vcell
    cmdn, kcmdn, trpn, ktrpn, csqn, kcsqn = 0.05, 0.00238, 0.07, 0.0005, 10.0, 0.8
    c, states = _states_along_ap(orc, orc.CRN, orc.MYOCYTE, 40.0)
    st = orc.make_stim(start=0.0, period_s1=1000.0, duration=2.0, amplitude=40.0, nstim1=1)
    Nai, Ki, Cai, CaUp, CaRel = 1, 2, 3, 4, 5
    for y in states:
        for t, stim in ((500.0, None), (1.0, st)):
            r = orc.model_eval(orc.CRN, orc.MYOCYTE, t, c, y, solve_vm=True, stim=stim)
            d = r["rates"]
            dCai_t = d[Cai] * (1 + trpn * ktrpn / (y[Cai] + ktrpn) ** 2 + cmdn * kcmdn / (y[Cai] + kcmdn) ** 2)
            dCaRel_t = d[CaRel] * (1 + csqn * kcsqn / (y[CaRel] + kcsqn) ** 2)
            lhs = F * (Vi * (d[Nai] + d[Ki] + 2 * dCai_t) + 2 * Vup * d[CaUp] + 2 * Vrel * dCaRel_t)
            # the K_i balance omits the stimulus (Courtemanche_1998.H:644 reads Istim before :652)
            rhs = -Cm * r["iion"]
            assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-9 * Cm)
            assert d[0] == pytest.approx(-(r["iion"] + r["istim"]), rel=1e-14)


def test_crn_published_action_potential(orc):
    """From the initial state of Courtemanche et al. (Am J Physiol 1998, Table 4) and a
    2 ms, 2 nA (20 pA/pF) stimulus at 1 Hz the paper's action potential comes back:
    resting potential -81 mV, overshoot +25 mV, APD90 about 300 ms."""
    y0 = [-81.18, 11.17, 139.0, 1.013e-4, 1.488, 1.488, 2.908e-3, 0.9649, 0.9775, 3.043e-2,
          0.9992, 4.966e-3, 0.9986, 3.296e-5, 1.869e-2, 1.367e-4, 0.9996, 0.7755, 0.0, 1.0,
          0.9992]
    st = orc.make_stim(start=10.0, period_s1=1000.0, duration=2.0, amplitude=20.0, nstim1=2)
    t, v, _ = orc.single_cell_run(orc.CRN, orc.MYOCYTE, st, 1e-5, 200000, integrator=0, y0=y0)
    beats = orc.detect_beats(t, v)
    assert len(beats) == 2
    for b in beats:
        assert -82.0 < b["v_base"] < -80.0
        assert 22.0 < b["v_peak"] < 28.0
        assert 290.0 < (b["t_repol90"] - b["t_up"]) * 1e3 < 315.0
    # the reference's own initial state (higher Na_i) shortens APD90; RL and RKF45 agree
    t1, v1, _ = orc.single_cell_run(orc.CRN, orc.MYOCYTE, st, 1e-5, 200000, integrator=0)
    t2, v2, _ = orc.single_cell_run(orc.CRN, orc.MYOCYTE, st, 1e-5, 200000, integrator=1)
    b1, b2 = orc.detect_beats(t1, v1)[-1], orc.detect_beats(t2, v2)[-1]
    assert 200.0 < (b1["t_repol90"] - b1["t_up"]) * 1e3 < 280.0
    assert abs((b2["t_repol90"] - b2["t_up"]) - (b1["t_repol90"] - b1["t_up"])) < 1e-3


@pytest.mark.parametrize("vs,gate", [(-47.13, 6), (-10.0, 15), (7.9, 20), (-14.1, 13),
                                     (3.3328, 13), (19.9, 14)])
def test_crn_removable_singularities(orc, vs, gate):
    """At each removable singularity of a rate expression the generated code substitutes
    a limit; the gate's rate there must join its values at V +- 1e-4 mV."""
    c, y = orc.model_init(orc.CRN, orc.MYOCYTE)
    rates = []
    for dv in (-1e-4, 0.0, 1e-4):
        yy = np.array(y)
        yy[0] = vs + dv
        yy[gate] = 0.5
        rates.append(orc.model_eval(orc.CRN, orc.MYOCYTE, 0.0, c, yy, solve_vm=False)["rates"][gate])
    mid = 0.5 * (rates[0] + rates[2])
    assert rates[1] == pytest.approx(mid, rel=1e-5, abs=1e-12)


def test_pys_charge_conservation(orc):
    # PerisYague_2022.H:214-263 volumes and buffer constants; chloride carries charge -1
    F, Cm, Vi, Vup, Vrel = 96.4867, 100.0, 13668.0, 1109.52, 96.48
    cmdn, kcmdn, trpn, ktrpn, csqn, kcsqn = 0.05, 0.00238, 0.07, 0.0005, 10.0, 0.8
    c, states = _states_along_ap(orc, orc.PYS, orc.MYOCYTE, 40.0)
    st = orc.make_stim(start=0.0, period_s1=1000.0, duration=2.0, amplitude=40.0, nstim1=1)
    Nai, Ki, Cli, Cai, CaUp, CaRel = 1, 2, 3, 4, 5, 6
    for y in states:
        for t, stim in ((500.0, None), (1.0, st)):
            r = orc.model_eval(orc.PYS, orc.MYOCYTE, t, c, y, solve_vm=True, stim=stim)
            d = r["rates"]
            dCai_t = d[Cai] * (1 + trpn * ktrpn / (y[Cai] + ktrpn) ** 2 + cmdn * kcmdn / (y[Cai] + kcmdn) ** 2)
            dCaRel_t = d[CaRel] * (1 + csqn * kcsqn / (y[CaRel] + kcsqn) ** 2)
            lhs = F * (Vi * (d[Nai] + d[Ki] - d[Cli] + 2 * dCai_t) + 2 * Vup * d[CaUp] + 2 * Vrel * dCaRel_t)
            rhs = -Cm * r["iion"]
            assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-9 * Cm)
            assert d[0] == pytest.approx(-(r["iion"] + r["istim"]), rel=1e-14)


@pytest.mark.parametrize("vs,gate,guarded", [(-47.13, 7, True), (7.9, 21, True),
                                             (-14.1, 13, False), (3.3328, 13, False),
                                             (18.80816, 14, False)])
def test_pys_singularities(orc, vs, gate, guarded):
    """Guarded singularities (m, w) return the expression's limit; the unguarded IKr/IKs
    quotients are 0/0 exactly at the singular voltage (PerisYague_2022.H:348-357) and join
    their neighbours on both sides."""
    c, y = orc.model_init(orc.PYS, orc.MYOCYTE)

    def rate(v):
        yy = np.array(y)
        yy[0], yy[gate] = v, 0.5
        return orc.model_eval(orc.PYS, orc.MYOCYTE, 0.0, c, yy, solve_vm=False)["rates"][gate]

    lo, hi = rate(vs - 1e-4), rate(vs + 1e-4)
    if guarded:
        assert rate(vs) == pytest.approx(0.5 * (lo + hi), rel=1e-5, abs=1e-12)
    else:
        assert np.isnan(rate(vs))
        # linear extrapolations to vs from either side agree: the quotient has a finite,
        # continuous limit there
        left, right = 2 * lo - rate(vs - 2e-4), 2 * hi - rate(vs + 2e-4)
        assert left == pytest.approx(right, rel=1e-6, abs=1e-12)


def test_pys_d_gate_guard_value(orc):
    """The d-gate time constant's guard at V = -5 mV returns 1.0 (PerisYague_2022.H:368),
    while the expression tends to 1/(6.24 * 2 * 0.035) on both sides."""
    c, y = orc.model_init(orc.PYS, orc.MYOCYTE)
    dinf = 1.0 / (1.0 + np.exp(0.0))
    for v, tau in ((-5.0, 1.0), (-5.0 + 1e-5, 1.0 / (6.24 * 2 * 0.035)),
                   (-5.0 - 1e-5, 1.0 / (6.24 * 2 * 0.035))):
        yy = np.array(y)
        yy[0], yy[15] = v, 0.25
        r = orc.model_eval(orc.PYS, orc.MYOCYTE, 0.0, c, yy, solve_vm=False)["rates"][15]
        assert r == pytest.approx((dinf - 0.25) / tau, rel=1e-4)


def test_stw_charge_conservation(orc):
    # Stewart_2009.H:264-358 volumes (uL), capacitance (uF) and buffer constants
    F, Cm, Vc, Vsr, Vss = 96485.3415, 0.185, 0.016404, 0.001094, 5.468e-05
    Bufc, Kbufc, Bufsr, Kbufsr, Bufss, Kbufss = 0.2, 0.001, 10.0, 0.3, 0.4, 0.00025
    c, states = _states_along_ap(orc, orc.STW, orc.MYOCYTE, 40.0)
    st = orc.make_stim(start=0.0, period_s1=1000.0, duration=2.0, amplitude=40.0, nstim1=1)
    Cai, CaSR, Cass, Nai, Ki = 14, 15, 16, 18, 19
    for y in states:
        for t, stim in ((500.0, None), (1.0, st)):
            r = orc.model_eval(orc.STW, orc.MYOCYTE, t, c, y, solve_vm=True, stim=stim)
            d = r["rates"]
            dCai_t = d[Cai] * (1 + Bufc * Kbufc / (y[Cai] + Kbufc) ** 2)
            dCaSR_t = d[CaSR] * (1 + Bufsr * Kbufsr / (y[CaSR] + Kbufsr) ** 2)
            dCass_t = d[Cass] * (1 + Bufss * Kbufss / (y[Cass] + Kbufss) ** 2)
            lhs = F * (Vc * (d[Nai] + d[Ki] + 2 * dCai_t) + 2 * Vsr * dCaSR_t + 2 * Vss * dCass_t)
            rhs = -Cm * r["iion"]
            assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-9 * Cm)
            assert d[0] == pytest.approx(-(r["iion"] + r["istim"]), rel=1e-14)


def test_stw_limit_cycle_returns_to_reference_state(orc):
    """The Purkinje cell fires on its own.  The reference's initial state (given to 17
    digits, Stewart_2009.H:358-377) is a point of that limit cycle: one cycle of the
    unstimulated cell brings all 20 states back to it."""
    import ctypes as C
    c, y0 = orc.model_init(orc.STW, orc.MYOCYTE)
    st = orc.make_stim(start=0.0, period_s1=1000.0, duration=0.0, amplitude=0.0, nstim1=0)
    yy = np.zeros(orc.MAX_STATES)
    yy[:20] = y0
    step = C.c_double(1e-3)
    t, dt, cross = 0.0, 0.05, []
    while len(cross) < 2 and t < 3000.0:
        prev = yy.copy()
        assert orc.lib().orc_rkf45_solve(orc.STW, orc.MYOCYTE, c, yy, 1, C.byref(st), t, t + dt,
                                         C.byref(step), 1e-12, 1e-10, 100000) >= 0
        t += dt
        # the slow diastolic rise through the initial potential
        if t > 200.0 and prev[0] < y0[0] <= yy[0] and yy[0] - prev[0] < 0.1 * dt:
            a = (y0[0] - prev[0]) / (yy[0] - prev[0])
            cross.append((t - dt + a * dt, prev[:20] + a * (yy[:20] - prev[:20])))
    assert len(cross) == 2
    period = cross[1][0] - cross[0][0]
    assert 1200.0 < period < 1250.0
    np.testing.assert_allclose(cross[0][1], y0, rtol=1e-6, atol=1e-12)


def test_stw_ical_ghk_limit(orc):
    """ICaL's GHK quotient (V - 15)/(exp(2 (V - 15) F/RT) - 1) is 0/0 at V = 15 mV exactly
    (Stewart_2009.H:467) and continuous through it."""
    c, y = orc.model_init(orc.STW, orc.MYOCYTE)

    def iion(v):
        yy = np.array(y)
        yy[0], yy[8] = v, 0.5  # d gate open
        return orc.model_eval(orc.STW, orc.MYOCYTE, 0.0, c, yy, solve_vm=False)["iion"]

    assert np.isnan(iion(15.0))
    left, right = 2 * iion(15.0 - 1e-4) - iion(15.0 - 2e-4), 2 * iion(15.0 + 1e-4) - iion(15.0 + 2e-4)
    assert left == pytest.approx(right, rel=1e-7)


# Grandi oracle constant slots (the GC_* order of oracle/models.c): conductances and
# permeabilities that carry K+ or Cl-, INaK (which moves K+), and the compartment volumes
GRD_K_CL_PATHS = (1, 3, 4, 5, 6, 9, 10, 11, 16, 18)   # gKr gKs gK1 gto gKur gClB GClCa IbarNaK pK gKp
GRD_VJN, GRD_VSL, GRD_VMYO, GRD_VSR = 20, 21, 22, 23


def test_grd_charge_conservation(orc):
    """Grandi's K_i and Cl_i are fixed (K_i's rate is 0), so the charge balance holds for
    the Na+ and Ca2+ currents: with the K+ and Cl- paths and INaK off, F times the
    volume-weighted change of total Na (free + NaB, three compartments) and twice that of
    total Ca (free + the buffers each compartment's balance subtracts: SLL/SLH at the
    membrane, TnC/CaM/Myo/SRB in the cytosol, Csqn in the SR) equals -Cm Iion.  Internal
    fluxes (SR uptake, leak, release, diffusion between compartments) must cancel."""
    F, Cm = 96485.0, 1.1e-10
    c0, states = _states_along_ap(orc, orc.GRD, orc.MYOCYTE, 40.0)
    c = c0.copy()
    c[list(GRD_K_CL_PATHS)] = 0.0
    Vjn, Vsl, Vmyo, Vsr = c[GRD_VJN], c[GRD_VSL], c[GRD_VMYO], c[GRD_VSR]
    (Cai, Cajn, Casl, Casr, Csqn, Nai, Najn, Nasl, CaM, Myoc, Myom, SLHjn, SLHsl, SLLjn,
     SLLsl, SRB, TnCHc, TnCHm, TnCL, NaBjn, NaBsl) = (1, 2, 3, 4, 5, 25, 26, 27, 28, 29, 30,
                                                      31, 32, 33, 34, 35, 36, 37, 38, 39, 40)
    st = orc.make_stim(start=0.0, period_s1=1000.0, duration=2.0, amplitude=40.0, nstim1=1)
    seen = 0.0
    for y in states:
        for t, stim in ((500.0, None), (1.0, st)):
            r = orc.model_eval(orc.GRD, orc.MYOCYTE, t, c, y, solve_vm=True, stim=stim)
            d = r["rates"]
            na = Vjn * (d[Najn] + d[NaBjn]) + Vsl * (d[Nasl] + d[NaBsl]) + Vmyo * d[Nai]
            ca = (Vjn * (d[Cajn] + d[SLLjn] + d[SLHjn]) + Vsl * (d[Casl] + d[SLLsl] + d[SLHsl]) +
                  Vmyo * (d[Cai] + d[TnCL] + d[TnCHc] + d[TnCHm] + d[CaM] + d[Myoc] + d[Myom] +
                          d[SRB]) + Vsr * (d[Casr] + d[Csqn]))
            assert F * (na + 2 * ca) == pytest.approx(-Cm * r["iion"], rel=1e-8,
                                                      abs=1e-9 * Cm * abs(r["iion"]) + 1e-30)
            assert d[0] == pytest.approx(-(r["iion"] + r["istim"]), rel=1e-14)
            seen = max(seen, abs(r["iion"]))
    assert seen > 1.0          # the balance was tested with currents flowing
    # with every path on, K_i stays fixed
    r = orc.model_eval(orc.GRD, orc.MYOCYTE, 500.0, c0, states[3], solve_vm=True)
    assert r["rates"][21] == 0.0


def test_grd_rush_larsen_converges_to_rkf45(orc):
    """The Rush-Larsen split (gates, linear buffer kinetics, and SR/junctional/
    subsarcolemmal Ca linearised in themselves, the other states frozen over a step) is a
    first-order discretisation of the same right-hand side: one paced atrial beat's APD90
    approaches the adaptive RKF45's (rest -73.4 mV, APD90 ~ 327 ms) with an error
    proportional to dt.  The error constant is large -- the fast Ca-buffer exchanges
    (Csqn with SR Ca, rate ~115 /ms; SLL/SLH with junctional Ca) are split -- so at 0.01 ms
    APD90 is ~40 ms long; the reference itself integrates this model with its adaptive ODE
    solvers, which the device runs as well.  The split stays stable at 0.02 ms."""
    st = orc.make_stim(start=10.0, period_s1=1000.0, duration=1.0, amplitude=40.0, nstim1=1)

    def apd(integ, dt, **kw):
        t, v, y = orc.single_cell_run(orc.GRD, orc.MYOCYTE, st, dt, int(0.6 / dt),
                                      integrator=integ, **kw)
        b = orc.detect_beats(t, v)
        assert len(b) == 1 and np.all(np.isfinite(y))
        return (b[0]["t_repol90"] - b[0]["t_up"]) * 1e3, b[0]["v_peak"]

    a_rk, p_rk = apd(1, 1e-5, abs_tol=1e-10, rel_tol=1e-8)
    assert 300.0 < a_rk < 360.0 and p_rk > 0.0
    errs = [apd(0, dt)[0] - a_rk for dt in (1e-5, 2e-6, 5e-7)]
    assert all(e > 0 for e in errs)
    assert 3.5 < errs[0] / errs[1] < 6.5          # dt ratio 5
    assert 3.0 < errs[1] / errs[2] < 5.0          # dt ratio 4
    assert errs[2] < 3.0
    apd(0, 2e-5)                                  # stable at the tissue step


@pytest.mark.parametrize("tissue", ["EPI", "MCELL", "ENDO"])
def test_tor_charge_conservation(orc, tissue):
    """ToRORd_dynCl: every ion is dynamic, chloride included (charge -1), so F times the
    volume-weighted change of Na, K, Cl and twice that of total Ca (the rapid-buffer
    factors undone, the SR pools included) equals -Acap Iion (Cm = 1 uF/cm^2)."""
    L, rad = 0.01, 0.0011
    Ageo = 2 * 3.14 * rad * rad + 2 * 3.14 * rad * L
    vcell = 1000 * 3.14 * rad * rad * L
    Acap, vmyo, vss, vnsr, vjsr = 2 * Ageo, 0.68 * vcell, 0.02 * vcell, 0.0552 * vcell, 0.0048 * vcell
    F = 96485.0
    tf = getattr(orc, tissue)
    cmdnmax = 0.05 * 1.3 if tf == orc.EPI else 0.05
    kmcmdn, trpnmax, kmtrpn = 0.00238, 0.07, 0.0005
    BSRmax, KmBSR, BSLmax, KmBSL = 0.047, 0.00087, 1.124, 0.0087
    csqnmax, kmcsqn = 10.0, 0.8
    c, states = _states_along_ap(orc, orc.TOR, tf, 53.0)
    I = dict(Nai=2, Nass=3, Ki=4, Kss=5, Cass=6, Cansr=7, Cajsr=8, Cai=9, Cli=10, Clss=11)
    seen = 0.0
    for y in states:
        r = orc.model_eval(orc.TOR, tf, 500.0, c, y, solve_vm=True)
        d = r["rates"]
        Cai, Cass, Cajsr = y[I["Cai"]], y[I["Cass"]], y[I["Cajsr"]]
        dCai_t = d[I["Cai"]] * (1 + cmdnmax * kmcmdn / (kmcmdn + Cai) ** 2 + trpnmax * kmtrpn / (kmtrpn + Cai) ** 2)
        dCass_t = d[I["Cass"]] * (1 + BSRmax * KmBSR / (KmBSR + Cass) ** 2 + BSLmax * KmBSL / (KmBSL + Cass) ** 2)
        dCajsr_t = d[I["Cajsr"]] * (1 + csqnmax * kmcsqn / (kmcsqn + Cajsr) ** 2)
        lhs = F * (vmyo * (d[I["Nai"]] + d[I["Ki"]] - d[I["Cli"]]) +
                   vss * (d[I["Nass"]] + d[I["Kss"]] - d[I["Clss"]]) +
                   2 * (vmyo * dCai_t + vss * dCass_t + vnsr * d[I["Cansr"]] + vjsr * dCajsr_t))
        assert lhs == pytest.approx(-Acap * r["iion"], rel=1e-9, abs=1e-12)
        seen = max(seen, abs(r["iion"]))
    assert seen > 0.5          # tested with currents flowing


@pytest.mark.parametrize("tissue", ["EPI", "MCELL", "ENDO"])
def test_tor_rush_larsen_converges_to_rkf45(orc, tissue):
    """The Rush-Larsen split of ToRORd_dynCl (gates; nca and the IKr Markov states linear
    in themselves) against the adaptive RKF45 on one paced beat from the reference's
    initial state: the published ToR-ORd action potential (rest about -89..-91 mV, APD90
    237 / 351 / 273 ms for epi / M / endo) and an APD90 error that shrinks with dt
    (measured 0.76 ms at 0.01 ms and 0.16 ms at 0.002 ms in M cells, below 0.15 ms in the
    others)."""
    tf = getattr(orc, tissue)
    st = orc.make_stim(start=10.0, period_s1=1000.0, duration=1.0, amplitude=53.0, nstim1=1)

    def run(integ, dt, **kw):
        t, v, y = orc.single_cell_run(orc.TOR, tf, st, dt, int(0.6 / dt), integrator=integ, **kw)
        b = orc.detect_beats(t, v)
        assert len(b) == 1 and np.all(np.isfinite(y))
        return (b[0]["t_repol90"] - b[0]["t_up"]) * 1e3, b[0]

    a_rk, b_rk = run(1, 1e-5, abs_tol=1e-10, rel_tol=1e-8)
    assert 220.0 < a_rk < 370.0 and -92.0 < b_rk["v_base"] < -88.0 and b_rk["v_peak"] > 20.0
    e1 = abs(run(0, 1e-5)[0] - a_rk)
    e2 = abs(run(0, 2e-6)[0] - a_rk)
    assert e1 < 5.0 and (e2 < 0.3 * e1 or e2 < 0.05)


def test_gaur_rush_larsen_converges_to_rkf45(orc):
    """The Rush-Larsen split of Gaur 2021 (gates and the two release fluxes; the SOICR
    timer and its refractory variable relaxed or grown exactly) against the adaptive RKF45
    on one paced beat from the reference's initial state: rest -87.3 mV, an overshoot near
    +38 mV, and an APD90 error of first order in dt (measured 2.77, 1.35 and 0.25 ms at
    0.02, 0.01 and 0.002 ms against 171 ms)."""
    st = orc.make_stim(start=10.0, period_s1=1000.0, duration=1.0, amplitude=40.0, nstim1=1)

    def run(integ, dt, **kw):
        t, v, y = orc.single_cell_run(orc.GAUR, orc.MYOCYTE, st, dt, int(0.6 / dt),
                                      integrator=integ, **kw)
        b = orc.detect_beats(t, v)
        assert len(b) == 1 and np.all(np.isfinite(y))
        return (b[0]["t_repol90"] - b[0]["t_up"]) * 1e3, b[0]

    a_rk, b_rk = run(1, 1e-5, abs_tol=1e-10, rel_tol=1e-8)
    assert 150.0 < a_rk < 200.0 and -88.0 < b_rk["v_base"] < -86.5 and 30.0 < b_rk["v_peak"] < 45.0
    e = [abs(run(0, dt)[0] - a_rk) for dt in (2e-5, 1e-5, 2e-6)]
    assert e[0] < 4.0
    assert 1.6 < e[0] / e[1] < 2.6 and 3.5 < e[1] / e[2] < 7.0     # first order
