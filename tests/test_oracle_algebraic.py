"""Pins of the oracle's ALGEBRAIC lists (oracle/models.c orc_model_algebraic:
BuenoOrovio_2008.H:166-225, AlievPanfilov_1996.H:94-120) and of the device library's
names for them, on the CPU.

No reference output holds an algebraic column, so the list is tied to the right-hand
sides the golden traces pin (orc_model_eval: the BuenoOrovio single-cell regression trace
and the AlievPanfilov reference trace): the total current equals the model's Iion, Istim
its stimulus, and in each Heaviside region the time constants and steady states equal
the Rush-Larsen (x_inf, tau) pairs of that right-hand side, each current vanishing where
its gate is closed.  A swapped index, a dropped term or a wrong branch fails one of
these.  The names are the reference's lists (BuenoOrovio_2008.H:30-43,
AlievPanfilov_1996.H:28-32)."""
import ctypes as C

import numpy as np

from cardiacfoam_amd import _capi

BO_NAMES = ["Jfi", "Jso", "Jsi", "Jion", "Istim", "tauSo", "tauO", "tauVMinus", "tauWMinus",
            "tauS", "vInfty", "wInfty"]
AP_NAMES = ["AV_eps", "Istim", "Iion_cm"]


def test_library_names_follow_the_reference():
    L = _capi.lib()
    buf = C.create_string_buffer(32)
    for model, want in ((0, BO_NAMES), (3, AP_NAMES)):
        assert L.cg_model_nalgebraic(model) == len(want)
        got = []
        for i in range(len(want)):
            _capi.check(L.cg_model_algebraic_name(model, i, buf, 32))
            got.append(buf.value.decode())
        assert got == want
    assert [L.cg_model_nalgebraic(m) for m in (1, 2, 4, 5, 6, 7, 8, 9, 10)] == [0] * 9
    assert L.cg_model_nalgebraic(11) == -1


def _bo_states(rng, n):
    u = np.concatenate([rng.uniform(-0.2, 1.6, n), [0.0, 0.006, 0.013, 0.3, 0.9]])
    return np.stack([u, rng.uniform(0, 1, len(u)), rng.uniform(0, 1, len(u)),
                     rng.uniform(0, 1, len(u))])


def test_bueno_orovio_list_against_the_right_hand_side(orc):
    """Every tissue, states across all four (v gate, w gate) regions.  The v gate is open
    (u below theta_v) where Jfi vanishes with v != 0: there the right-hand side's
    Rush-Larsen pair is (vInfty, tauVMinus); closed, its x_inf is 0.  Likewise for w with
    Jsi, where the open-gate leak makes Jso the whole current left after Jfi, and the
    closed gate's Jsi is -w s / tau_si and Jso = 1/tauSo."""
    rng = np.random.default_rng(0)
    st = orc.make_stim(start=1.0, period_s1=10.0, duration=2.0, amplitude=0.4, nstim1=3)
    tsi = orc.const_index(orc.BO, "tauSi")
    regions = set()
    for tis in (orc.EPI, 2, 3):
        c, _ = orc.model_init(orc.BO, tis)
        Y = _bo_states(rng, 400)
        for i in range(Y.shape[1]):
            y = Y[:, i]
            t = rng.uniform(0, 40)
            a = orc.model_algebraic(orc.BO, tis, t, y, stim=st, consts=c)
            e = orc.model_eval(orc.BO, tis, t, c, y, solve_vm=True, stim=st)
            assert len(a) == 12
            assert a[3] == e["iion"] and a[4] == e["istim"]
            assert a[0] + a[1] + a[2] == a[3]
            assert e["rates"][0] == -a[3] - a[4]
            assert a[9] == e["tau"][3]                        # tauS, every region
            if a[0] == 0.0 and y[1] != 0.0:
                assert e["tau"][1] == a[7] and e["xinf"][1] == a[10]
                regions.add("v open")
            else:
                assert e["xinf"][1] == 0.0
                regions.add("v closed")
            if a[2] == 0.0 and y[2] * y[3] != 0.0:
                assert e["tau"][2] == a[8] and e["xinf"][2] == a[11]
                assert a[1] == a[3] - a[0]
                regions.add("w open")
            else:
                assert a[2] == -y[2] * y[3] / c[tsi]
                assert e["xinf"][2] == 0.0
                assert abs(a[1] - 1.0 / a[5]) <= 1e-15 * abs(a[1])
                regions.add("w closed")
    assert regions == {"v open", "v closed", "w open", "w closed"}


def test_bueno_orovio_time_constants_by_region(orc):
    """tauO and tauVMinus switch at their own thresholds (theta_o, theta_v^-): the leak
    current below theta_w is (u - u_o)/tauO, and tauVMinus is the v gate's tau below
    theta_v, each checked against the right-hand side's Rush-Larsen pair."""
    c, _ = orc.model_init(orc.BO, orc.EPI)
    seen_o, seen_vm = set(), set()
    for u in np.linspace(-0.1, 0.29, 391):
        y = np.array([u, 0.7, 0.4, 0.2])
        a = orc.model_algebraic(orc.BO, orc.EPI, 0.0, y, consts=c)
        e = orc.model_eval(orc.BO, orc.EPI, 0.0, c, y)
        assert e["tau"][1] == a[7] and e["xinf"][1] == a[10]
        if a[2] == 0.0:                       # below theta_w: Jso = (u - u_o) / tauO
            uo = u - a[1] * a[6]
            assert abs(uo) <= 1e-12           # u_o = 0 in every BuenoOrovio tissue
            seen_o.add(a[6])
        seen_vm.add(a[7])
    assert len(seen_o) == 2 and len(seen_vm) == 2


def test_aliev_panfilov_list_against_the_right_hand_side(orc):
    rng = np.random.default_rng(1)
    c, _ = orc.model_init(orc.AP, 4)
    st = orc.make_stim(start=1.0, period_s1=10.0, duration=2.0, amplitude=0.5, nstim1=3)
    one = c.copy()
    one[3], one[5] = 1.0, 0.0             # eps = 1: the bare recovery term
    for _ in range(500):
        y = np.array([rng.uniform(-0.1, 1.1), rng.uniform(0.0, 2.5)])
        t = rng.uniform(0, 40)
        a = orc.model_algebraic(orc.AP, 4, t, y, stim=st, consts=c)
        e = orc.model_eval(orc.AP, 4, t, c, y, solve_vm=True, stim=st)
        bare = orc.model_eval(orc.AP, 4, t, one, y, solve_vm=True, stim=st)["rates"][1]
        assert len(a) == 3
        assert a[2] == e["iion"] and a[1] == e["istim"]
        assert e["rates"][0] == -a[2] - a[1]
        assert abs(e["rates"][1] - a[0] * bare) <= 1e-14 * max(1.0, abs(e["rates"][1]))


def test_planted_mistake_fails_a_pin(orc):
    """A swapped pair of the list (tauVMinus <-> tauWMinus) fails the Rush-Larsen pin."""
    c, _ = orc.model_init(orc.BO, orc.EPI)
    y = np.array([0.05, 0.7, 0.4, 0.2])
    a = orc.model_algebraic(orc.BO, orc.EPI, 0.0, y, consts=c)
    e = orc.model_eval(orc.BO, orc.EPI, 0.0, c, y)
    a[7], a[8] = a[8], a[7]
    assert not (e["tau"][1] == a[7] and e["tau"][2] == a[8])
