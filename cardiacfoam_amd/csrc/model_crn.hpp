// Courtemanche-Ramirez-Nattel 1998 human atrial model (21 states) on gfx950.
// Equations: src/ionicModels/Courtemanche/Courtemanche_1998.H:458-658 (rates),
// :337-455 (constants and initial state); wrapper Courtemanche.C:106-147 (V = 1000 Vm in
// mV, Im = Iion_cm).  The reference accepts tissue "myocyte" only (Courtemanche.C:92-95).
// The K_i balance omits the stimulus: the reference reads ALGEBRAIC[Istim] before it is
// assigned (Courtemanche_1998.H:644 vs :652) from the zeroed scratch of derivatives().
//
// Integration: the 15 gates (m h j oa oi ua ui xr xs d f fCa u v w) are Rush-Larsen;
// Na_i K_i Ca_i Ca_up Ca_rel are forward Euler.  Everything that depends on V alone
// (13 voltage-gated steady states and time constants, rectification and exchanger
// factors) is evaluated once per step in VTerms and reused by every RKF45 stage.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct Courtemanche {
    static constexpr int NS = 21;
    static constexpr int NP = 15;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    enum { V, Nai, Ki, Cai, CaUp, CaRel, m, h, j, oa, oi, ua, ui, xr, xs, d, f, fCa, u, v, w };
    // the conductances and maximal fluxes a caller may rescale, and the INaK sigma factor
    enum { gNa, gK1, gto, gKur, gKr, gKs, gCaL, gbCa, gbNa, INaKmax, INaCamax, IpCamax,
           Iupmax, Krel, sigma };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        const double tab[NP] = {7.8, 0.09, 0.1652, 0.005, 0.029411765, 0.12941176, 0.12375,
                                0.001131, 0.0006744375, 0.59933874, 1600.0, 0.275, 0.005,
                                30.0, (std::exp(140.0 / 67.3) - 1.0) / 7.0};
        for (int i = 0; i < NP; ++i) p[i] = tab[i];
        const double y0[NS] = {
            -81.9463303822041098, 13.8169746305367962, 136.355229902154434,
            1.23092247890489894e-04, 1.54668119199095355, 1.07650740580354909,
            2.56385228666526068e-03, 0.970298907063270155, 0.981123905023234988,
            2.91755626557170314e-02, 0.999342865333055497, 4.58838038240151104e-03,
            0.991468962753066063, 8.33819909884048389e-04, 1.86683180787284714e-02,
            1.24231529593716656e-04, 0.951907788168154578, 0.739682838459564729,
            -1.97647749727073971e-40, 1.0, 0.999233799248152699};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // fixed constants of CourtemancheinitConsts (Courtemanche_1998.H:341-431)
    static constexpr double Fc = 96.4867, Rc = 8.3143, Tc = 310.0;
    static constexpr double RTF = Rc * Tc / Fc, FRT = 1.0 / RTF;
    static constexpr double Cm = 100.0, Vcell = 20100.0, Vi = 0.68 * Vcell,
                            Vrel = 0.0048 * Vcell, Vup = 0.0552 * Vcell;
    static constexpr double Cao = 1.8, Ko = 5.4, Nao = 140.0;
    static constexpr double CMDNmax = 0.05, CSQNmax = 10.0, KmCMDN = 0.00238, KmCSQN = 0.8,
                            KmTRPN = 0.0005, TRPNmax = 0.07, Caupmax = 15.0, Kup = 0.00092;
    static constexpr double ECaL = 65.0, taufCa = 2.0, tautr = 180.0, KQ10 = 3.0;
    static constexpr double KmCa = 1.38, KmNa = 87.5, gam = 0.35, ksat = 0.1, KmKo = 1.5,
                            KmNai = 10.0;
    static constexpr double c1 = 3.4175e-13, c2 = 1.367e-15, tauu = 8.0;

    // x / (1 - exp(-x/s)), the removable singularity of the alpha/beta rate expressions.
    // FP64 keeps the generated guard (|x| < eps -> s); FP32 takes the series
    // s (1 + z/2 + z^2/12 - z^4/720), z = x/s, for |z| < 0.05 (truncation < 1e-12),
    // where 1 - exp(-z) would cancel.
    template <class T> CG_HD static T xe(T x, T s, T eps)
    {
        if constexpr (is_fp32<T>()) {
            const T z = x / s;
            if (fabsf(z) < 0.05f) {
                const T z2 = z * z;
                return s * (T(1) + T(0.5) * z + z2 * (T(1) / T(12) - z2 * (T(1) / T(720))));
            }
            return x / (T(1) - cexp(-z));
        } else {
            return fabs(x) < eps ? s : x / (T(1) - cexp(-x / s));
        }
    }

    template <class T> struct VTerms {
        T minf, mtau, hinf, htau, jinf, jtau, oainf, oatau, oiinf, oitau, uainf, uatau,
            uiinf, uitau, xrinf, xrtau, xsinf, xstau, dinf, dtau, finf, ftau, winf, wtau;
        T gkur, rk1, rkr, eg, eg1, fnak;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T* p)
    {
        VTerms<T> q;
        // INa (Courtemanche_1998.H:605-625)
        T am;
        if constexpr (is_fp32<T>())
            am = T(0.32) * xe(Vv + T(47.13), T(10), T(0));
        else
            am = Vv == T(-47.13) ? T(3.2)
                                 : T(0.32) * (Vv + T(47.13)) / (T(1) - cexp(T(-0.1) * (Vv + T(47.13))));
        const T bm = T(0.08) * cexp(-Vv / T(11));
        q.minf = am / (am + bm);
        q.mtau = T(1) / (am + bm);
        T ah, bh, aj, bj;
        if (Vv < T(-40)) {
            ah = T(0.135) * cexp((Vv + T(80)) / T(-6.8));
            bh = T(3.56) * cexp(T(0.079) * Vv) + T(310000) * cexp(T(0.35) * Vv);
            aj = (T(-127140) * cexp(T(0.2444) * Vv) - T(3.474e-05) * cexp(T(-0.04391) * Vv)) *
                 (Vv + T(37.78)) / (T(1) + cexp(T(0.311) * (Vv + T(79.23))));
            bj = T(0.1212) * cexp(T(-0.01052) * Vv) / (T(1) + cexp(T(-0.1378) * (Vv + T(40.14))));
        } else {
            ah = T(0);
            bh = T(1) / (T(0.13) * (T(1) + cexp((Vv + T(10.66)) / T(-11.1))));
            aj = T(0);
            bj = T(0.3) * cexp(T(-2.535e-07) * Vv) / (T(1) + cexp(T(-0.1) * (Vv + T(32))));
        }
        q.hinf = ah / (ah + bh);
        q.htau = T(1) / (ah + bh);
        q.jinf = aj / (aj + bj);
        q.jtau = T(1) / (aj + bj);
        // Ito and IKur share the activation time constant (:586-587, :628-629)
        const T aoa = T(0.65) / (cexp((Vv + T(10)) / T(-8.5)) + cexp((Vv - T(30)) / T(-59)));
        const T boa = T(0.65) / (T(2.5) + cexp((Vv + T(82)) / T(17)));
        const T tact = T(1) / (aoa + boa) / T(KQ10);
        q.oainf = T(1) / (T(1) + cexp((Vv + T(20.47)) / T(-17.54)));
        q.oatau = tact;
        const T aoi = T(1) / (T(18.53) + cexp((Vv + T(113.7)) / T(10.95)));
        const T boi = T(1) / (T(35.56) + cexp((Vv + T(1.26)) / T(-7.44)));
        q.oiinf = T(1) / (T(1) + cexp((Vv + T(43.1)) / T(5.3)));
        q.oitau = T(1) / (aoi + boi) / T(KQ10);
        q.uainf = T(1) / (T(1) + cexp((Vv + T(30.3)) / T(-9.6)));
        q.uatau = tact;
        const T aui = T(1) / (T(21) + cexp((Vv - T(185)) / T(-28)));
        const T bui = T(1) / cexp((Vv - T(158)) / T(-16));
        q.uiinf = T(1) / (T(1) + cexp((Vv - T(99.45)) / T(27.48)));
        q.uitau = T(1) / (aui + bui) / T(KQ10);
        q.gkur = T(1) + T(10) / (T(1) + cexp((Vv - T(15)) / T(-13)));
        // IKr, IKs (:566-581)
        const T axr = T(0.0003) * xe(Vv + T(14.1), T(5), T(1e-06));
        const T bxr = T(7.3898e-05) * xe(-(Vv - T(3.3328)), T(5.1237), T(1e-07));
        q.xrinf = T(1) / (T(1) + cexp((Vv + T(14.1)) / T(-6.5)));
        q.xrtau = T(1) / (axr + bxr);
        q.rkr = T(1) / (T(1) + cexp((Vv + T(15)) / T(22.4)));
        const T axs = T(4e-05) * xe(Vv - T(19.9), T(17), T(1e-06));
        const T bxs = T(3.5e-05) * xe(-(Vv - T(19.9)), T(9), T(1e-06));
        q.xsinf = T(1) / csqrt(T(1) + cexp((Vv - T(19.9)) / T(-12.7)));
        q.xstau = T(0.5) / (axs + bxs);
        // ICaL d, f (:517-522); d's time constant is (1 - e)/(0.035 x (1 + e)), e = exp(-x/6.24)
        const T xd = Vv + T(10);
        q.dinf = T(1) / (T(1) + cexp(xd / T(-8)));
        {
            const T e = cexp(xd / T(-6.24));
            if constexpr (is_fp32<T>())
                q.dtau = T(1) / (T(0.035) * (T(1) + e) * xe(xd, T(6.24), T(0)));
            else
                q.dtau = fabs(xd) < T(1e-06) ? T(1) / (T(6.24) * T(2) * T(0.035))
                                             : (T(1) - e) / (T(0.035) * xd * (T(1) + e));
        }
        q.finf = T(1) / (T(1) + cexp((Vv + T(28)) / T(6.9)));
        q.ftau = T(9) / (T(0.0197) * cexp(-(T(0.0337) * T(0.0337)) * sq(xd)) + T(0.02));
        // release gate w (:548-549): 6 (1 - e)/((1 + 0.3 e) x), e = exp(-x/5), x = V - 7.9
        q.winf = T(1) - T(1) / (T(1) + cexp(-(Vv - T(40)) / T(17)));
        {
            const T xw = Vv - T(7.9);
            const T e = cexp(-xw / T(5));
            if constexpr (is_fp32<T>())
                q.wtau = T(6) / ((T(1) + T(0.3) * e) * xe(xw, T(5), T(0)));
            else
                q.wtau = T(6) * (fabs(xw) < T(1e-06) ? T(2) / T(13)
                                                      : (T(1) - e) / ((T(1) + T(0.3) * e) * xw));
        }
        q.rk1 = T(1) / (T(1) + cexp(T(0.07) * (Vv + T(80))));
        q.eg = cexp(T(gam) * Vv * T(FRT));
        q.eg1 = cexp(T(gam - 1.0) * Vv * T(FRT));
        q.fnak = T(1) / (T(1) + T(0.1245) * cexp(T(-0.1) * Vv * T(FRT)) +
                         T(0.0365) * p[sigma] * cexp(-Vv * T(FRT)));
        return q;
    }

    template <class T> struct Currents {
        T INa, IK1, Ito, IKur, IKr, IKs, ICaL, IpCa, INaK, INaCa, IbNa, IbCa;
        CG_HD T total() const
        {
            return INa + IK1 + Ito + IKur + IKr + IKs + ICaL + IpCa + INaK + INaCa + IbNa + IbCa;
        }
    };

    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        // gs: per-cell multipliers of gNa, gKr, gCaL, gKs (1 when unscaled)
        Currents<T> c;
        const T ECa = T(0.5 * RTF) * clog(T(Cao) / y[Cai]);
        const T EK = T(RTF) * clog(T(Ko) / y[Ki]);
        const T ENa = T(RTF) * clog(T(Nao) / y[Nai]);
        const T dvk = Vv - EK;
        c.INa = gs[0] * p[gNa] * cube(y[m]) * y[h] * y[j] * (Vv - ENa);
        c.IK1 = p[gK1] * dvk * q.rk1;
        c.Ito = p[gto] * cube(y[oa]) * y[oi] * dvk;
        c.IKur = p[gKur] * q.gkur * cube(y[ua]) * y[ui] * dvk;
        c.IKr = gs[1] * p[gKr] * y[xr] * dvk * q.rkr;
        c.IKs = gs[3] * p[gKs] * sq(y[xs]) * dvk;
        c.ICaL = gs[2] * p[gCaL] * y[d] * y[f] * y[fCa] * (Vv - T(ECaL));
        c.IpCa = p[IpCamax] * y[Cai] / (T(0.0005) + y[Cai]);
        const T nai3 = cube(y[Nai]);
        c.INaCa = p[INaCamax] * (q.eg * nai3 * T(Cao) - q.eg1 * T(Nao * Nao * Nao) * y[Cai]) /
                  (T((KmNa * KmNa * KmNa + Nao * Nao * Nao) * (KmCa + Cao)) *
                   (T(1) + T(ksat) * q.eg1));
        c.INaK = p[INaKmax] * q.fnak * T(Ko / (Ko + KmKo)) /
                 (T(1) + cpow(T(KmNai) / y[Nai], T(1.5)));
        c.IbNa = p[gbNa] * (Vv - ENa);
        c.IbCa = p[gbCa] * (Vv - ECa);
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy; V has a rate only
    // when solve_v (-Iion - Istim).
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, const T* gs, bool solve_v,
                             T istim, const VTerms<T>& q)
    {
        const Currents<T> c = currents(y[V], y, p, q, gs);
        // SR release driven by the flux signal Fn (:547-558)
        const T Irel = p[Krel] * sq(y[u]) * y[v] * y[w] * (y[CaRel] - y[Cai]);
        const T Fn = T(1e-12 * Vrel) * Irel - T(5e-13 / Fc * Cm) * (T(0.5) * c.ICaL - T(0.2) * c.INaCa);
        const T eu = cexp(-(Fn - T(c1)) / T(c2));
        out.gate(u, T(1) / (T(1) + eu), T(tauu));
        out.gate(v, T(1) - T(1) / (T(1) + cexp(-(Fn - T(0.2 * c1)) / T(c2))),
                 T(1.91) + T(2.09) / (T(1) + eu));
        out.gate(w, q.winf, q.wtau);
        out.gate(m, q.minf, q.mtau);
        out.gate(h, q.hinf, q.htau);
        out.gate(j, q.jinf, q.jtau);
        out.gate(oa, q.oainf, q.oatau);
        out.gate(oi, q.oiinf, q.oitau);
        out.gate(ua, q.uainf, q.uatau);
        out.gate(ui, q.uiinf, q.uitau);
        out.gate(xr, q.xrinf, q.xrtau);
        out.gate(xs, q.xsinf, q.xstau);
        out.gate(d, q.dinf, q.dtau);
        out.gate(f, q.finf, q.ftau);
        out.gate(fCa, T(1) / (T(1) + y[Cai] / T(0.00035)), T(taufCa));
        // calcium handling (:509-512, :531, :636-641)
        const T Iup = p[Iupmax] / (T(1) + T(Kup) / y[Cai]);
        const T Ileak = p[Iupmax] * y[CaUp] / T(Caupmax);
        const T Itr = (y[CaUp] - y[CaRel]) / T(tautr);
        const T dCaRel = (Itr - Irel) / (T(1) + T(CSQNmax * KmCSQN) / sq(y[CaRel] + T(KmCSQN)));
        const T dCaUp = Iup - (Ileak + Itr * T(Vrel / Vup));
        const T B1 = (T(2) * c.INaCa - (c.IpCa + c.ICaL + c.IbCa)) * T(Cm / (2.0 * Vi * Fc)) +
                     (T(Vup) * (Ileak - Iup) + Irel * T(Vrel)) / T(Vi);
        const T B2 = T(1) + T(TRPNmax * KmTRPN) / sq(y[Cai] + T(KmTRPN)) +
                     T(CMDNmax * KmCMDN) / sq(y[Cai] + T(KmCMDN));
        const T cvf = T(Cm / (Vi * Fc));
        out.rate(Nai, (T(-3) * c.INaK - (T(3) * c.INaCa + c.IbNa + c.INa)) * cvf);
        out.rate(Ki, (T(2) * c.INaK - (c.IK1 + c.Ito + c.IKur + c.IKr + c.IKs)) * cvf);
        out.rate(Cai, B1 / B2);
        out.rate(CaUp, dCaUp);
        out.rate(CaRel, dCaRel);
        if (solve_v) out.rate(V, -c.total() - istim);
    }

    // one RL/FE step of dt in place
    template <class T>
    CG_HD static void step(T* y, T dt, const T* p, const T* gs, bool solve_v, T istim,
                            const VTerms<T>& q)
    {
        RlStep<T> out{y, dt};
        rates(y, out, p, gs, solve_v, istim, q);
    }

    // dy/dt at y (dV/dt = 0 unless V is integrated with the other states)
    template <class T>
    CG_HD static void deriv(const T* y, T* dy, const T* p, const T* gs, bool solve_v, T istim,
                             const VTerms<T>& q)
    {
        Deriv<T> out{y, dy};
        dy[V] = T(0);
        rates(y, out, p, gs, solve_v, istim, q);
    }
};

} // namespace cg
