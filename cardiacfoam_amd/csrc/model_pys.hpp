// Peris-Yague 2022 human atrial model with chloride (22 states) on gfx950.
// Equations: src/ionicModels/PerisYague/PerisYague_2022.H:296-445 (rates), :211-294
// (constants and initial state); wrapper PerisYague.C:97-146 (V = 1000 Vm in mV,
// Im = Iion_cm).  Key "PerisYague_2022" (PerisYague.H:124), tissue "myocyte" only
// (PerisYague.C:92-95).  Kept as the reference has them: the h-gate beta above -40 mV is
// 0.13/(1 + exp(-(V + 10.66)/11.1)) (:333); the d-gate time constant is 1.0 inside its
// guard |V + 5| < 1e-6 (:368); the IKr/IKs alpha/beta quotients are unguarded in FP64
// (0/0 only at V = -14.1, 3.3328, 18.80816 mV exactly); IbK has gbK = 0 and enters
// neither Iion nor a balance; no stimulus in the K_i balance.
//
// Integration: 16 gates (m h j ua uif uis xr xs d f fCa qCa u v w) are Rush-Larsen;
// Na_i K_i Cl_i Ca_i Ca_up Ca_rel forward Euler.  V-only terms are evaluated once per step.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct PerisYague {
    static constexpr int NS = 22;
    static constexpr int NP = 15;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    enum { V, Nai, Ki, Cli, Cai, CaUp, CaRel, m, h, j, ua, uif, uis, xr, xs, d, f, fCa, qCa, u,
           v, w };
    // rescalable conductances and maximal fluxes, and the INaK sigma factor
    enum { gNa, gK1, gClCa, gKur, gKr, gKs, gCaL, gbCa, gbNa, INaKmax, INaCamax, IpCamax,
           Iupmax, krel, sigma };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        const double tab[NP] = {13.99, 0.08218, 0.15731, 0.45539, 0.01730, 0.0594, 0.06574,
                                0.00113, 0.000674, 0.94935, 2304.0, 0.275, 0.005, 30.0,
                                (std::exp(140.0 / 67.3) - 1.0) / 7.0};
        for (int i = 0; i < NP; ++i) p[i] = tab[i];
        const double y0[NS] = {-76.0, 11.17, 139.0, 30.0, 0.0001013, 1.488, 1.488, 2.91e-3,
                               0.965, 0.978, 4.96e-3, 1.0, 1.0, 3.29e-5, 1.87e-2, 1.37e-4,
                               0.999, 0.775, 0.0, 0.0235, 1.0, 0.999};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // fixed constants of PerisYague_2022initConsts (PerisYague_2022.H:214-270)
    static constexpr double Fc = 96.4867, Rc = 8.3143, Tc = 310.0;
    static constexpr double RTF = Rc * Tc / Fc, FRT = 1.0 / RTF;
    static constexpr double Cm = 100.0, Vi = 13668.0, Vup = 1109.52, Vrel = 96.48;
    static constexpr double Ko = 5.4, Nao = 140.0, Cao = 1.8, Clo = 132.0;
    static constexpr double KQ10 = 3.0, gam = 0.35, KmNai = 10.0, KmKo = 1.5, KmNa = 87.5,
                            KmCa = 1.38, ksat = 0.1, Kup = 0.00092, Caupmax = 15.0;
    static constexpr double CMDNmax = 0.05, TRPNmax = 0.07, CSQNmax = 10.0, KmCMDN = 0.00238,
                            KmTRPN = 0.0005, KmCSQN = 0.8, tautr = 180.0;
    static constexpr double c1 = 1.367e-13, c2 = 13.67e-16, tauu = 8.0, taufCa = 2.0;

    // x / (1 - exp(-x/s)).  FP64: the reference's expression, with its guard
    // (|x| < eps -> s) where it has one (eps = 0: none).  FP32: the series
    // s (1 + z/2 + z^2/12 - z^4/720), z = x/s, for |z| < 0.05, where 1 - exp(-z) cancels.
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
        T minf, mtau, hinf, htau, jinf, jtau, uainf, uatau, uinf, uiftau, uistau, xrinf, xrtau,
            xsinf, xstau, dinf, dtau, finf, ftau, winf, wtau;
        T gkur, rk1, rkr, eg, eg1, fnak;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T* p)
    {
        VTerms<T> q;
        // INa (PerisYague_2022.H:322-343)
        const T am = T(0.32) * xe(Vv + T(47.13), T(10), T(1e-6));
        const T bm = T(0.08) * cexp(-Vv / T(11));
        q.minf = am / (am + bm);
        q.mtau = T(1.7) / (am + bm);
        T ah, bh, aj, bj;
        if (Vv < T(-40)) {
            ah = T(0.135) * cexp(-(Vv + T(80)) / T(6.8));
            bh = T(3.56) * cexp(T(0.079) * Vv) + T(3.1e5) * cexp(T(0.35) * Vv);
            aj = (T(-127140) * cexp(T(0.2444) * Vv) - T(3.474e-5) * cexp(T(-0.04391) * Vv)) *
                 (Vv + T(37.78)) / (T(1) + cexp(T(0.311) * (Vv + T(79.23))));
            bj = T(0.1212) * cexp(T(-0.01052) * Vv) / (T(1) + cexp(T(-0.1378) * (Vv + T(40.14))));
        } else {
            ah = T(0);
            bh = T(0.13) / (T(1) + cexp(-(Vv + T(10.66)) / T(11.1)));
            aj = T(0);
            bj = T(0.3) * cexp(T(-2.535e-7) * Vv) / (T(1) + cexp(T(-0.1) * (Vv + T(32))));
        }
        q.hinf = ah / (ah + bh);
        q.htau = T(2) / (ah + bh);
        q.jinf = aj / (aj + bj);
        q.jtau = T(2) / (aj + bj);
        q.rk1 = T(1) / (T(1) + cexp(T(0.063) * (Vv + T(70))));
        // IKr (:348-353), IKs (:355-360)
        const T axr = T(0.0003) * xe(Vv + T(14.1), T(5), T(0));
        const T bxr = T(7.3898e-5) * xe(-(Vv - T(3.3328)), T(5.1237), T(0));
        q.xrinf = axr / (axr + bxr);
        q.xrtau = T(1) / (axr + bxr);
        q.rkr = T(1) / (T(1) + cexp((Vv + T(15)) / T(22.4)));
        const T axs = T(4e-5) * xe(Vv - T(18.80816), T(17), T(0));
        const T bxs = T(3.5e-5) * xe(-(Vv - T(18.80816)), T(9), T(0));
        q.xsinf = T(1) / csqrt(T(1) + cexp(-(Vv - T(19.9)) / T(12.7)));
        q.xstau = T(0.5) / (axs + bxs);
        // IKur (:362-375)
        q.gkur = p[gKur] * (T(0.005) + T(0.05) / (T(1) + cexp(-(Vv - T(15)) / T(13))));
        q.uainf = T(1) / (T(1) + cexp(-(Vv + T(30.3)) / T(9.6)));
        const T aua = T(0.65) / (cexp(-(Vv + T(10)) / T(8.5)) + cexp(-(Vv - T(30)) / T(59)));
        const T bua = T(0.65) / (T(2.5) + cexp((Vv + T(82)) / T(17)));
        q.uatau = T(1) / (T(KQ10) * (aua + bua));
        q.uinf = T(1) / (T(1) + cexp((Vv + T(17.5)) / T(5.85)));
        q.uiftau = T(400) + T(1068) * cexp(-sq(Vv / T(50)));
        q.uistau = T(2000) + T(60000) * cexp(-sq((Vv + T(39.3)) / T(30)));
        // ICaL (:377-384): d's time constant (1 - e)/(0.035 x (1 + e)), e = exp(-x/6.24)
        const T xd = Vv + T(5);
        q.dinf = T(1) / (T(1) + cexp(-xd / T(8)));
        {
            const T e = cexp(-xd / T(6.24));
            if (fabs(xd) < T(1e-6))
                q.dtau = T(1);
            else if constexpr (is_fp32<T>())
                q.dtau = T(1) / (T(0.035) * (T(1) + e) * xe(xd, T(6.24), T(0)));
            else
                q.dtau = (T(1) - e) / (T(0.035) * xd * (T(1) + e));
        }
        q.finf = T(1) / (T(1) + cexp((Vv + T(28)) / T(6.9)));
        q.ftau = T(9) / (T(0.0197) * cexp(T(-0.03372) * sq(xd)) + T(0.02));
        // release gate w (:408-410)
        q.winf = T(1) - T(1) / (T(1) + cexp(-(Vv - T(40)) / T(17)));
        {
            const T xw = Vv - T(7.9);
            const T e = cexp(-xw / T(5));
            if constexpr (is_fp32<T>())
                q.wtau = T(6) / ((T(1) + T(0.3) * e) * xe(xw, T(5), T(0)));
            else
                q.wtau = T(6) * (fabs(xw) < T(1e-6) ? T(2) / T(13)
                                                     : (T(1) - e) / ((T(1) + T(0.3) * e) * xw));
        }
        q.eg = cexp(T(gam) * Vv * T(FRT));
        q.eg1 = cexp(T(gam - 1.0) * Vv * T(FRT));
        q.fnak = T(1) / (T(1) + T(0.1245) * cexp(T(-0.1) * Vv * T(FRT)) +
                         T(0.0365) * p[sigma] * cexp(-Vv * T(FRT)));
        return q;
    }

    template <class T> struct Currents {
        T INa, IK1, IKur, IKr, IKs, ICaL, IpCa, INaK, INaCa, IbNa, IbCa, IClCa;
        CG_HD T total() const
        {
            return INa + IK1 + IKur + IKr + IKs + ICaL + IpCa + INaK + INaCa + IbNa + IbCa + IClCa;
        }
    };

    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        // gs: per-cell multipliers of gNa, gKr, gCaL, gKs (1 when unscaled)
        Currents<T> c;
        const T ENa = T(RTF) * clog(T(Nao) / y[Nai]);
        const T EK = T(RTF) * clog(T(Ko) / y[Ki]);
        const T ECa = T(0.5 * RTF) * clog(T(Cao) / y[Cai]);
        const T ECl = -T(RTF) * clog(T(Clo) / y[Cli]);
        const T dvk = Vv - EK;
        c.INa = gs[0] * p[gNa] * cube(y[m]) * y[h] * y[j] * (Vv - ENa);
        c.IK1 = p[gK1] * (dvk - T(5)) * q.rk1;
        c.IKr = gs[1] * p[gKr] * y[xr] * dvk * q.rkr;
        c.IKs = gs[3] * p[gKs] * sq(y[xs]) * dvk;
        c.IKur = q.gkur * cube(y[ua]) * (T(0.25) * y[uif] + T(0.75) * y[uis]) * dvk;
        c.ICaL = gs[2] * p[gCaL] * y[d] * y[f] * y[fCa] * (Vv - T(65));
        c.INaK = p[INaKmax] * q.fnak * T(Ko / (Ko + KmKo)) / (T(1) + cpow(T(KmNai) / y[Nai], T(1.5)));
        c.INaCa = p[INaCamax] * (q.eg * cube(y[Nai]) * T(Cao) - q.eg1 * T(Nao * Nao * Nao) * y[Cai]) /
                  (T((KmNa * KmNa * KmNa + Nao * Nao * Nao) * (KmCa + Cao)) * (T(1) + T(ksat) * q.eg1));
        c.IbNa = p[gbNa] * (Vv - ENa);
        c.IbCa = p[gbCa] * (Vv - ECa);
        c.IpCa = p[IpCamax] * y[Cai] / (T(0.0005) + y[Cai]);
        c.IClCa = p[gClCa] * y[qCa] * (Vv - ECl);
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy; V has a rate only
    // when solve_v (-Iion - Istim).  Every state is read before the policy writes it.
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, const T* gs, bool solve_v,
                             T istim, const VTerms<T>& q)
    {
        const Currents<T> c = currents(y[V], y, p, q, gs);
        const T Irel = p[krel] * sq(y[u]) * y[v] * y[w] * (y[CaRel] - y[Cai]);
        const T Itr = (y[CaUp] - y[CaRel]) / T(tautr);
        const T Iup = p[Iupmax] / (T(1) + T(Kup) / y[Cai]);
        const T Ileak = p[Iupmax] * y[CaUp] / T(Caupmax);
        const T Fn = T(1e-12 * Vrel) * Irel - T(5e-13 / Fc * Cm) * (T(0.5) * c.ICaL - T(0.2) * c.INaCa);
        const T eu = cexp(-(Fn - T(c1)) / T(c2));
        const T B1 = (T(2) * c.INaCa - (c.IpCa + c.ICaL + c.IbCa)) * T(Cm / (2.0 * Vi * Fc)) +
                     (T(Vup) * (Ileak - Iup) + Irel * T(Vrel)) / T(Vi);
        const T B2 = T(1) + T(TRPNmax * KmTRPN) / sq(y[Cai] + T(KmTRPN)) +
                     T(CMDNmax * KmCMDN) / sq(y[Cai] + T(KmCMDN));
        const T dCaRel = (Itr - Irel) / (T(1) + T(CSQNmax * KmCSQN) / sq(y[CaRel] + T(KmCSQN)));
        const T dCaUp = Iup - (Ileak + Itr * T(Vrel / Vup));
        const T fcainf = T(1) / (T(1) + y[Cai] / T(0.00035));
        out.gate(qCa, T(1) - cube(T(1) / (T(1) + Fn / T(1.1e-10))), T(2));
        out.gate(u, T(1) / (T(1) + eu), T(tauu));
        out.gate(v, T(1) - T(1) / (T(1) + cexp(-(Fn - T(0.2 * c1)) / T(c2))),
                 T(1.91) + T(2.09) / (T(1) + eu));
        out.gate(w, q.winf, q.wtau);
        out.gate(m, q.minf, q.mtau);
        out.gate(h, q.hinf, q.htau);
        out.gate(j, q.jinf, q.jtau);
        out.gate(ua, q.uainf, q.uatau);
        out.gate(uif, q.uinf, q.uiftau);
        out.gate(uis, q.uinf, q.uistau);
        out.gate(xr, q.xrinf, q.xrtau);
        out.gate(xs, q.xsinf, q.xstau);
        out.gate(d, q.dinf, q.dtau);
        out.gate(f, q.finf, q.ftau);
        out.gate(fCa, fcainf, T(taufCa));
        const T cvf = T(Cm / (Vi * Fc));
        out.rate(Nai, (T(-3) * c.INaK - (T(3) * c.INaCa + c.IbNa + c.INa)) * cvf);
        out.rate(Ki, (T(2) * c.INaK - (c.IK1 + c.IKur + c.IKr + c.IKs)) * cvf);
        out.rate(Cli, c.IClCa * cvf);
        out.rate(Cai, B1 / B2);
        out.rate(CaUp, dCaUp);
        out.rate(CaRel, dCaRel);
        if (solve_v) out.rate(V, -c.total() - istim);
    }

    template <class T>
    CG_HD static void step(T* y, T dt, const T* p, const T* gs, bool solve_v, T istim,
                            const VTerms<T>& q)
    {
        RlStep<T> out{y, dt};
        rates(y, out, p, gs, solve_v, istim, q);
    }

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
