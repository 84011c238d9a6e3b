// Stewart et al. 2009 human Purkinje-fibre model (20 states) on gfx950.
// Equations: src/ionicModels/Stewart/Stewart_2009.H:380-530 (rates), :262-377
// (constants and initial state); wrapper Stewart.C:102-155 (V = 1000 Vm in mV,
// Im = Iion_cm).  Tissue "myocyte" only (Stewart.C:93-96).  In tissue mode dV/dt is 0
// (the reference leaves RATES[V] unwritten there, :524-528).
//
// Integration: the 13 voltage/Ca gates (y Xr1 Xr2 Xs m h j d f f2 fCass s r) and the RyR
// state R' (dR'/dt = k4 - (k2 Ca_ss + k4) R', a linear gate) are Rush-Larsen; Ca_i Ca_SR
// Ca_ss Na_i K_i forward Euler.  V-only terms are evaluated once per step.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct Stewart {
    static constexpr int NS = 20;
    static constexpr int NP = 17;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    enum { V, y_, Xr1, Xr2, Xs, m, h, j, d, f, f2, fCass, s, r, Cai, CaSR, Cass, Rp, Nai, Ki };
    // rescalable conductances and maximal fluxes
    enum { gNa, gK1, gto, gsus, gKr, gKs, gCaL, gbca, gbna, gpCa, gpK, gfK, gfNa, PNaK, KNaCa,
           Vmaxup, Vrel };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        const double tab[NP] = {130.5744, 0.065, 0.08184, 0.0227, 0.0918, 0.2352, 3.98e-05,
                                0.000592, 0.00029, 0.1238, 0.0146, 0.0234346, 0.0145654,
                                2.724, 1000.0, 0.006375, 0.102};
        for (int i = 0; i < NP; ++i) p[i] = tab[i];
        const double y0[NS] = {
            -69.1370441635924067, 4.57562667986601973e-02, 5.50281999719088016e-03,
            0.313213286437995009, 9.53708522974788982e-03, 4.17391656294996971e-02,
            0.190678733735144990, 0.238219836154029002, 2.87906256206415015e-04,
            0.989328560287986991, 0.995474890442184956, 0.999955429598213041,
            0.963861017995010005, 1.03618091196912006e-03, 1.01878186157052002e-04,
            3.10836886659417022, 4.46818714055411022e-04, 0.991580051907845039,
            8.80420286531673035, 136.781894160227012};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // fixed constants of StewartinitConsts (Stewart_2009.H:264-358)
    static constexpr double Fc = 96485.3415, Rc = 8314.472, Tc = 310.0;
    static constexpr double RTF = Rc * Tc / Fc, FRT = Fc / (Rc * Tc);
    static constexpr double Cm = 0.185, Vc = 0.016404, Vsr = 0.001094, Vss = 5.468e-05;
    static constexpr double Ko = 5.4, Nao = 140.0, Cao = 2.0, Pkna = 0.03;
    static constexpr double KpCa = 0.0005, Bufc = 0.2, Bufsr = 10.0, Bufss = 0.4, EC = 1.5,
                            Kbufc = 0.001, Kbufsr = 0.3, Kbufss = 0.00025, Kup = 0.00025,
                            Vleak = 0.00036, Vxfer = 0.0038, k1p = 0.15, k2p = 0.045, k3 = 0.06,
                            k4 = 0.005, maxsr = 2.5, minsr = 1.0;
    static constexpr double Ksat = 0.1, KmCa = 1.38, KmNai = 87.5, alpha = 2.5, gam = 0.35,
                            KmNa = 40.0, Kmk = 1.0;

    template <class T> struct VTerms {
        T dinf, dtau, f2inf, f2tau, finf, ftau, minf, mtau, hinf, htau, jtau, yinf, ytau,
            xr1inf, xr1tau, xr2inf, xr2tau, xsinf, xstau, rinf, rtau, sinf, stau;
        T nak, sus, rk1, ipk, e2, ghk, ncx_num, ncx_den;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T*)
    {
        VTerms<T> q;
        // ICaL d, f2, f (Stewart_2009.H:382-397)
        const T ad = T(1.4) / (T(1) + cexp((T(-35) - Vv) / T(13))) + T(0.25);
        const T bd = T(1.4) / (T(1) + cexp((Vv + T(5)) / T(5)));
        const T gd = T(1) / (T(1) + cexp((T(50) - Vv) / T(20)));
        q.dinf = T(1) / (T(1) + cexp((T(-8) - Vv) / T(7.5)));
        q.dtau = ad * bd + gd;
        q.f2inf = T(0.67) / (T(1) + cexp((Vv + T(35)) / T(7))) + T(0.33);
        q.f2tau = T(562) * cexp(-sq(Vv + T(27)) / T(240)) + T(31) / (T(1) + cexp((T(25) - Vv) / T(10))) +
                  T(80) / (T(1) + cexp((Vv + T(30)) / T(10)));
        q.finf = T(1) / (T(1) + cexp((Vv + T(20)) / T(7)));
        q.ftau = T(1102.5) * cexp(-sq(Vv + T(27)) / T(225)) + T(200) / (T(1) + cexp((T(13) - Vv) / T(10))) +
                 T(180) / (T(1) + cexp((Vv + T(30)) / T(10))) + T(20);
        // INa m, h, j (:400-419)
        T ah, bh, aj, bj;
        if (Vv < T(-40)) {
            ah = T(0.057) * cexp(-(Vv + T(80)) / T(6.8));
            bh = T(2.7) * cexp(T(0.079) * Vv) + T(310000) * cexp(T(0.3485) * Vv);
            aj = (T(-25428) * cexp(T(0.2444) * Vv) - T(6.948e-06) * cexp(T(-0.04391) * Vv)) *
                 (Vv + T(37.78)) / (T(1) + cexp(T(0.311) * (Vv + T(79.23))));
            bj = T(0.02424) * cexp(T(-0.01052) * Vv) / (T(1) + cexp(T(-0.1378) * (Vv + T(40.14))));
        } else {
            ah = T(0);
            bh = T(0.77) / (T(0.13) * (T(1) + cexp((Vv + T(10.66)) / T(-11.1))));
            aj = T(0);
            bj = T(0.6) * cexp(T(0.057) * Vv) / (T(1) + cexp(T(-0.1) * (Vv + T(32))));
        }
        q.hinf = T(1) / sq(T(1) + cexp((Vv + T(71.55)) / T(7.43)));
        q.htau = T(1) / (ah + bh);
        q.jtau = T(1) / (aj + bj);
        const T am = T(1) / (T(1) + cexp((T(-60) - Vv) / T(5)));
        const T bm = T(0.1) / (T(1) + cexp((Vv + T(35)) / T(5))) + T(0.1) / (T(1) + cexp((Vv - T(50)) / T(200)));
        q.minf = T(1) / sq(T(1) + cexp((T(-56.86) - Vv) / T(9.03)));
        q.mtau = am * bm;
        // If y, IKr, IKs, Ito (:420-449)
        q.yinf = T(1) / (T(1) + cexp((Vv + T(80.6)) / T(6.8)));
        q.ytau = T(4000) / (cexp(T(-2.9) - T(0.04) * Vv) + cexp(T(3.6) + T(0.11) * Vv));
        q.xr1inf = T(1) / (T(1) + cexp((T(-26) - Vv) / T(7)));
        q.xr1tau = T(450) / (T(1) + cexp((T(-45) - Vv) / T(10))) * (T(6) / (T(1) + cexp((Vv + T(30)) / T(11.5))));
        q.xr2inf = T(1) / (T(1) + cexp((Vv + T(88)) / T(24)));
        q.xr2tau = T(3) / (T(1) + cexp((T(-60) - Vv) / T(20))) * (T(1.12) / (T(1) + cexp((Vv - T(60)) / T(20))));
        q.xsinf = T(1) / (T(1) + cexp((T(-5) - Vv) / T(14)));
        q.xstau = T(1400) / csqrt(T(1) + cexp((T(5) - Vv) / T(6))) * (T(1) / (T(1) + cexp((Vv - T(35)) / T(15)))) + T(80);
        q.rinf = T(1) / (T(1) + cexp((T(20) - Vv) / T(13)));
        q.rtau = T(10.45) * cexp(-sq(Vv + T(40)) / T(1800)) + T(7.3);
        q.sinf = T(1) / (T(1) + cexp((Vv + T(27)) / T(13)));
        q.stau = T(85) * cexp(-sq(Vv + T(25)) / T(320)) + T(5) / (T(1) + cexp((Vv - T(40)) / T(5))) + T(42);
        // voltage factors of the currents
        q.nak = T(1) / (T(1) + T(0.1245) * cexp(T(-0.1) * Vv * T(FRT)) + T(0.0353) * cexp(-Vv * T(FRT)));
        q.sus = T(1) / (T(1) + cexp((T(5) - Vv) / T(17)));
        q.rk1 = T(1) / (T(1) + cexp(T(0.1) * (Vv + T(75.44))));
        q.ipk = T(1) / (T(1) + cexp((T(25) - Vv) / T(5.98)));
        q.e2 = cexp(T(2) * (Vv - T(15)) * T(FRT));
        q.ghk = ghk_quot(Vv - T(15), T(2 * FRT), q.e2);  // (V - 15)/(e2 - 1)
        q.ncx_num = cexp(T(gam) * Vv * T(FRT));
        q.ncx_den = cexp(T(gam - 1.0) * Vv * T(FRT));
        return q;
    }

    template <class T> struct Currents {
        T IK1, Ito, Isus, IKr, IKs, ICaL, INaK, INa, IbNa, INaCa, IbCa, IpK, IpCa, IfNa, IfK;
        CG_HD T total() const
        {
            return IK1 + Ito + Isus + IKr + IKs + ICaL + INaK + INa + IbNa + INaCa + IbCa + IpK +
                   IpCa + (IfNa + IfK);
        }
    };

    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        // gs: per-cell multipliers of gNa, gKr, gCaL, gKs (1 when unscaled)
        Currents<T> c;
        const T EK = T(RTF) * clog(T(Ko) / y[Ki]);
        const T ENa = T(RTF) * clog(T(Nao) / y[Nai]);
        const T ECa = T(0.5 * RTF) * clog(T(Cao) / y[Cai]);
        const T EKs = T(RTF) * clog(T(Ko + Pkna * Nao) / (y[Ki] + T(Pkna) * y[Nai]));
        const T dvk = Vv - EK;
        c.INaK = p[PNaK] * T(Ko / (Ko + Kmk)) * y[Nai] / (y[Nai] + T(KmNa)) * q.nak;
        c.Isus = p[gsus] * q.sus * dvk;
        c.Ito = p[gto] * y[r] * y[s] * dvk;
        c.ICaL = gs[2] * p[gCaL] * y[d] * y[f] * y[f2] * y[fCass] * T(4 * Fc * FRT) *
                 (T(0.25) * y[Cass] * q.e2 - T(Cao)) * q.ghk;
        c.IbCa = p[gbca] * (Vv - ECa);
        c.IfK = y[y_] * p[gfK] * dvk;
        c.IfNa = y[y_] * p[gfNa] * (Vv - ENa);
        c.IK1 = p[gK1] * q.rk1 * (dvk - T(8));
        c.IpK = p[gpK] * dvk * q.ipk;
        c.IKr = gs[1] * p[gKr] * y[Xr1] * y[Xr2] * dvk;  // sqrt(Ko/5.4) = 1 at Ko = 5.4
        c.IKs = gs[3] * p[gKs] * sq(y[Xs]) * (Vv - EKs);
        c.IpCa = p[gpCa] * y[Cai] / (y[Cai] + T(KpCa));
        c.INaCa = p[KNaCa] * (q.ncx_num * cube(y[Nai]) * T(Cao) - q.ncx_den * T(Nao * Nao * Nao) * y[Cai] * T(alpha)) /
                  (T((KmNai * KmNai * KmNai + Nao * Nao * Nao) * (KmCa + Cao)) * (T(1) + T(Ksat) * q.ncx_den));
        c.INa = gs[0] * p[gNa] * cube(y[m]) * y[h] * y[j] * (Vv - ENa);
        c.IbNa = p[gbna] * (Vv - ENa);
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy; V has a rate only
    // when solve_v (-Iion - Istim).  Every state is read before the policy writes it.
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, const T* gs, bool solve_v,
                             T istim, const VTerms<T>& q)
    {
        const Currents<T> c = currents(y[V], y, p, q, gs);
        const T bufc = T(1) / (T(1) + T(Bufc * Kbufc) / sq(y[Cai] + T(Kbufc)));
        const T bufsr = T(1) / (T(1) + T(Bufsr * Kbufsr) / sq(y[CaSR] + T(Kbufsr)));
        const T bufss = T(1) / (T(1) + T(Bufss * Kbufss) / sq(y[Cass] + T(Kbufss)));
        const T Ileak = T(Vleak) * (y[CaSR] - y[Cai]);
        const T Iup = p[Vmaxup] / (T(1) + T(Kup * Kup) / sq(y[Cai]));
        const T Ixfer = T(Vxfer) * (y[Cass] - y[Cai]);
        const T kcasr = T(maxsr) - T(maxsr - minsr) / (T(1) + sq(T(EC) / y[CaSR]));
        const T k1 = T(k1p) / kcasr, k2 = T(k2p) * kcasr;
        const T css2 = sq(y[Cass]);
        const T O = k1 * css2 * y[Rp] / (T(k3) + k1 * css2);
        const T Irel = p[Vrel] * O * (y[CaSR] - y[Cass]);
        const T cq = T(1) + sq(y[Cass] / T(0.05));
        const T dCai = bufc * ((Ileak - Iup) * T(Vsr / Vc) + Ixfer -
                               (c.IbCa + c.IpCa - T(2) * c.INaCa) * T(Cm / (2.0 * Vc * Fc)));
        const T dCass = bufss * (-c.ICaL * T(Cm / (2.0 * Vss * Fc)) + Irel * T(Vsr / Vss) -
                                 Ixfer * T(Vc / Vss));
        const T dCaSR = bufsr * (Iup - (Irel + Ileak));
        const T rp_b = k2 * y[Cass] + T(k4);
        out.lin(Rp, T(k4), rp_b);
        out.gate(fCass, T(0.6) / cq + T(0.4), T(80) / cq + T(2));
        out.gate(d, q.dinf, q.dtau);
        out.gate(f2, q.f2inf, q.f2tau);
        out.gate(f, q.finf, q.ftau);
        out.gate(m, q.minf, q.mtau);
        out.gate(h, q.hinf, q.htau);
        out.gate(j, q.hinf, q.jtau);
        out.gate(y_, q.yinf, q.ytau);
        out.gate(Xr1, q.xr1inf, q.xr1tau);
        out.gate(Xr2, q.xr2inf, q.xr2tau);
        out.gate(Xs, q.xsinf, q.xstau);
        out.gate(r, q.rinf, q.rtau);
        out.gate(s, q.sinf, q.stau);
        const T cvf = T(Cm / (Vc * Fc));
        out.rate(Cai, dCai);
        out.rate(Cass, dCass);
        out.rate(CaSR, dCaSR);
        out.rate(Ki, -(c.IK1 + c.Ito + c.IfK + c.Isus + c.IKr + c.IKs + c.IpK - T(2) * c.INaK) * cvf);
        out.rate(Nai, -(c.INa + c.IbNa + c.IfNa + T(3) * c.INaK + T(3) * c.INaCa) * cvf);
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
