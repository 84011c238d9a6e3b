// Gaur et al. 2021 human ventricular myocyte with spontaneous SR release (29 states) on
// gfx950.  Equations: src/ionicModels/Gaur/Gaur_2021.H:649-921 (rates), :460-647 (constants
// and initial state); wrapper Gaur.C:106-155 (V = 1000 Vm in mV, Im = Iion_cm).  Tissue
// "myocyte" only (Gaur.C:93-96).  The model's own factors are kept as written (e.g.
// Acap / (2 F vmyo) in the K_i and Na_i balances).
//
// Integration (Rush-Larsen): the Hodgkin-Huxley gates and the two release fluxes as gates;
// ITo_aa linear in itself; the SOICR timer and its refractory variable either relax to 0
// (tau 1 us / 0.1 ms, taken exactly) or grow at rate 1; CaMKt and the concentrations
// forward Euler.  V-only terms are evaluated once per step.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct Gaur {
    static constexpr int NS = 29;
    static constexpr int NP = 14;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    // STATES order of Gaur_2021Names.H
    enum { V, nai, nass, ki, kss, cai, cai2, cass, cansr, cajsr, cacsr, m, h, j, ml, hl, d, fca,
           xr, xs1, xs2, aa, Jrel2, Jrel1, CaMKt, tjsrol, A, fs, ff };
    // rescalable conductances and permeabilities of initConsts (:496-606); E_Cl (:491),
    // which needs a log
    enum { GNa, GKr, PCa, GKs, GNaL, GK1, Gto, Gncx, Pnak, GKb, PNab, PCab, GpCa, ECl };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        const double tab[ECl] = {14.838, 0.5 * 0.015, 0.0002, 0.021, 0.0075, 0.75 * 0.2,
                                0.4 * 0.5, 1.0 * 0.0008, 1.0 * 30.0, 0.003, 3.75e-10,
                                1.0 * 2.5e-08, 0.0575};
        for (int i = 0; i < ECl; ++i) p[i] = tab[i];
        p[ECl] = Rc * Tc / Fc * log(cli / clo);
        const double y0[NS] = {
            -87.3, 6.43, 6.43, 140.76, 140.76, 6.75e-05, 6.81e-05, 6.75e-05, 1.37, 1.31, 1.3,
            0.0022, 0.631, 0.631, 0.001, 0.301, 4.42e-07, 0.988, 0.153, 0.054, 0.046, 0.995,
            8.59e-10, 9.71e-22, 8.23910999999999914e-03, 100.0, 100.0, 0.0, 0.0};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // fixed constants of initConsts (:463-642)
    static constexpr double Fc = 96485.0, Rc = 8314.0, Tc = 310.0;
    static constexpr double cao = 1.0 * 1.8, cli = 19.53, clo = 100.0, ko = 5.4, nao = 140.0;
    static constexpr double L = 0.017, PI_ = 3.14, rad = 0.0011, vmyo1frac = 0.4;
    static constexpr double Ageo = 2.0 * PI_ * rad * rad + 2.0 * PI_ * rad * L;
    static constexpr double vcell = 1000.0 * PI_ * rad * rad * L;
    static constexpr double Acap = 2.0 * Ageo;
    static constexpr double vjsr = 0.0048 * vcell / 2.0, vmyo = 0.68 * vcell,
                            vnsr = 0.0552 * vcell, vss = 0.02 * vcell / 2.0, vcsr = vjsr;
    static constexpr double vmyo1 = vmyo * vmyo1frac, vnsr1 = vnsr / 2.0, vnsr2 = vnsr / 2.0,
                            vmyo2 = vmyo - vmyo1;
    static constexpr double KmCaMK = 0.065;

    template <class T> struct VTerms {
        T hinf, tauh, tauj, minf, taum, hlinf, mlinf, taul, dinf, taud, finf, tauff, taufs;
        T xrinf, tauxr, rkr, xsinf, tauxs1, tauxs2, rk1, alpha_aa, beta_aa, rito2, xkb;
        T hca, hna, Knai, Knao, e1, e2, FRT;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T*)
    {
        VTerms<T> q;
        // I_Na (:671-689)
        T ah, bh, aj, bj;
        if (Vv >= T(-40)) {
            ah = T(0);
            aj = T(0);
            bh = T(0.77) / (T(0.13) * (T(1) + cexp(-(Vv + T(10.66)) / T(11.1))));
            bj = T(0.6) * cexp(T(0.057) * Vv) / (T(1) + cexp(T(-0.1) * (Vv + T(32))));
        } else {
            ah = T(0.057) * cexp(-(Vv + T(80)) / T(6.8));
            aj = (T(-25428) * cexp(T(0.2444) * Vv) - T(6.948e-06) * cexp(T(-0.04391) * Vv)) *
                 (Vv + T(37.78)) / (T(1) + cexp(T(0.311) * (Vv + T(79.23))));
            bh = T(2.7) * cexp(T(0.079) * Vv) + T(310000) * cexp(T(0.3485) * Vv);
            bj = T(0.02424) * cexp(T(-0.01052) * Vv) / (T(1) + cexp(T(-0.1378) * (Vv + T(40.14))));
        }
        q.tauh = T(1) / (ah + bh);
        q.tauj = T(1) / (aj + bj);
        q.hinf = T(1) / sq(T(1) + cexp((Vv + T(71.55)) / T(7.43)));
        q.minf = T(1) / sq(T(1) + cexp((T(-56.86) - Vv) / T(9.03)));
        q.taum = (T(1) / (T(1) + cexp((T(-60) - Vv) / T(5)))) *
                 (T(0.1) / (T(1) + cexp((Vv + T(35)) / T(5))) + T(0.1) / (T(1) + cexp((Vv - T(50)) / T(200))));
        // INaL (:691-699); the generated code substitutes 10 at V = -47.13
        const T w = Vv + T(47.13);
        T aml;
        if constexpr (is_fp32<T>()) {
            aml = T(0.32) * ghk_quot(w, T(-0.1), cexp(T(-0.1) * w)) * T(-1);
        } else {
            aml = T(0.32) * (Vv == T(-47.13) ? T(10) : -w / (cexp(T(-0.1) * w) - T(1)));
        }
        const T bml = T(0.08) * cexp(-Vv / T(11));
        q.hlinf = T(1) / (T(1) + cexp((Vv + T(91)) / T(6.1)));
        q.mlinf = aml / (aml + bml);
        q.taul = T(1) / (aml + bml);
        // ICaL gates (:701-724)
        q.dinf = T(1) / (T(1) + cexp(-(Vv - T(3.4)) / T(6.2)));
        q.taud = T(0.6) + T(1) / (cexp(T(-0.05) * (Vv + T(6))) + cexp(T(0.09) * (Vv + T(14))));
        q.finf = T(1) / (T(1) + cexp((Vv - T(-22.9)) / T(4.9))) + T(0.35) / (T(1) + cexp((T(45) - Vv) / T(20)));
        q.tauff = T(7) + T(1) / (T(0.0045) * cexp(-(Vv + T(20)) / T(10)) + T(0.0045) * cexp((Vv + T(20)) / T(10)));
        q.taufs = T(70) + T(1) / (T(3.5e-05) * cexp(-(Vv + T(5)) / T(4)) + T(3.5e-05) * cexp((Vv + T(5)) / T(6)));
        // IKr, IKs, IK1, ITo, IKb (:726-753, :853-855)
        q.rkr = T(1) / (T(1) + cexp((Vv + T(22)) / T(15)));
        q.tauxr = T(12.98) + T(1) / (T(0.3652) * cexp((Vv - T(31.66)) / T(3.869)) +
                                     T(4.123e-05) * cexp(-(Vv - T(47.78)) / T(20.38)));
        q.xrinf = T(1) / (T(1) + cexp(-(Vv + T(56.8)) / T(17.8)));
        q.tauxs1 = T(300) + T(1) / (T(1e-06) * cexp((Vv + T(50)) / T(20)) + T(0.04) * cexp(-(Vv + T(50)) / T(20)));
        q.tauxs2 = T(1) / (T(0.01) * cexp((Vv - T(50)) / T(20)) + T(0.0193) * cexp(-(Vv + T(66.54)) / T(31)));
        q.xsinf = T(1) / (T(1) + cexp(-(Vv - T(25.1)) / T(37.1)));
        q.rk1 = T(1) / (T(1) + cexp((Vv + T(79.3) - T(2.6 * ko)) / T(19.6)));
        q.alpha_aa = T(0.025) / (T(1) + cexp((Vv + T(58)) / T(5)));
        q.beta_aa = T(1) / (T(5) * (T(1) + cexp((Vv + T(19)) / T(-9))));
        q.rito2 = T(1) / (T(1) + cexp(-(Vv + T(10)) / T(5)));
        q.xkb = T(1) / (T(1) + cexp(-(Vv - T(14.48)) / T(18.34)));
        // exchanger and pump voltage factors, GHK exponentials
        q.FRT = T(Fc / (Rc * Tc));
        const T vfrt = Vv * q.FRT;
        q.hca = cexp(T(0.167) * vfrt);
        q.hna = cexp(T(0.5224) * vfrt);
        q.Knai = T(9.073) * cexp(T(-0.155) * vfrt / T(3));
        q.Knao = T(27.78) * cexp(T(1.0 - -0.155) * vfrt / T(3));
        q.e1 = cexp(vfrt);
        q.e2 = cexp(T(2) * vfrt);
        return q;
    }

    // vfrt / (e^vfrt - 1) and 2 vfrt / (e^{2 vfrt} - 1), 1 at V = 0 (the generated guards)
    template <class T> CG_HD static void ghk(T Vv, const VTerms<T>& q, T& g1, T& g2)
    {
        if constexpr (is_fp32<T>()) {
            g1 = ghk_quot(Vv, q.FRT, q.e1) * q.FRT;
            g2 = ghk_quot(Vv, T(2) * q.FRT, q.e2) * (T(2) * q.FRT);
        } else {
            const T vfrt = Vv * q.FRT;
            g1 = Vv == T(0) ? T(1) : vfrt / (q.e1 - T(1));
            g2 = Vv == T(0) ? T(1) : T(2) * vfrt / (q.e2 - T(1));
        }
    }

    // Na/Ca exchanger of one compartment (:771-819)
    template <class T> CG_HD static T ncx(T na, T ca, const VTerms<T>& q, T gfrac)
    {
        const T kna1 = T(15), kna2 = T(5), kna3 = T(88.12), kcaon = T(1500000), kcaoff = T(5000);
        const T wca = T(60000), wna = T(60000), wnaca = T(5000);
        const T h10 = T(12.5) + T(1) + T(nao) / kna1 * (T(1) + T(nao) / kna2);
        const T h11 = T(nao * nao) / (h10 * kna1 * kna2);
        const T k1 = (T(1) / h10) * T(cao) * kcaon, k2 = kcaoff, k5 = kcaoff;
        const T allo = T(1) / (T(1) + sq(T(0.00015) / ca));
        const T h4 = T(1) + na / kna1 * (T(1) + na / kna2);
        const T h1 = T(1) + na / kna3 * (T(1) + q.hna);
        const T h5 = na * na / (h4 * kna1 * kna2);
        const T h7 = T(1) + T(nao) / kna3 * (T(1) + T(1) / q.hna);
        const T h2 = na * q.hna / (kna3 * h1);
        const T h8 = T(nao) / (kna3 * q.hna * h7);
        const T k6 = (T(1) / h4) * ca * kcaon;
        const T k3pp = h8 * wnaca, k4pp = h2 * wnaca;
        const T k3 = (T(1) / h7) * wca + k3pp, k4 = (T(1) / h1) * wca / q.hca + k4pp;
        const T k7 = h5 * h2 * wna, k8 = h8 * h11 * wna;
        const T x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
        const T x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
        const T x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
        const T x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
        const T sx = x1 + x2 + x3 + x4;
        const T E1 = x1 / sx, E2 = x2 / sx, E3 = x3 / sx, E4 = x4 / sx;
        const T JCa = E2 * k2 - E1 * k1;
        const T JNa = T(3) * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
        return gfrac * allo * (JNa + T(2) * JCa);
    }

    template <class T> struct Currents {
        T INa, INaL, ICaL, ICaNa, ICaK, IKr, IKs, IK1, ITo, INaCai, INaCass, INaK, INab, IKb,
          IpCa, ICab, Jrel, CaMKa;
        CG_HD T total() const
        {
            return INa + INaL + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + ITo + INaCai + INaCass +
                   INaK + INab + IKb + IpCa + ICab;
        }
    };

    // membrane currents (:656-875); gs: per-cell multipliers of GNa, GKr, PCa, GKs
    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        Currents<T> c;
        const T RTF = T(Rc * Tc / Fc);
        const T CaMKb = T(0.05) * (T(1) - y[CaMKt]) / (T(1) + T(0.0015) / y[cass]);
        c.CaMKa = CaMKb + y[CaMKt];
        const T EK = RTF * clog(T(ko) / y[ki]);
        const T ENa = RTF * clog(T(nao) / y[nai]);
        const T EKs = RTF * clog(T(ko + 0.01833 * nao) / (y[ki] + T(0.01833) * y[nai]));
        const T dvk = Vv - EK;
        T g1, g2;
        ghk(Vv, q, g1, g2);
        c.INa = gs[0] * p[GNa] * cube(y[m]) * y[h] * y[j] * (Vv - ENa);
        c.INaL = p[GNaL] * cube(y[ml]) * y[hl] * (Vv - ENa);
        const T dff = y[d] * (y[ff] * y[fs]) * y[fca];
        const T pca = gs[2] * p[PCa];
        c.ICaL = pca * (T(2 * Fc) * (y[cass] * q.e2 - T(0.341 * cao)) * g2) * dff;
        c.ICaK = T(0.0003574) * pca * (T(Fc) * (T(0.75) * y[kss] * q.e1 - T(0.75 * ko)) * g1) * dff;
        c.ICaNa = T(0.00125) * pca * (T(Fc) * (T(0.75) * y[nass] * q.e1 - T(0.75 * nao)) * g1) * dff;
        c.IKr = gs[1] * p[GKr] * y[xr] * q.rkr * dvk;                // sqrt(ko / 5.4) = 1
        const T KsCa = T(1) + T(0.6) / (T(1) + cpow(T(3.8e-05) / y[cai], T(1.4)));
        c.IKs = gs[3] * p[GKs] * KsCa * y[xs1] * y[xs2] * (Vv - EKs);
        c.IK1 = p[GK1] * q.rk1 * dvk;
        // release flux and the Ca-activated ITo2 (:755-769, :906-909)
        const T tj = y[tjsrol];
        const T grel = T(2) * (T(1) - cexp(-tj / T(0.5))) * cexp(-tj / T(5));
        c.Jrel = y[Jrel1] + grel * (y[cajsr] - y[cass]);
        const T kito2 = T(1) - T(1) / (T(1) + sq(c.Jrel / T(0.4)));
        c.ITo = p[Gto] * y[aa] * q.rito2 * kito2 * (Vv - p[ECl]);
        c.INaCai = ncx(y[nai], y[cai], q, T(0.8) * p[Gncx]);
        c.INaCass = ncx(y[nass], y[cass], q, T(0.2) * p[Gncx]);
        {
            const T ni = y[nai] / q.Knai, no = T(nao) / q.Knao, kk = y[ki] / T(0.5);
            const T kko = T(ko / 0.3582);
            const T P = T(4.2) / (T(1) + T(1e-07 / 1.698e-07) + y[nai] / T(224) + y[ki] / T(292));
            const T dni = cube(T(1) + ni) + sq(T(1) + kk) - T(1);
            const T dno = cube(T(1) + no) + sq(T(1) + kko) - T(1);
            const T a1 = T(949.5) * cube(ni) / dni;
            const T a3 = T(1899.0) * sq(kko) / dno;
            const T b2 = T(39.4) * cube(no) / dno;
            const T b3 = T(79300.0) * P * T(1e-07) / T(1.0 + 9.8 / 1.698e-07);
            const T b4 = T(40.0) * sq(kk) / dni;
            const T a2 = T(687.2), a4 = T(639.0 * 9.8 / 1.698e-07 / (1.0 + 9.8 / 1.698e-07));
            const T b1 = T(182.4 * 0.05);
            const T x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
            const T x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
            const T x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
            const T x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
            const T sx = x1 + x2 + x3 + x4;
            c.INaK = p[Pnak] * (T(3) * (x1 / sx * a3 - x2 / sx * b3) + T(2) * (x4 / sx * b1 - x3 / sx * a1));
        }
        c.IKb = p[GKb] * q.xkb * dvk;
        c.INab = p[PNab] * T(Fc) * (y[nai] * q.e1 - T(nao)) * g1;
        c.ICab = p[PCab] * T(2 * Fc) * (y[cai] * q.e2 - T(0.341 * cao)) * g2;
        c.IpCa = p[GpCa] * y[cai] / (T(0.0005) + y[cai]);
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy; V has a rate only when
    // solve_v (-Iion - Istim).  Every state is read before the policy writes any.
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, const T* gs, bool solve_v,
                             T istim, const VTerms<T>& q)
    {
        const Currents<T> c = currents(y[V], y, p, q, gs);
        const T Iion = c.total();
        const T CaMKb = c.CaMKa - y[CaMKt];
        const T dCaMKt = T(0.05) * CaMKb * (CaMKb + y[CaMKt]) - T(0.00068) * y[CaMKt];
        const T fcamk = T(1) / (T(1) + T(KmCaMK) / c.CaMKa);
        const T camk8 = T(1) + T(1) / (T(1) + cpow(T(KmCaMK) / c.CaMKa, T(8)));
        const T Jdiff = (y[cass] - y[cai]) / T(0.2);
        const T JdiffK = (y[kss] - y[ki]) / T(2);
        const T JdiffNa = (y[nass] - y[nai]) / T(2);
        const T Jgap = (y[cai] - y[cai2]) / T(3);
        // SR uptake (:877-890)
        const T Jleak = T(0.005) * y[cansr] / T(15);
        const T Jtr = (y[cansr] - y[cajsr]) / T(100), Jtr2 = (y[cansr] - y[cacsr]) / T(100);
        const T Jup = (T(1) - fcamk) * (T(0.005) * y[cai] / (y[cai] + T(0.001))) +
                      fcamk * (T(3.0 * 0.005) * y[cai] / (y[cai] + T(0.001 - 0.0002))) - Jleak;
        const T Jup2 = (T(1) - fcamk) * (T(0.005) * y[cai2] / (y[cai2] + T(0.001))) +
                       fcamk * (T(3.0 * 0.005) * y[cai2] / (y[cai2] + T(0.001 - 0.0002))) - Jleak;
        // buffers and concentrations (:892-903)
        const T Bcacsr = T(1) / (T(1) + T(10.0 * 0.8) / sq(T(0.8) + y[cacsr]));
        const T Bcai = T(1) / (T(1) + T(0.05 * 0.00238) / sq(T(0.00238) + y[cai]) +
                               T(0.07 * 0.0005) / sq(T(0.0005) + y[cai]));
        const T Bcai2 = T(1) / (T(1) + T(0.05 * 0.00238) / sq(T(0.00238) + y[cai2]) +
                                T(0.07 * 0.0005) / sq(T(0.0005) + y[cai2]));
        const T Bcajsr = T(1) / (T(1) + T(10.0 * 0.8) / sq(T(0.8) + y[cajsr]));
        const T Bcass = T(1) / (T(1) + T(0.047 * 0.00087) / sq(T(0.00087) + y[cass]) +
                                T(1.124 * 0.0087) / sq(T(0.0087) + y[cass]));
        const T A2F = T(Acap / (2.0 * Fc));
        const T dcansr = Jup * T(vnsr1 / vnsr) + Jup2 * T(vnsr2 / vnsr) - Jtr * T(vjsr / vnsr) -
                         Jtr2 * T(vcsr / vnsr);
        const T dki = -(c.IKr + c.IKs + c.IK1 + c.IKb - T(2) * c.INaK) * A2F / T(vmyo) +
                      JdiffK * T(vss / vmyo);
        const T dkss = -c.ICaK * A2F / T(vss) - JdiffK;
        const T dnai = -(c.INa + c.INaL + T(3) * c.INaCai + T(3) * c.INaK + c.INab) * A2F / T(vmyo) +
                       JdiffNa * T(vss / vmyo);
        const T dnass = -(c.ICaNa + T(3) * c.INaCass) * A2F / T(vss) - JdiffNa;
        const T dcacsr = Bcacsr * (Jtr2 - y[Jrel2]);
        const T dcai = Bcai * (-(c.IpCa + c.ICab - T(2) * c.INaCai) * A2F / T(2 * vmyo1) -
                               Jup * T(vnsr1 / vmyo1) + Jdiff * T(vss / vmyo1) - Jgap);
        const T dcai2 = Bcai2 * (y[Jrel2] * T(vcsr / vmyo2) + Jgap * T(vmyo2 / vmyo1) -
                                 Jup2 * T(vnsr2 / vmyo2));
        const T dcajsr = Bcajsr * (Jtr - c.Jrel);
        const T dcass = Bcass * (-(c.ICaL - T(2) * c.INaCass) * A2F / T(2 * vss) +
                                 c.Jrel * T(vjsr / vss) - Jdiff);
        // release gates (:757-768, :906-913)
        T trel1 = T(20) * camk8 / (T(1) + cpow(T(0.5) / y[cajsr], T(8)));
        trel1 = T(0.001) > trel1 ? T(0.001) : trel1;
        T trel2 = T(50) * camk8 / (T(1) + cpow(T(0.5) / y[cacsr], T(8)));
        trel2 = T(0.001) > trel2 ? T(0.001) : trel2;
        const T Rel1 = (-c.ICaL + T(2) * c.INaCass) * T(Acap / (2.0 * vss * Fc)) +
                       c.Jrel * T(vjsr / vss) - Jdiff;
        const T Rel2 = Jgap + y[Jrel2] * T(vcsr / vmyo2) - Jup2 * T(vnsr2 / vmyo2);
        const T Jrel1_inf = Rel1 > T(0) ? T(60) * Rel1 * camk8 / (T(1) + cpow(T(0.75) / y[cajsr], T(8))) : T(0);
        const T Jrel2_inf = Rel2 > T(0) ? T(250) * Rel2 * camk8 / (T(1) + cpow(T(0.75) / y[cacsr], T(8))) : T(0);
        // ICaL inactivation targets (:707, :715)
        const T fca_inf = T(0.3) / (T(1) - (c.ICaL > T(0) ? T(0) : c.ICaL / T(0.05))) +
                          T(0.55) / (T(1) + y[cass] / T(0.003)) + T(0.15);
        const T taufca = T(10) * fcamk + T(0.5) + T(1) / (T(1) + y[cass] / T(0.003));
        const bool a_decays = y[tjsrol] < T(5);
        const bool t_resets = y[cajsr] > T(10) && y[A] > T(45);
        const T Av = y[A], tv = y[tjsrol], aav = y[aa];

        // writes
        out.gate(h, q.hinf, q.tauh);
        out.gate(j, q.hinf, q.tauj);
        out.gate(m, q.minf, q.taum);
        out.gate(hl, q.hlinf, T(600));
        out.gate(ml, q.mlinf, q.taul);
        out.gate(d, q.dinf, q.taud);
        out.gate(fca, fca_inf, taufca);
        out.gate(ff, q.finf, q.tauff);
        out.gate(fs, q.finf, q.taufs);
        out.gate(xr, q.xrinf, q.tauxr);
        out.gate(xs1, q.xsinf, q.tauxs1);
        out.gate(xs2, q.xsinf, q.tauxs2);
        out.lin(aa, q.alpha_aa, q.alpha_aa + q.beta_aa,
                q.alpha_aa * (T(1) - aav) - q.beta_aa * aav);
        if (a_decays) out.lin(A, T(0), T(10), -Av / T(0.1));
        else out.rate(A, T(1));
        if (t_resets) out.lin(tjsrol, T(0), T(1000), -tv / T(0.001));
        else out.rate(tjsrol, T(1));
        out.gate(Jrel1, Jrel1_inf, trel1);
        out.gate(Jrel2, Jrel2_inf, trel2);
        out.rate(CaMKt, dCaMKt);
        out.rate(cansr, dcansr);
        out.rate(ki, dki);
        out.rate(kss, dkss);
        out.rate(nai, dnai);
        out.rate(nass, dnass);
        out.rate(cacsr, dcacsr);
        out.rate(cai, dcai);
        out.rate(cai2, dcai2);
        out.rate(cajsr, dcajsr);
        out.rate(cass, dcass);
        if (solve_v) out.rate(V, -Iion - istim);
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
