// Grandi et al. 2011 human atrial myocyte (41 states) on gfx950.
// Equations: src/ionicModels/Grandi/Grandi_2011.H:660-931 (rates), :414-657 (constants
// with AF = ISO = RA = 0 as GrandiinitConsts sets them, and the initial state); wrapper
// Grandi.C:106-155 (V = 1000 Vm in mV, Im = Iion_cm).  Tissue "myocyte" only
// (Grandi.C:93-96).  K_i is constant (its rate is 0).  In tissue mode dV/dt is 0.
//
// Integration (Rush-Larsen): the Hodgkin-Huxley gates (d f xr xs ikur_r ikur_s h j m hl x
// y) as gates; first-order kinetics linear in their own state (ml, fCaB_jn/sl, the Ca and
// Na buffers, Csqn) as linear gates with the other states frozen over the step; Ca_sr,
// Ca_jn and Ca_sl, which are stiff at a 10 us step (SR release, the small junction volume
// and the buffer on-rates give -d(rate)/dx ~ 1e2..1e3 per ms), as linear in themselves
// with the diagonal coefficient b of those terms frozen (x' = a - b x, a = rate + b x);
// the RyR states, Ca_i and the Na pools forward Euler.  Each rate is written in the
// reference's own form; the linear split only sets the Rush-Larsen (x_inf, tau).
#pragma once
#include "cg_common.hpp"

namespace cg {

struct Grandi {
    static constexpr int NS = 41;
    static constexpr int NP = 22;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    // STATES order of Grandi_2011Names.H
    enum { V, Cai, Cajn, Casl, Casr, Csqn, d, f, fCaBjn, fCaBsl, xr, xs, kur_r, kur_s, h, j, m,
           hl, ml, x, y_, Ki, ri, ro, rr, Nai, Najn, Nasl, CaM, Myoc, Myom, SLHjn, SLHsl,
           SLLjn, SLLsl, SRB, TnCHc, TnCHm, TnCL, NaBjn, NaBsl };
    // rescalable conductances, permeabilities and maximal fluxes; sigma (INaK) and E_Cl,
    // constants of initConsts that need exp/log
    enum { gNa, gKr, pCa, gKs, gK1, gto, gKur, gCaB, gNaB, gClB, GClCa, IbarNaK, IbarNCX,
           IbarSLCaP, VmaxSRCaP, ks, pK, pNa, gKp, gNaL, sigma, ECl };

    // geometry of GrandiinitConsts (Grandi_2011.H:440-470)
    static constexpr double PI = 3.14159265358979312, len = 100.0, rad = 10.25;
    static constexpr double Vcell = PI * (rad * rad) * len * 1e-15;
    static constexpr double Vjn = 0.0539 * 0.01 * Vcell, Vmyo = 0.65 * Vcell,
                            Vsl = 0.02 * Vcell, Vsr = 0.035 * Vcell;
    static constexpr double Fc = 96485.0, Rc = 8314.0, Tc = 310.0, FRT = Fc / Rc / Tc;
    static constexpr double Cmem = 1.1e-10;
    static constexpr double Fjn = 0.11, Fsl = 1.0 - 0.11, FjnL = 0.9, FslL = 1.0 - 0.9;
    static constexpr double Ko = 5.4, Nao = 140.0, Cao = 1.8, Mgi = 1.0;
    static constexpr double BmaxSLLj = 0.0046 * Vmyo / Vjn * 0.1, BmaxSLLsl = Vmyo / Vsl * 0.0374,
                            BmaxSLHj = 0.00165 * Vmyo / Vjn * 0.1, BmaxSLHsl = Vmyo / Vsl * 0.0134,
                            BmaxCsqn = Vmyo / Vsr * 0.14;
    // diffusion between the compartments (Grandi_2011.H:462-468 evaluated)
    static constexpr double JCaslmyo = 3.72425607984805052e-12, JCajnsl = 8.24130542277896849e-13,
                            JNaslmyo = 1.63862792221979433e-12, JNajnsl = 1.83127823220607955e-14;
    static constexpr double koff_tncl = 0.0196, kleak = 5.348e-06;

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        const double tab[20] = {23.0, 0.035, 0.00027, 0.0035, 0.0525, 0.165, 0.045,
                                6.06430000000000033e-04, 0.000597, 0.009, 0.0548, 1.26, 3.15,
                                0.0471, 0.0053114, 25.0, 1.35e-07, 7.5e-09, 0.002, 0.0};
        for (int i = 0; i < 20; ++i) p[i] = tab[i];
        p[sigma] = (exp(140.0 / 67.3) - 1.0) / 7.0;
        p[ECl] = 1.0 / FRT * log(15.0 / 150.0);
        const double y0[NS] = {
            -7.34336366728778671e+01, 2.10808768153058460e-04, 3.25814677291117296e-04,
            2.33018340557575125e-04, 5.02305826642838293e-01, 1.13337536953687845e+00,
            2.16850216379767157e-05, 9.98384427312367095e-01, 4.49572164109603364e-02,
            3.28512098597005947e-02, 1.31290096227093382e-03, 7.49436760722081534e-03,
            3.93548562883350357e-04, 9.58234428284286399e-01, 3.15482710277587786e-01,
            2.48034071360795916e-01, 1.89326933812916480e-02, 3.79829335413739144e-02,
            1.01974216400706526e-02, 1.37939236359928058e-03, 9.45874848392074696e-01, 120.0,
            5.01323282772066123e-07, 2.01567245823636694e-06, 8.00819151705148946e-01,
            9.15199678386256998e+00, 9.15153381546177336e+00, 9.15182798281732346e+00,
            7.02128101897185673e-04, 3.94923428392655786e-03, 1.35538532457244482e-01,
            1.03674364292988680e-01, 1.90759804527589089e-01, 1.35640688636079511e-02,
            2.14063418881809235e-02, 4.45327242854324807e-03, 1.27856586024588575e-01,
            5.69999505293381902e-03, 1.83143535034222225e-02, 3.61396062660070427e+00,
            7.88607791910409195e-01};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // (1 - exp(-w/6)) / w of the d gate's tau: the reference's form in FP64; in FP32 through
    // the series of ghk_quot near w = 0, where 1 - exp cancels
    template <class T> CG_HD static T dtau_frac(T w)
    {
        if constexpr (is_fp32<T>()) {
            const T e = cexp(w * T(1.0 / 6.0));
            return (T(1) / e) / ghk_quot(w, T(1.0 / 6.0), e);
        } else {
            return (T(1) - cexp(-w / T(6))) / w;
        }
    }

    template <class T> struct VTerms {
        T xrinf, xrtau, krinf, krtau, ksinf, kstau, hjinf, htau, jtau, minf, mtau, hlinf, aml,
            bml, xinf, xtau, yinf, ytau, dinf, dtau, finf, ftau, xsinf, xstau;
        T fnak, e1, e2, ghk1, ghk2, en, en1, kp, rkr;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T* p)
    {
        VTerms<T> q;
        // IKr, IKur (Grandi_2011.H:663-676)
        q.xrinf = T(1) / (T(1) + cexp(-(Vv + T(10)) / T(5)));
        q.xrtau = T(550) / (T(1) + cexp((T(-22) - Vv) / T(9))) * T(6) / (T(1) + cexp((Vv + T(11)) / T(9))) +
                  T(230) / (T(1) + cexp((Vv + T(40)) / T(20)));
        q.krinf = T(1) / (T(1) + cexp((Vv + T(6)) / T(-8.6)));
        q.krtau = T(9) / (T(1) + cexp((Vv + T(5)) / T(12))) + T(0.5);
        q.ksinf = T(1) / (T(1) + cexp((Vv + T(7.5)) / T(10)));
        q.kstau = T(590) / (T(1) + cexp((Vv + T(60)) / T(10))) + T(3050);
        // INa h, j, m and INaL (:677-700)
        q.hjinf = T(1) / sq(T(1) + cexp((Vv + T(71.55)) / T(7.43)));
        T ah, bh, aj, bj;
        if (Vv >= T(-40)) {
            ah = T(0);
            bh = T(0.77) / (T(0.13) * (T(1) + cexp(-(Vv + T(10.66)) / T(11.1))));
            aj = T(0);
            bj = T(0.6) * cexp(T(0.057) * Vv) / (T(1) + cexp(T(-0.1) * (Vv + T(32))));
        } else {
            ah = T(0.057) * cexp(-(Vv + T(80)) / T(6.8));
            bh = T(2.7) * cexp(T(0.079) * Vv) + T(3.1e5) * cexp(T(0.3485) * Vv);
            aj = (T(-2.5428e4) * cexp(T(0.2444) * Vv) - T(6.948e-6) * cexp(T(-0.04391) * Vv)) *
                 (Vv + T(37.78)) / (T(1) + cexp(T(0.311) * (Vv + T(79.23))));
            bj = T(0.02424) * cexp(T(-0.01052) * Vv) / (T(1) + cexp(T(-0.1378) * (Vv + T(40.14))));
        }
        q.htau = T(1) / (ah + bh);
        q.jtau = T(1) / (aj + bj);
        q.minf = T(1) / sq(T(1) + cexp(-(T(56.86) + Vv) / T(9.03)));
        q.mtau = T(0.1292) * cexp(-sq((Vv + T(45.79)) / T(15.54))) +
                 T(0.06487) * cexp(-sq((Vv - T(4.823)) / T(51.12)));
        q.hlinf = T(1) / (T(1) + cexp((Vv + T(91)) / T(6.1)));
        const T wm = Vv + T(47.13);
        q.aml = T(-0.32) * ghk_quot(wm, T(-0.1), cexp(T(-0.1) * wm));  // 0.32 w / (1 - e^{-w/10})
        q.bml = T(0.08) * cexp(-Vv / T(11));
        // Ito, ICaL, IKs gates (:701-733)
        q.xinf = T(1) / (T(1) + cexp(-(Vv + T(1)) / T(11)));
        q.xtau = T(3.5) * cexp(-sq(Vv / T(30))) + T(1.5);
        q.yinf = T(1) / (T(1) + cexp((Vv + T(40.5)) / T(11.5)));
        q.ytau = T(25.635) * cexp(-sq((Vv + T(52.45)) / T(15.8827))) + T(24.14);
        q.dinf = T(1) / (T(1) + cexp(-(Vv + T(9)) / T(6)));
        q.dtau = q.dinf * dtau_frac(Vv + T(9)) / T(0.035);
        q.finf = T(1) / (T(1) + cexp((Vv + T(30)) / T(7))) + T(0.2) / (T(1) + cexp((T(50) - Vv) / T(20)));
        q.ftau = T(1) / (T(0.0197) * cexp(-sq(T(0.0337) * (Vv + T(25)))) + T(0.02));
        q.xsinf = T(1) / (T(1) + cexp(-(Vv + T(3.8)) / T(14.25)));
        q.xstau = T(990.1) / (T(1) + cexp(-(Vv + T(2.436)) / T(14.12)));
        // voltage factors of the currents
        q.fnak = T(1) / (T(1) + T(0.1245) * cexp(T(-0.1) * Vv * T(FRT)) +
                         T(0.0365) * p[sigma] * cexp(-Vv * T(FRT)));
        q.e1 = cexp(Vv * T(FRT));
        q.e2 = cexp(T(2) * Vv * T(FRT));
        q.ghk1 = ghk_quot(Vv, T(FRT), q.e1);        // V / (e^{V F/RT} - 1)
        q.ghk2 = ghk_quot(Vv, T(2 * FRT), q.e2);    // V / (e^{2 V F/RT} - 1)
        q.en = cexp(T(0.35) * Vv * T(FRT));
        q.en1 = cexp(T(0.35 - 1.0) * Vv * T(FRT));
        q.kp = T(1) / (T(1) + cexp(T(7.488) - Vv / T(5.98)));
        q.rkr = T(1) / (T(1) + cexp((Vv + T(74)) / T(24)));
        return q;
    }

    template <class T> struct Currents {
        T INatjn, INatsl, IClCa, IClB, ICatjn, ICatsl, IKt;
        CG_HD T total() const { return (INatjn + INatsl) + (IClCa + IClB) + (ICatjn + ICatsl) + IKt; }
    };

    // membrane currents (Grandi_2011.H:734-921); gs: per-cell multipliers of gNa, gKr, pCa
    // (ICaL), gKs
    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        Currents<T> c;
        const T ECajn = T(0.5 / FRT) * clog(T(Cao) / y[Cajn]);
        const T ECasl = T(0.5 / FRT) * clog(T(Cao) / y[Casl]);
        const T EK = T(1.0 / FRT) * clog(T(Ko) / y[Ki]);
        const T ENajn = T(1.0 / FRT) * clog(T(Nao) / y[Najn]);
        const T ENasl = T(1.0 / FRT) * clog(T(Nao) / y[Nasl]);
        const T EKs = T(1.0 / FRT) * clog(T(Ko + 0.01833 * Nao) / (y[Ki] + T(0.01833) * y[Nai]));
        // INaK
        const T rnj = T(11) / y[Najn], rns = T(11) / y[Nasl];
        const T nak = p[IbarNaK] * q.fnak * T(Ko);
        const T INaKjn = T(Fjn) * nak / (T(1) + sq(sq(rnj))) / T(Ko + 1.5);
        const T INaKsl = T(Fsl) * nak / (T(1) + sq(sq(rns))) / T(Ko + 1.5);
        // L-type channel: Ca, K, Na through it
        const T pca = gs[2] * p[pCa];
        const T ca4 = T(4.0 * Fc * FRT);
        const T ibarcajn = pca * ca4 * (T(0.341) * y[Cajn] * q.e2 - T(0.341 * Cao)) * q.ghk2;
        const T ibarcasl = pca * ca4 * (T(0.341) * y[Casl] * q.e2 - T(0.341 * Cao)) * q.ghk2;
        const T ibark = p[pK] * T(Fc * FRT) * (T(0.75) * y[Ki] * q.e1 - T(0.75 * Ko)) * q.ghk1;
        const T ibarnajn = p[pNa] * T(Fc * FRT) * (T(0.75) * y[Najn] * q.e1 - T(0.75 * Nao)) * q.ghk1;
        const T ibarnasl = p[pNa] * T(Fc * FRT) * (T(0.75) * y[Nasl] * q.e1 - T(0.75 * Nao)) * q.ghk1;
        const T dfq = y[d] * y[f] * T(0.45);
        const T ojn = T(1) - y[fCaBjn], osl = T(1) - y[fCaBsl];
        const T ICaLCajn = T(FjnL) * ibarcajn * dfq * ojn;
        const T ICaLCasl = T(FslL) * ibarcasl * dfq * osl;
        const T ICaLK = ibark * dfq * (T(FjnL) * ojn + T(FslL) * osl);
        const T ICaLNajn = T(FjnL) * ibarnajn * dfq * ojn;
        const T ICaLNasl = T(FslL) * ibarnasl * dfq * osl;
        // chloride
        const T dvcl = Vv - p[ECl];
        c.IClB = p[gClB] * dvcl;
        c.IClCa = T(Fjn) * p[GClCa] / (T(1) + T(0.1) / y[Cajn]) * dvcl +
                  T(Fsl) * p[GClCa] / (T(1) + T(0.1) / y[Casl]) * dvcl;
        // potassium
        const T dvk = Vv - EK;
        const T k1a = T(1.02) / (T(1) + cexp(T(0.2385) * (dvk - T(59.215))));
        const T k1b = (T(0.49124) * cexp(T(0.08032) * (dvk + T(5.476))) +
                       cexp(T(0.06175) * (dvk - T(594.31)))) /
                      (T(1) + cexp(T(-0.5143) * (dvk + T(4.753))));
        const T IK1 = p[gK1] * (k1a / (k1a + k1b)) * dvk;
        const T IKp = (T(Fjn) + T(Fsl)) * p[gKp] * q.kp * dvk;
        const T IKr = gs[1] * p[gKr] * y[xr] * q.rkr * dvk;
        const T IKs = (T(Fjn) + T(Fsl)) * gs[3] * p[gKs] * sq(y[xs]) * (Vv - EKs);
        const T IKur = p[gKur] * y[kur_r] * y[kur_s] * dvk;
        const T Ito = p[gto] * y[x] * y[y_] * dvk;
        // sodium
        const T gna = gs[0] * p[gNa] * cube(y[m]) * y[h] * y[j];
        const T INajn = T(Fjn) * gna * (Vv - ENajn), INasl = T(Fsl) * gna * (Vv - ENasl);
        const T INaBjn = T(Fjn) * p[gNaB] * (Vv - ENajn), INaBsl = T(Fsl) * p[gNaB] * (Vv - ENasl);
        const T gnal = p[gNaL] * cube(y[ml]) * y[hl];
        const T INaLjn = T(Fjn) * gnal * (Vv - ENajn), INaLsl = T(Fsl) * gnal * (Vv - ENasl);
        // NCX and sarcolemmal Ca pump
        constexpr double KmCai = 0.00359, KmCao = 1.3, KmNai = 12.29, KmNao = 87.5;
        constexpr double Nao3 = Nao * Nao * Nao, KmNao3 = KmNao * KmNao * KmNao;
        const T Kajn = T(1) / (T(1) + sq(T(0.000384) / y[Cajn]));
        const T Kasl = T(1) / (T(1) + sq(T(0.000384) / y[Casl]));
        const T nj3 = cube(y[Najn]), ns3 = cube(y[Nasl]);
        const T s3jn = T(KmCai * Nao3) * (T(1) + cube(y[Najn] / T(KmNai))) +
                       T(KmNao3) * y[Cajn] * (T(1) + y[Cajn] / T(KmCai)) + T(KmCao) * nj3 +
                       nj3 * T(Cao) + T(Nao3) * y[Cajn];
        const T s3sl = T(KmCai * Nao3) * (T(1) + cube(y[Nasl] / T(KmNai))) +
                       T(KmNao3) * y[Casl] * (T(1) + y[Casl] / T(KmCai)) + T(KmCao) * ns3 +
                       ns3 * T(Cao) + T(Nao3) * y[Casl];
        const T ncx = p[IbarNCX] / (T(1) + T(0.27) * q.en1);
        const T INaCajn = T(Fjn) * ncx * Kajn * (q.en * nj3 * T(Cao) - q.en1 * T(Nao3) * y[Cajn]) / s3jn;
        const T INaCasl = T(Fsl) * ncx * Kasl * (q.en * ns3 * T(Cao) - q.en1 * T(Nao3) * y[Casl]) / s3sl;
        const T pb = cpow(T(0.0005), T(1.6));
        const T paj = cpow(y[Cajn], T(1.6)), pas = cpow(y[Casl], T(1.6));
        const T IpCajn = T(Fjn) * p[IbarSLCaP] * paj / (paj + pb);
        const T IpCasl = T(Fsl) * p[IbarSLCaP] * pas / (pas + pb);
        const T ICaBjn = T(Fjn) * p[gCaB] * (Vv - ECajn), ICaBsl = T(Fsl) * p[gCaB] * (Vv - ECasl);
        // totals per compartment (:900-921)
        c.ICatjn = ICaLCajn + ICaBjn + IpCajn - T(2) * INaCajn;
        c.ICatsl = ICaLCasl + ICaBsl + IpCasl - T(2) * INaCasl;
        c.INatjn = INajn + INaBjn + T(3) * INaCajn + T(3) * INaKjn + ICaLNajn + INaLjn;
        c.INatsl = INasl + INaBsl + T(3) * INaCasl + T(3) * INaKsl + ICaLNasl + INaLsl;
        c.IKt = Ito + IKr + IKs + IK1 - T(2) * (INaKjn + INaKsl) + ICaLK + IKp + IKur;
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy; V has a rate only when
    // solve_v (-Iion - Istim).  Every state is read before the policy writes any.
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, const T* gs, bool solve_v,
                             T istim, const VTerms<T>& q)
    {
        const Currents<T> c = currents(y[V], y, p, q, gs);
        const T iion = c.total();
        // buffers: kon C (Bmax - B) - koff B (Grandi_2011.H:824-850)
        const T cai = y[Cai], cajn = y[Cajn], casl = y[Casl], casr = y[Casr];
        const T dNaBjn = T(0.0001) * y[Najn] * (T(7.561) - y[NaBjn]) - T(0.001) * y[NaBjn];
        const T dNaBsl = T(0.0001) * y[Nasl] * (T(1.65) - y[NaBsl]) - T(0.001) * y[NaBsl];
        const T dCaM = T(34) * cai * (T(0.024) - y[CaM]) - T(0.238) * y[CaM];
        const T myo = T(0.14) - y[Myoc] - y[Myom];
        const T dMyoc = T(13.8) * cai * myo - T(0.00046) * y[Myoc];
        const T dMyom = T(0.0157 * Mgi) * myo - T(5.7e-05) * y[Myom];
        const T dSLHjn = T(100) * cajn * (T(BmaxSLHj) - y[SLHjn]) - T(0.03) * y[SLHjn];
        const T dSLHsl = T(100) * casl * (T(BmaxSLHsl) - y[SLHsl]) - T(0.03) * y[SLHsl];
        const T dSLLjn = T(100) * cajn * (T(BmaxSLLj) - y[SLLjn]) - T(1.3) * y[SLLjn];
        const T dSLLsl = T(100) * casl * (T(BmaxSLLsl) - y[SLLsl]) - T(1.3) * y[SLLsl];
        const T dSRB = T(100) * cai * (T(0.0171) - y[SRB]) - T(0.06) * y[SRB];
        const T tnch = T(0.14) - y[TnCHc] - y[TnCHm];
        const T dTnCHc = T(2.37) * cai * tnch - T(3.2e-05) * y[TnCHc];
        const T dTnCHm = T(0.003 * Mgi) * tnch - T(0.00333) * y[TnCHm];
        const T dTnCL = T(32.7) * cai * (T(0.07) - y[TnCL]) - T(koff_tncl) * y[TnCL];
        const T dCsqn = T(100) * casr * (T(BmaxCsqn) - y[Csqn]) - T(65) * y[Csqn];
        const T JCaBcyt = dTnCL + dTnCHc + dTnCHm + dCaM + dMyoc + dMyom + dSRB;
        // RyR and SR fluxes (:751-772)
        const T RI = T(1) - y[rr] - y[ro] - y[ri];
        const T jrel = p[ks] * y[ro];
        const T JSRCarel = jrel * (casr - cajn);
        const T JSRleak = (casr - cajn) * T(kleak);
        const T fi = cpow(cai / T(2.5 * 0.000246), T(1.787));
        const T fr = cpow(casr / T(1.7), T(1.787));
        const T Jserca = p[VmaxSRCaP] * (fi - fr) / (T(1) + fi + fr);
        const T kCaSR = T(15) - T(14) / (T(1) + cpow(T(0.45) / casr, T(2.5)));
        const T kiSRCa = T(0.5) * kCaSR, koSRCa = T(10) / kCaSR;
        const T cj2 = sq(cajn);
        const T dri = kiSRCa * cajn * y[ro] - T(0.005) * y[ri] - (T(0.06) * y[ri] - koSRCa * cj2 * RI);
        const T dro = koSRCa * cj2 * y[rr] - T(0.06) * y[ro] - (kiSRCa * cajn * y[ro] - T(0.005) * y[ri]);
        const T drr = T(0.005) * RI - kiSRCa * cajn * y[rr] - (koSRCa * cj2 * y[rr] - T(0.06) * y[ro]);
        // concentrations (:851-899)
        constexpr double cjn = Cmem / (Vjn * 2.0 * Fc), csl = Cmem / (Vsl * 2.0 * Fc);
        const T dCai = -Jserca * T(Vsr / Vmyo) - JCaBcyt + T(JCaslmyo / Vmyo) * (casl - cai);
        const T dNai = T(JNaslmyo / Vmyo) * (y[Nasl] - y[Nai]);
        const T rCasr = Jserca - (JSRleak * T(Vmyo / Vsr) + JSRCarel) - dCsqn;
        const T bCasr = jrel + T(kleak * Vmyo / Vsr) + T(100) * (T(BmaxCsqn) - y[Csqn]);
        const T rCajn = -c.ICatjn * T(cjn) + T(JCajnsl / Vjn) * (casl - cajn) - (dSLLjn + dSLHjn) +
                        JSRCarel * T(Vsr / Vjn) + JSRleak * T(Vmyo / Vjn);
        const T bCajn = jrel * T(Vsr / Vjn) + T(kleak * Vmyo / Vjn) + T(JCajnsl / Vjn) +
                        T(100) * ((T(BmaxSLLj) - y[SLLjn]) + (T(BmaxSLHj) - y[SLHjn]));
        const T rCasl = -c.ICatsl * T(csl) + T(JCajnsl / Vsl) * (cajn - casl) +
                        T(JCaslmyo / Vsl) * (cai - casl) - (dSLLsl + dSLHsl);
        const T bCasl = T(JCajnsl / Vsl + JCaslmyo / Vsl) +
                        T(100) * ((T(BmaxSLLsl) - y[SLLsl]) + (T(BmaxSLHsl) - y[SLHsl]));
        const T rNajn = -c.INatjn * T(Cmem / (Vjn * Fc)) + T(JNajnsl / Vjn) * (y[Nasl] - y[Najn]) - dNaBjn;
        const T rNasl = -c.INatsl * T(Cmem / (Vsl * Fc)) + T(JNajnsl / Vsl) * (y[Najn] - y[Nasl]) +
                        T(JNaslmyo / Vsl) * (y[Nai] - y[Nasl]) - dNaBsl;
        // (a, b) of the linear kinetics, all read before any write
        const T aMyoc = T(13.8) * cai * (T(0.14) - y[Myom]), aMyom = T(0.0157 * Mgi) * (T(0.14) - y[Myoc]);
        const T aTnCHc = T(2.37) * cai * (T(0.14) - y[TnCHm]), aTnCHm = T(0.003 * Mgi) * (T(0.14) - y[TnCHc]);
        const T aml = q.aml;
        // writes
        out.gate(xr, q.xrinf, q.xrtau);
        out.gate(kur_r, q.krinf, q.krtau);
        out.gate(kur_s, q.ksinf, q.kstau);
        out.gate(h, q.hjinf, q.htau);
        out.gate(j, q.hjinf, q.jtau);
        out.gate(m, q.minf, q.mtau);
        out.gate(hl, q.hlinf, T(600));
        out.lin(ml, aml, aml + q.bml, aml * (T(1) - y[ml]) - q.bml * y[ml]);
        out.gate(x, q.xinf, q.xtau);
        out.gate(y_, q.yinf, q.ytau);
        out.gate(d, q.dinf, q.dtau);
        out.gate(f, q.finf, q.ftau);
        out.gate(xs, q.xsinf, q.xstau);
        out.lin(fCaBjn, T(1.7) * cajn, T(1.7) * cajn + T(0.0119),
                T(1.7) * cajn * (T(1) - y[fCaBjn]) - T(0.0119) * y[fCaBjn]);
        out.lin(fCaBsl, T(1.7) * casl, T(1.7) * casl + T(0.0119),
                T(1.7) * casl * (T(1) - y[fCaBsl]) - T(0.0119) * y[fCaBsl]);
        out.lin(NaBjn, T(0.0001 * 7.561) * y[Najn], T(0.0001) * y[Najn] + T(0.001), dNaBjn);
        out.lin(NaBsl, T(0.0001 * 1.65) * y[Nasl], T(0.0001) * y[Nasl] + T(0.001), dNaBsl);
        out.lin(CaM, T(34.0 * 0.024) * cai, T(34) * cai + T(0.238), dCaM);
        out.lin(Myoc, aMyoc, T(13.8) * cai + T(0.00046), dMyoc);
        out.lin(Myom, aMyom, T(0.0157 * Mgi + 5.7e-05), dMyom);
        out.lin(SLHjn, T(100 * BmaxSLHj) * cajn, T(100) * cajn + T(0.03), dSLHjn);
        out.lin(SLHsl, T(100 * BmaxSLHsl) * casl, T(100) * casl + T(0.03), dSLHsl);
        out.lin(SLLjn, T(100 * BmaxSLLj) * cajn, T(100) * cajn + T(1.3), dSLLjn);
        out.lin(SLLsl, T(100 * BmaxSLLsl) * casl, T(100) * casl + T(1.3), dSLLsl);
        out.lin(SRB, T(100 * 0.0171) * cai, T(100) * cai + T(0.06), dSRB);
        out.lin(TnCHc, aTnCHc, T(2.37) * cai + T(3.2e-05), dTnCHc);
        out.lin(TnCHm, aTnCHm, T(0.003 * Mgi + 0.00333), dTnCHm);
        out.lin(TnCL, T(32.7 * 0.07) * cai, T(32.7) * cai + T(koff_tncl), dTnCL);
        out.lin(Csqn, T(100 * BmaxCsqn) * casr, T(100) * casr + T(65), dCsqn);
        out.lin(Casr, rCasr + bCasr * casr, bCasr, rCasr);
        out.lin(Cajn, rCajn + bCajn * cajn, bCajn, rCajn);
        out.lin(Casl, rCasl + bCasl * casl, bCasl, rCasl);
        out.rate(ri, dri);
        out.rate(ro, dro);
        out.rate(rr, drr);
        out.rate(Cai, dCai);
        out.rate(Nai, dNai);
        out.rate(Najn, rNajn);
        out.rate(Nasl, rNasl);
        out.rate(Ki, T(0));
        if (solve_v) out.rate(V, -iion - istim);
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
