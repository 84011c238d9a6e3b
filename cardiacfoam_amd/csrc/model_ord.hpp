// O'Hara-Rudy 2011 (41 states) on gfx950.
// Equations: src/ionicModels/ORd/ORd_2011.H:695-989; wrapper ORd.C:110-150
// (V = 1000 Vm in mV, Im = Iion_cm).  The reference accepts tissue "myocyte" only
// (ORd.C:93-96), which selects the endocardial column of every tissue switch.
//
// ORdinitConsts (ORd_2011.H:493-692) never assigns the INaK and SERCA constants
// (CONSTANTS_ is zero-filled, ORd.C:54), which makes INaK NaN in the reference; the
// published ORd-2011 values are used instead.  The K_i balance omits the stimulus, as
// the reference's derivatives() sees a zeroed Istim (ORd_2011.H:962 vs :982).
//
// Integration: 30 gates + nca (linear, xinf = anca k2n/km2n, tau = 1/km2n) are
// Rush-Larsen; CaMKt and the concentrations are forward Euler.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct ORd {
    static constexpr int NS = 41;
    static constexpr int NP = 15;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    enum { V, CaMKt, Nai, Nass, Ki, Kss, Cass, Cansr, Cajsr, Cai, m, hf, hs, j, hsp, jp, mL,
           hL, hLp, a, iF, iS, ap, iFp, iSp, d, ff, fs, fcaf, fcas, jca, ffp, fcafp, nca,
           xrf, xrs, xs1, xs2, xk1, Jrelnp, Jrelp };
    enum { GNa, GNaL, Gto, PCa, GKr, GKs, GK1, GKb, Gncx, Pnak, upScale, cmdnmax, PCab,
           PNab, GpCa };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        const double tab[NP] = {75.0, 0.0075, 0.02, 0.0001, 0.046, 0.0034, 0.1908, 0.003,
                                0.0008, 30.0, 1.0, 0.05, 2.5e-8, 3.75e-10, 0.0005};
        for (int i = 0; i < NP; ++i) p[i] = tab[i];
        const double y0[NS] = {-87.0, 0.0, 7.0, 7.0, 145.0, 145.0, 1e-4, 1.2, 1.2, 1e-4,
                               0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0,
                               0.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0,
                               1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // fixed constants of ORdinitConsts (ORd_2011.H:497-692)
    static constexpr double Fc = 96485.0, Rc = 8314.0, Tc = 310.0;
    static constexpr double cao = 1.8, ko = 5.4, nao = 140.0;
    static constexpr double L = 0.01, rad = 0.0011;
    static constexpr double Ageo = 2.0 * 3.14 * rad * rad + 2.0 * 3.14 * rad * L;
    static constexpr double vcell = 1000.0 * 3.14 * rad * rad * L;
    static constexpr double Acap = 2.0 * Ageo;
    static constexpr double vjsr = 0.0048 * vcell, vmyo = 0.68 * vcell,
                            vnsr = 0.0552 * vcell, vss = 0.02 * vcell;

    template <class T> struct VTerms {
        T hca, hna, Knai, Knao, e1, e2, vffrt, rk1, xkb, Axrf, rkr, AiF, Afcaf;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv)
    {
        VTerms<T> q;
        const T FRT = T(Fc / (Rc * Tc));
        q.hca = cexp(T(0.167) * Vv * FRT);
        q.hna = cexp(T(0.5224) * Vv * FRT);
        q.Knai = T(9.073) * cexp(T(-0.155) * Vv * FRT / T(3));
        q.Knao = T(27.78) * cexp((T(1) - T(-0.155)) * Vv * FRT / T(3));
        const T vfrt = Vv * FRT;
        q.vffrt = Vv * T(Fc) * FRT;
        q.e1 = cexp(vfrt);
        q.e2 = cexp(T(2) * vfrt);
        q.rk1 = T(1) / (T(1) + cexp((Vv + T(105.8) - T(2.6 * ko)) / T(9.493)));
        q.xkb = T(1) / (T(1) + cexp(-(Vv - T(14.48)) / T(18.34)));
        q.Axrf = T(1) / (T(1) + cexp((Vv + T(54.81)) / T(38.21)));
        q.rkr = T(1) / (T(1) + cexp((Vv + T(55)) / T(75))) * (T(1) / (T(1) + cexp((Vv - T(10)) / T(30))));
        q.AiF = T(1) / (T(1) + cexp((Vv - T(213.6)) / T(151.2)));
        q.Afcaf = T(0.3) + T(0.6) / (T(1) + cexp((Vv - T(10)) / T(10)));
        return q;
    }

    // Na/Ca exchanger of one compartment (ORd_2011.H:715-777)
    template <class T>
    CG_HD static T ncx(T na, T ca, const VTerms<T>& q, T gfrac)
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
        const T h6 = T(1) / h4;
        const T h7 = T(1) + T(nao) / kna3 * (T(1) + T(1) / q.hna);
        const T h2 = na * q.hna / (kna3 * h1);
        const T h3 = T(1) / h1;
        const T h8 = T(nao) / (kna3 * q.hna * h7);
        const T h9 = T(1) / h7;
        const T k6 = h6 * ca * kcaon;
        const T k3pp = h8 * wnaca, k4pp = h2 * wnaca;
        const T k3 = h9 * wca + k3pp, k4 = h3 * wca / q.hca + k4pp;
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
        T INa, INaL, Ito, ICaL, ICaNa, ICaK, IKr, IKs, IK1, INaCai, INaCass, INaK, INab, IKb,
          IpCa, ICab, CaMKa;
        CG_HD T total() const
        {
            return INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCai + INaCass +
                   INaK + INab + IKb + IpCa + ICab;
        }
    };

    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        // gs: per-cell multipliers of GNa, GKr, PCa, GKs
        Currents<T> c;
        const T RTF = T(Rc * Tc / Fc);
        const T FRT = T(Fc / (Rc * Tc));
        const T CaMKb = T(0.05) * (T(1) - y[CaMKt]) / (T(1) + T(0.0015) / y[Cass]);
        c.CaMKa = CaMKb + y[CaMKt];
        const T fp = T(1) / (T(1) + T(0.15) / c.CaMKa);  // fINap = fItop = fINaLp = fICaLp
        const T EK = RTF * clog(T(ko) / y[Ki]);
        const T ENa = RTF * clog(T(nao) / y[Nai]);
        const T EKs = RTF * clog((T(ko) + T(0.01833) * T(nao)) / (y[Ki] + T(0.01833) * y[Nai]));
        // INa
        const T hh = T(0.99) * y[hf] + T(0.01) * y[hs];
        const T hp = T(0.99) * y[hf] + T(0.01) * y[hsp];
        c.INa = gs[0] * p[GNa] * (Vv - ENa) * cube(y[m]) *
                ((T(1) - fp) * hh * y[j] + fp * hp * y[jp]);
        // INaL
        c.INaL = p[GNaL] * (Vv - ENa) * y[mL] * ((T(1) - fp) * y[hL] + fp * y[hLp]);
        // Ito
        const T AiS = T(1) - q.AiF;
        const T ii = q.AiF * y[iF] + AiS * y[iS];
        const T ip = q.AiF * y[iFp] + AiS * y[iSp];
        c.Ito = p[Gto] * (Vv - EK) * ((T(1) - fp) * y[a] * ii + fp * y[ap] * ip);
        // ICaL, ICaNa, ICaK (GHK)
        const T f_ = T(0.6) * y[ff] + T(0.4) * y[fs];
        const T fca = q.Afcaf * y[fcaf] + (T(1) - q.Afcaf) * y[fcas];
        const T fcap = q.Afcaf * y[fcafp] + (T(1) - q.Afcaf) * y[fcas];
        const T fp_ = T(0.6) * y[ffp] + T(0.4) * y[fs];
        const T gnp = f_ * (T(1) - y[nca]) + y[jca] * fca * y[nca];
        const T gpp = fp_ * (T(1) - y[nca]) + y[jca] * fcap * y[nca];
        T PhiCaL, PhiCaNa, PhiCaK;
        if constexpr (is_fp32<T>()) {
            const T FF = T(Fc) * FRT, g1 = ghk_quot(Vv, FRT, q.e1), g2 = ghk_quot(Vv, T(2) * FRT, q.e2);
            PhiCaL = T(4) * FF * (y[Cass] * q.e2 - T(0.341) * T(cao)) * g2;
            PhiCaNa = FF * (T(0.75) * y[Nass] * q.e1 - T(0.75) * T(nao)) * g1;
            PhiCaK = FF * (T(0.75) * y[Kss] * q.e1 - T(0.75) * T(ko)) * g1;
        } else {
            PhiCaL = T(4) * q.vffrt * (y[Cass] * q.e2 - T(0.341) * T(cao)) / (q.e2 - T(1));
            PhiCaNa = q.vffrt * (T(0.75) * y[Nass] * q.e1 - T(0.75) * T(nao)) / (q.e1 - T(1));
            PhiCaK = q.vffrt * (T(0.75) * y[Kss] * q.e1 - T(0.75) * T(ko)) / (q.e1 - T(1));
        }
        const T pca = gs[2] * p[PCa], pcap = T(1.1) * pca;
        c.ICaL = (T(1) - fp) * pca * PhiCaL * y[d] * gnp + fp * pcap * PhiCaL * y[d] * gpp;
        c.ICaNa = (T(1) - fp) * T(0.00125) * pca * PhiCaNa * y[d] * gnp +
                  fp * T(0.00125) * pcap * PhiCaNa * y[d] * gpp;
        c.ICaK = (T(1) - fp) * T(0.0003574) * pca * PhiCaK * y[d] * gnp +
                 fp * T(0.0003574) * pcap * PhiCaK * y[d] * gpp;
        // IKr, IKs, IK1, IKb
        const T xr = q.Axrf * y[xrf] + (T(1) - q.Axrf) * y[xrs];
        c.IKr = gs[1] * p[GKr] * csqrt(T(ko / 5.4)) * xr * q.rkr * (Vv - EK);
        const T KsCa = T(1) + T(0.6) / (T(1) + cpow(T(3.8e-05) / y[Cai], T(1.4)));
        c.IKs = gs[3] * p[GKs] * KsCa * y[xs1] * y[xs2] * (Vv - EKs);
        c.IK1 = p[GK1] * csqrt(T(ko)) * q.rk1 * y[xk1] * (Vv - EK);
        c.IKb = p[GKb] * q.xkb * (Vv - EK);
        // INaCa
        c.INaCai = ncx(y[Nai], y[Cai], q, T(0.8) * p[Gncx]);
        c.INaCass = ncx(y[Nass], y[Cass], q, T(0.2) * p[Gncx]);
        // INaK (published constants)
        {
            const T Kki = T(0.5), Kko = T(0.3582), MgADP = T(0.05), MgATP = T(9.8),
                    Kmgatp = T(1.698e-7), Hp = T(1.0e-7), eP = T(4.2), Khp = T(1.698e-7),
                    Knap = T(224), Kxkur = T(292);
            const T P = eP / (T(1) + Hp / Khp + y[Nai] / Knap + y[Ki] / Kxkur);
            const T dn_i = cube(T(1) + y[Nai] / q.Knai) + sq(T(1) + y[Ki] / Kki) - T(1);
            const T dn_o = cube(T(1) + T(nao) / q.Knao) + sq(T(1) + T(ko) / Kko) - T(1);
            const T a1 = T(949.5) * cube(y[Nai] / q.Knai) / dn_i;
            const T b1 = T(182.4) * MgADP, a2 = T(687.2);
            const T b2 = T(39.4) * cube(T(nao) / q.Knao) / dn_o;
            const T a3 = T(1899.0) * sq(T(ko) / Kko) / dn_o;
            const T b3 = T(79300.0) * P * Hp / (T(1) + MgATP / Kmgatp);
            const T a4 = T(639.0) * MgATP / Kmgatp / (T(1) + MgATP / Kmgatp);
            const T b4 = T(40.0) * sq(y[Ki] / Kki) / dn_i;
            const T x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
            const T x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
            const T x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
            const T x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
            const T sx = x1 + x2 + x3 + x4;
            const T JnakK = T(2) * (x4 / sx * b1 - x3 / sx * a1);
            const T JnakNa = T(3) * (x1 / sx * a3 - x2 / sx * b3);
            c.INaK = p[Pnak] * (JnakNa + JnakK);
        }
        if constexpr (is_fp32<T>()) {
            const T FF = T(Fc) * FRT;
            c.INab = p[PNab] * FF * (y[Nai] * q.e1 - T(nao)) * ghk_quot(Vv, FRT, q.e1);
            c.ICab = p[PCab] * T(4) * FF * (y[Cai] * q.e2 - T(0.341) * T(cao)) *
                     ghk_quot(Vv, T(2) * FRT, q.e2);
        } else {
            c.INab = p[PNab] * q.vffrt * (y[Nai] * q.e1 - T(nao)) / (q.e1 - T(1));
            c.ICab = p[PCab] * T(4) * q.vffrt * (y[Cai] * q.e2 - T(0.341) * T(cao)) / (q.e2 - T(1));
        }
        c.IpCa = p[GpCa] * y[Cai] / (T(0.0005) + y[Cai]);
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy (ORd_2011.H:795-989);
    // V has a rate only when solve_v (-Iion - Istim).
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, int tissue, const T* gs,
                             bool solve_v, T istim, const VTerms<T>& q)
    {
        (void)tissue;
        const T Vv = y[V];
        const Currents<T> c = currents(Vv, y, p, q, gs);
        const T Iion = c.total();
        const T CaMKb = c.CaMKa - y[CaMKt];
        const T fp = T(1) / (T(1) + T(0.15) / c.CaMKa);
        // fluxes
        const T Jdiff = (y[Cass] - y[Cai]) / T(0.2);
        const T JdiffK = (y[Kss] - y[Ki]) / T(2);
        const T JdiffNa = (y[Nass] - y[Nai]) / T(2);
        const T Jtr = (y[Cansr] - y[Cajsr]) / T(100);
        const T Jleak = T(0.0039375) * y[Cansr] / T(15);
        const T Jupnp = p[upScale] * T(0.004375) * y[Cai] / (y[Cai] + T(0.00092));
        const T Jupp = p[upScale] * T(2.75) * T(0.004375) * y[Cai] / (y[Cai] + T(0.00092) - T(0.00017));
        const T Jup = (T(1) - fp) * Jupnp + fp * Jupp - Jleak;
        const T Jrel = (T(1) - fp) * y[Jrelnp] + fp * y[Jrelp];
        // buffers
        const T Bcajsr = T(1) / (T(1) + T(10.0 * 0.8) / sq(T(0.8) + y[Cajsr]));
        const T Bcass = T(1) / (T(1) + T(0.047 * 0.00087) / sq(T(0.00087) + y[Cass]) +
                                T(1.124 * 0.0087) / sq(T(0.0087) + y[Cass]));
        const T Bcai = T(1) / (T(1) + p[cmdnmax] * T(0.00238) / sq(T(0.00238) + y[Cai]) +
                               T(0.07 * 0.0005) / sq(T(0.0005) + y[Cai]));
        const T AF = T(Acap / Fc);
        const T dCaMKt = T(0.05) * CaMKb * (CaMKb + y[CaMKt]) - T(0.00068) * y[CaMKt];
        const T dNai = -(c.INa + c.INaL + T(3) * c.INaCai + T(3) * c.INaK + c.INab) * AF / T(vmyo) +
                       JdiffNa * T(vss / vmyo);
        const T dNass = -(c.ICaNa + T(3) * c.INaCass) * AF / T(vss) - JdiffNa;
        const T dKi = -(c.Ito + c.IKr + c.IKs + c.IK1 + c.IKb - T(2) * c.INaK) * AF / T(vmyo) +
                      JdiffK * T(vss / vmyo);
        const T dKss = -c.ICaK * AF / T(vss) - JdiffK;
        const T dCai = Bcai * (-(c.IpCa + c.ICab - T(2) * c.INaCai) * AF / T(2 * vmyo) -
                               Jup * T(vnsr / vmyo) + Jdiff * T(vss / vmyo));
        const T dCass = Bcass * (-(c.ICaL - T(2) * c.INaCass) * AF / T(2 * vss) +
                                 Jrel * T(vjsr / vss) - Jdiff);
        const T dCansr = Jup - Jtr * T(vjsr / vnsr);
        const T dCajsr = Bcajsr * (Jtr - Jrel);
        // state-dependent gate targets at the step start
        const T anca = T(1) / (T(1000) / y[jca] + sq(sq(T(1) + T(0.002) / y[Cass])));
        const T km2n = y[jca];
        T trel = T(4.75) / (T(1) + T(0.0123) / y[Cajsr]);
        trel = trel < T(0.001) ? T(0.001) : trel;
        T trelp = T(1.25 * 4.75) / (T(1) + T(0.0123) / y[Cajsr]);
        trelp = trelp < T(0.001) ? T(0.001) : trelp;
        const T relf = T(1) / (T(1) + T(1) * cpow(T(1.5) / y[Cajsr], T(8)));
        const T Jrel_inf = T(0.5 * 4.75) * -c.ICaL * relf;
        const T Jrel_infp = T(0.5 * 1.25 * 4.75) * -c.ICaL * relf;

        // voltage-gated RL (ORd_2011.H:830-930)
        {
            const T txk1 = T(122.2) / (cexp(-(Vv + T(127.2)) / T(20.36)) + cexp((Vv + T(236.8)) / T(69.33)));
            const T xk1ss = T(1) / (T(1) + cexp(-(Vv + T(2.5538 * ko + 144.59)) / T(1.5692 * ko + 3.8115)));
            out.gate(xk1, xk1ss, txk1);
            const T txrf = T(12.98) + T(1) / (T(0.3652) * cexp((Vv - T(31.66)) / T(3.869)) +
                                              T(4.123e-05) * cexp(-(Vv - T(47.78)) / T(20.38)));
            const T txrs = T(1.865) + T(1) / (T(0.06629) * cexp((Vv - T(34.7)) / T(7.355)) +
                                              T(1.128e-05) * cexp(-(Vv - T(29.74)) / T(25.94)));
            const T xrss = T(1) / (T(1) + cexp(-(Vv + T(8.337)) / T(6.789)));
            out.gate(xrf, xrss, txrf);
            out.gate(xrs, xrss, txrs);
            const T txs1 = T(817.3) + T(1) / (T(0.0002326) * cexp((Vv + T(48.28)) / T(17.8)) +
                                              T(0.001292) * cexp(-(Vv + T(210)) / T(230)));
            const T txs2 = T(1) / (T(0.01) * cexp((Vv - T(50)) / T(20)) + T(0.0193) * cexp(-(Vv + T(66.54)) / T(31)));
            const T xs1ss = T(1) / (T(1) + cexp(-(Vv + T(11.6)) / T(8.932)));
            out.gate(xs1, xs1ss, txs1);
            out.gate(xs2, xs1ss, txs2);
            const T hss = T(1) / (T(1) + cexp((Vv + T(82.9)) / T(6.086)));
            const T hssp = T(1) / (T(1) + cexp((Vv + T(89.1)) / T(6.086)));
            const T thf = T(1) / (T(1.432e-05) * cexp(-(Vv + T(1.196)) / T(6.285)) + T(6.149) * cexp((Vv + T(0.5096)) / T(20.27)));
            const T ths = T(1) / (T(0.009794) * cexp(-(Vv + T(17.95)) / T(28.05)) + T(0.3343) * cexp((Vv + T(5.73)) / T(56.66)));
            const T tj = T(2.038) + T(1) / (T(0.02136) * cexp(-(Vv + T(100.6)) / T(8.281)) + T(0.3052) * cexp((Vv + T(0.9941)) / T(38.45)));
            const T mss = T(1) / (T(1) + cexp(-(Vv + T(39.57)) / T(9.871)));
            const T tm = T(1) / (T(6.765) * cexp((Vv + T(11.64)) / T(34.77)) + T(8.552) * cexp(-(Vv + T(77.42)) / T(5.955)));
            out.gate(hf, hss, thf);
            out.gate(hs, hss, ths);
            out.gate(hsp, hssp, T(3) * ths);
            out.gate(m, mss, tm);
            out.gate(j, hss, tj);
            out.gate(jp, hss, T(1.46) * tj);
            const T ass = T(1) / (T(1) + cexp(-(Vv - T(14.34)) / T(14.82)));
            const T assp = T(1) / (T(1) + cexp(-(Vv - T(24.34)) / T(14.82)));
            const T dti_dev = T(1.354) + T(0.0001) / (cexp((Vv - T(167.4)) / T(15.89)) + cexp(-(Vv - T(12.23)) / T(0.2154)));
            const T dti_rec = T(1) - T(0.5) / (T(1) + cexp((Vv + T(70)) / T(20)));
            const T iss = T(1) / (T(1) + cexp((Vv + T(43.94)) / T(5.711)));
            const T ta_ = T(1.0515) / (T(1) / (T(1.2089) * (T(1) + cexp(-(Vv - T(18.4099)) / T(29.3814)))) +
                                       T(3.5) / (T(1) + cexp((Vv + T(100)) / T(29.3814))));
            const T tiF = T(4.562) + T(1) / (T(0.3933) * cexp(-(Vv + T(100)) / T(100)) + T(0.08004) * cexp((Vv + T(50)) / T(16.59)));
            const T tiS = T(23.62) + T(1) / (T(0.001416) * cexp(-(Vv + T(96.52)) / T(59.05)) + T(1.78e-08) * cexp((Vv + T(114.1)) / T(8.079)));
            out.gate(a, ass, ta_);
            out.gate(ap, assp, ta_);
            out.gate(iF, iss, tiF);
            out.gate(iS, iss, tiS);
            out.gate(iFp, iss, dti_dev * dti_rec * tiF);
            out.gate(iSp, iss, dti_dev * dti_rec * tiS);
            const T hLss = T(1) / (T(1) + cexp((Vv + T(87.61)) / T(7.488)));
            const T hLssp = T(1) / (T(1) + cexp((Vv + T(93.81)) / T(7.488)));
            const T mLss = T(1) / (T(1) + cexp(-(Vv + T(42.85)) / T(5.264)));
            out.gate(hL, hLss, T(200));
            out.gate(mL, mLss, tm);
            out.gate(hLp, hLssp, T(600));
            const T dss = T(1) / (T(1) + cexp(-(Vv + T(3.94)) / T(4.23)));
            const T fss = T(1) / (T(1) + cexp((Vv + T(19.58)) / T(3.696)));
            const T td = T(0.6) + T(1) / (cexp(T(-0.05) * (Vv + T(6))) + cexp(T(0.09) * (Vv + T(14))));
            const T tfcaf = T(7) + T(1) / (T(0.04) * cexp(-(Vv - T(4)) / T(7)) + T(0.04) * cexp((Vv - T(4)) / T(7)));
            const T tfcas = T(100) + T(1) / (T(0.00012) * cexp(-Vv / T(3)) + T(0.00012) * cexp(Vv / T(7)));
            const T tff = T(7) + T(1) / (T(0.0045) * cexp(-(Vv + T(20)) / T(10)) + T(0.0045) * cexp((Vv + T(20)) / T(10)));
            const T tfs = T(1000) + T(1) / (T(3.5e-05) * cexp(-(Vv + T(5)) / T(4)) + T(3.5e-05) * cexp((Vv + T(5)) / T(6)));
            out.gate(d, dss, td);
            out.gate(ff, fss, tff);
            out.gate(fs, fss, tfs);
            out.gate(fcaf, fss, tfcaf);
            out.gate(fcafp, fss, T(2.5) * tfcaf);
            out.gate(fcas, fss, tfcas);
            out.gate(ffp, fss, T(2.5) * tff);
            out.gate(jca, fss, T(75));
        }
        out.lin(nca, anca * T(1000), km2n);  // dn/dt = anca k2n - n km2n
        out.gate(Jrelnp, Jrel_inf, trel);
        out.gate(Jrelp, Jrel_infp, trelp);
        out.rate(CaMKt, dCaMKt);
        out.rate(Nai, dNai);
        out.rate(Nass, dNass);
        out.rate(Ki, dKi);
        out.rate(Kss, dKss);
        out.rate(Cai, dCai);
        out.rate(Cass, dCass);
        out.rate(Cansr, dCansr);
        out.rate(Cajsr, dCajsr);
        if (solve_v) out.rate(V, -Iion - istim);
    }

    // one RL/FE step of dt in place
    template <class T>
    CG_HD static void step(T* y, T dt, const T* p, int tissue, const T* gs, bool solve_v,
                            T istim, const VTerms<T>& q)
    {
        RlStep<T> out{y, dt};
        rates(y, out, p, tissue, gs, solve_v, istim, q);
    }

    // dy/dt at y (dV/dt = 0 unless V is integrated with the other states)
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int tissue, const T* gs, bool solve_v,
                             T istim, const VTerms<T>& q)
    {
        Deriv<T> out{y, d};
        d[V] = T(0);
        rates(y, out, p, tissue, gs, solve_v, istim, q);
    }
};

} // namespace cg
