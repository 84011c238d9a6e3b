// ToR-ORd with dynamic chloride (Tomek et al. 2019, 2023 revision; 45 states) on gfx950.
// Equations: src/ionicModels/ToRORd_dynCl/ToRORd_dynCl_2023.H:848-1177 (rates), :571-846
// (constants and the endo / M / epi initial states); wrapper ToRORd_dynCl.C:106-155
// (V = 1000 Vm in mV, Im = Iion_cm).  Tissues epicardialCells, mCells, endocardialCells
// (ToRORd_dynCl.C:93-96).  The K_i balance omits the stimulus, as the reference's
// derivatives() sees a zeroed Istim (:1119 reads it before :1172 assigns it).  I_katp is
// 0 (fkatp = 0 in initConsts) and is kept as the reference writes it.
//
// Integration (Rush-Larsen): the Hodgkin-Huxley gates and the RyR release fluxes as gates;
// nca_i, nca_ss and the five IKr Markov states as linear in their own state with the
// others frozen over the step; CaMKt and the concentrations forward Euler.  V-only terms
// are evaluated once per step.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct ToRORd {
    static constexpr int NS = 45;
    static constexpr int NP = 25;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    // STATES order of ToRORd_dynCl_2023Names.H
    enum { V, CaMKt, Nai, Nass, Ki, Kss, Cass, Cansr, Cajsr, Cai, Cli, Clss, m, h, j, hp, jp,
           mL, hL, hLp, a, iF, iS, ap, iFp, iSp, d, ff, fs, fcaf, fcas, jca, ffp, fcafp, nca_ss,
           nca_i, C1, C2, C3, I, O, xs1, xs2, Jrelnp, Jrelp };
    // tissue-dependent and rescalable constants of initConsts; constA and the extracellular
    // activity coefficients gamma_cao, gamma_ko (= gamma_nao), which need pow/exp
    enum { GNa, GKr, PCa, GKs, GK1, GKb, GNaL, Gto, Gncx, Pnak, cmdnmax, upScale, GpCa, GClCa,
           GClb, PCab, PNab, Jrel_b, Jup_b, bt, gkatp, fkatp, constA, gcao, gko };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue < 1 || tissue > 3) return false;
        const bool epi = tissue == 1, mid = tissue == 2;
        const double PCa_b = 8.37570000000000046e-05;
        p[GNa] = 11.7802;
        p[GKr] = epi ? 0.0321 * 1.3 : mid ? 0.0321 * 0.8 : 0.0321;
        p[PCa] = epi ? PCa_b * 1.2 : mid ? PCa_b * 2.0 : PCa_b;
        p[GKs] = epi ? 0.0011 * 1.4 : 0.0011;
        p[GK1] = epi ? 0.6992 * 1.2 : mid ? 0.6992 * 1.3 : 0.6992;
        p[GKb] = epi ? 0.0189 * 0.6 : 0.0189;
        p[GNaL] = epi ? 0.0279 * 0.6 : 0.0279;
        p[Gto] = (epi || mid) ? 0.16 * 2.0 : 0.16;
        p[Gncx] = epi ? 0.0034 * 1.1 : mid ? 0.0034 * 1.4 : 0.0034;
        p[Pnak] = epi ? 15.4509 * 0.9 : mid ? 15.4509 * 0.7 : 15.4509;
        p[cmdnmax] = epi ? 0.05 * 1.3 : 0.05;
        p[upScale] = epi ? 1.3 : 1.0;
        p[GpCa] = 0.0005;
        p[GClCa] = 0.2843;
        p[GClb] = 0.00198;
        p[PCab] = 5.91940000000000005e-08;
        p[PNab] = 1.9239e-09;
        p[Jrel_b] = 1.5378;
        p[Jup_b] = 1.0;
        p[bt] = 4.75;
        p[gkatp] = 4.3195;
        p[fkatp] = 0.0;
        // :736-744, with Io = 0.5 (nao + ko + clo + 4 cao) / 1000
        p[constA] = 1820000.0 * pow(74.0 * Tc, -1.5);
        const double Io = 0.5 * (nao + ko + clo + 4.0 * cao) / 1000.0;
        const double lgo = sqrt(Io) / (1.0 + sqrt(Io)) - 0.3 * Io;
        p[gcao] = exp(-p[constA] * 4.0 * lgo);
        p[gko] = exp(-p[constA] * 1.0 * lgo);
        // initial states (:782-846), columns endo / M / epi
        static const double y0[3][NS] = {
            {-89.74808, 1.10e-02, 12.39736, 12.3977, 147.7115, 147.7114, 6.50e-05, 1.528001,
             1.525693, 7.45e-05, 29.20698, 29.20696, 6.52e-04, 0.8473267, 0.8471657, 0.7018454,
             0.8469014, 1.35e-04, 0.5566017, 0.3115491, 8.90e-04, 0.9996716, 0.5988908,
             4.53e-04, 0.9996716, 0.6620692, 1.59e-31, 1.0, 0.9401791, 1.0, 0.9999014,
             0.9999846, 1.0, 1.0, 4.90e-04, 8.33e-04, 0.9982511, 7.94e-04, 6.53e-04,
             9.80e-06, 2.92e-04, 0.243959, 1.59e-04, 1.81e-22, 4.36e-21},
            {-91.33918, 2.01882e-02, 15.94867, 15.94922, 156.7131, 156.7130, 6.64e-05,
             2.012225, 2.016415, 8.30e-05, 48.91277, 48.91274, 4.62e-04, 0.8739077, 0.8737841,
             0.7478972, 0.8735375, 9.99e-05, 0.5986118, 0.3339899, 7.99e-04, 0.9997514,
             0.5702538, 4.07e-04, 0.9997514, 0.6351927, -8.33e-30, 1.0, 0.9183587, 1.0,
             0.9997540, 0.9999743, 1.0, 1.0, 5.34e-04, 1.26e-03, 0.9983451, 7.09e-04,
             5.91e-04, 1.06e-05, 3.45e-04, 0.2642293, 1.33e-04, -1.30e-21, -7.61e-20},
            {-90.74563, 1.27e-02, 13.40062, 13.40094, 152.3639, 152.3638, 5.75e-05, 1.806794,
             1.805047, 6.62e-05, 34.31721, 34.31719, 5.25e-04, 0.8645148, 0.8644571, 0.7313656,
             0.8643527, 1.12e-04, 0.5916536, 0.3476812, 8.32e-04, 0.9997242, 0.9997235,
             4.24e-04, 0.9997242, 0.9997241, -2.49e-36, 1.0, 0.9510602, 1.0, 0.9999377,
             0.9999886, 1.0, 1.0, 3.05e-04, 5.27e-04, 0.9984733, 7.39e-04, 6.03e-04,
             5.68e-06, 1.79e-04, 0.2233584, 1.42e-04, 6.78e-25, -1.58e-23}};
        const double* src = y0[tissue == 3 ? 0 : mid ? 1 : 2];
        for (int i = 0; i < NS; ++i) y[i] = src[i];
        return true;
    }

    // fixed constants of initConsts (:573-770)
    static constexpr double Fc = 96485.0, Rc = 8314.0, Tc = 310.0;
    static constexpr double cao = 1.8, clo = 150.0, ko = 5.0, nao = 140.0;
    static constexpr double L = 0.01, rad = 0.0011;
    static constexpr double Ageo = 2.0 * 3.14 * rad * rad + 2.0 * 3.14 * rad * L;
    static constexpr double vcell = 1000.0 * 3.14 * rad * rad * L;
    static constexpr double Acap = 2.0 * Ageo;
    static constexpr double vjsr = 0.0048 * vcell, vmyo = 0.68 * vcell,
                            vnsr = 0.0552 * vcell, vss = 0.02 * vcell;

    template <class T> struct VTerms {
        T th, tj, hss, hssp, mss, tm, hLss, hLssp, mLss;        // INa, INaL
        T AiF, ass, assp, iss, tau_a, tiF, tiS, dti, dep;        // Ito (dep: epicardial factor)
        T Afcaf, dss, fss, jcass, tff, tfs, tfcaf, tfcas, td;   // ICaL
        T xs1ss, txs1, txs2, xkb;                                // IKs, IKb
        T al, be, al2, be2, ali, bei, alCI, beIC;                // IKr Markov rates
        T hca, hna, Knai, Knao, e1, e2, vffrt, FRT;              // NCX, NaK, GHK
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T*)
    {
        VTerms<T> q;
        // INa (:909-930)
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
        q.th = T(1) / (ah + bh);
        q.tj = T(1) / (aj + bj);
        q.hss = T(1) / sq(T(1) + cexp((Vv + T(71.55)) / T(7.43)));
        q.hssp = T(1) / sq(T(1) + cexp((Vv + T(77.55)) / T(7.43)));
        q.mss = T(1) / sq(T(1) + cexp(-(Vv + T(56.86)) / T(9.03)));
        q.tm = T(0.1292) * cexp(-sq((Vv + T(45.79)) / T(15.54))) +
               T(0.06487) * cexp(-sq((Vv - T(4.823)) / T(51.12)));
        q.hLss = T(1) / (T(1) + cexp((Vv + T(87.61)) / T(7.488)));
        q.hLssp = T(1) / (T(1) + cexp((Vv + T(93.81)) / T(7.488)));
        q.mLss = T(1) / (T(1) + cexp(-(Vv + T(42.85)) / T(5.264)));
        // Ito (:946-973)
        q.AiF = T(1) / (T(1) + cexp((Vv - T(213.6)) / T(151.2)));
        q.ass = T(1) / (T(1) + cexp(-(Vv - T(14.34)) / T(14.82)));
        q.assp = T(1) / (T(1) + cexp(-(Vv - T(24.34)) / T(14.82)));
        q.iss = T(1) / (T(1) + cexp((Vv + T(43.94)) / T(5.711)));
        q.tau_a = T(1.0515) / (T(1) / (T(1.2089) * (T(1) + cexp(-(Vv - T(18.4099)) / T(29.3814)))) +
                               T(3.5) / (T(1) + cexp((Vv + T(100)) / T(29.3814))));
        q.tiF = T(4.562) + T(1) / (T(0.3933) * cexp(-(Vv + T(100)) / T(100)) +
                                   T(0.08004) * cexp((Vv + T(50)) / T(16.59)));
        q.tiS = T(23.62) + T(1) / (T(0.001416) * cexp(-(Vv + T(96.52)) / T(59.05)) +
                                   T(1.78e-08) * cexp((Vv + T(114.1)) / T(8.079)));
        q.dti = (T(1.354) + T(0.0001) / (cexp((Vv - T(167.4)) / T(15.89)) +
                                         cexp(-(Vv - T(12.23)) / T(0.2154)))) *
                (T(1) - T(0.5) / (T(1) + cexp((Vv + T(70)) / T(20))));
        q.dep = T(1) - T(0.95) / (T(1) + cexp((Vv + T(70)) / T(5)));
        // ICaL gates (:975-1008)
        q.Afcaf = T(0.3) + T(0.6) / (T(1) + cexp((Vv - T(10)) / T(10)));
        q.dss = Vv >= T(31.4978) ? T(1) : T(1.0763) * cexp(T(-1.007) * cexp(T(-0.0829) * Vv));
        q.fss = T(1) / (T(1) + cexp((Vv + T(19.58)) / T(3.696)));
        q.jcass = T(1) / (T(1) + cexp((Vv + T(18.08)) / T(2.7916)));
        q.tfcaf = T(7) + T(1) / (T(0.04) * cexp(-(Vv - T(4)) / T(7)) + T(0.04) * cexp((Vv - T(4)) / T(7)));
        q.tfcas = T(100) + T(1) / (T(0.00012) * cexp(-Vv / T(3)) + T(0.00012) * cexp(Vv / T(7)));
        q.tff = T(7) + T(1) / (T(0.0045) * cexp(-(Vv + T(20)) / T(10)) + T(0.0045) * cexp((Vv + T(20)) / T(10)));
        q.tfs = T(1000) + T(1) / (T(3.5e-05) * cexp(-(Vv + T(5)) / T(4)) + T(3.5e-05) * cexp((Vv + T(5)) / T(6)));
        q.td = T(0.6) + T(1) / (cexp(T(-0.05) * (Vv + T(6))) + cexp(T(0.09) * (Vv + T(14))));
        // IKs, IKb (:897-906)
        q.xs1ss = T(1) / (T(1) + cexp(-(Vv + T(11.6)) / T(8.932)));
        q.txs1 = T(817.3) + T(1) / (T(0.0002326) * cexp((Vv + T(48.28)) / T(17.8)) +
                                    T(0.001292) * cexp(-(Vv + T(210)) / T(230)));
        q.txs2 = T(1) / (T(0.01) * cexp((Vv - T(50)) / T(20)) + T(0.0193) * cexp(-(Vv + T(66.54)) / T(31)));
        q.xkb = T(1) / (T(1) + cexp(-(Vv - T(10.8968)) / T(23.9871)));
        // IKr Markov rates and the exchanger / pump voltage factors (:1050-1066)
        q.FRT = T(Fc / (Rc * Tc));
        const T vfrt = Vv * q.FRT;
        q.vffrt = Vv * T(Fc) * q.FRT;
        q.al = T(0.1161) * cexp(T(0.299) * vfrt);
        q.be = T(0.2442) * cexp(T(-1.604) * vfrt);
        q.al2 = T(0.0578) * cexp(T(0.971) * vfrt);
        q.be2 = T(0.000349) * cexp(T(-1.062) * vfrt);
        q.ali = T(0.2533) * cexp(T(0.5953) * vfrt);
        q.bei = T(0.06525) * cexp(T(-0.8209) * vfrt);
        q.alCI = T(5.2e-05) * cexp(T(1.525) * vfrt);
        q.beIC = q.be2 * q.bei * q.alCI / (q.al2 * q.ali);
        q.hca = cexp(T(0.167) * vfrt);
        q.hna = cexp(T(0.5224) * vfrt);
        q.Knai = T(9.073) * cexp(T(-0.155) * vfrt / T(3));
        q.Knao = T(27.78) * cexp(T(1.0 - -0.155) * vfrt / T(3));
        q.e1 = cexp(vfrt);
        q.e2 = cexp(T(2) * vfrt);
        return q;
    }

    // Na/Ca exchanger of one compartment (:1013-1023, :1065-1166), gfrac = its share of Gncx
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
        T INa, INaL, Ito, ICaLi, ICaLss, ICaNai, ICaNass, ICaKi, ICaKss, IKr, IKs, IK1, INaCai,
          INaCass, INaK, INab, IKb, IpCa, ICab, IClCaj, IClCasl, IClb, Ikatp, fp;
        CG_HD T total() const
        {
            return INa + INaL + Ito + (ICaLss + ICaLi) + (ICaNass + ICaNai) + (ICaKss + ICaKi) +
                   IKr + IKs + IK1 + INaCai + INaCass + INaK + INab + IKb + IpCa + ICab +
                   (IClCaj + IClCasl) + IClb + Ikatp;
        }
    };

    // the ionic currents (:850-1062); gs: per-cell multipliers of GNa, GKr, PCa, GKs
    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        Currents<T> c;
        const T RTF = T(Rc * Tc / Fc);
        const T CaMKb = T(0.05) * (T(1) - y[CaMKt]) / (T(1) + T(0.0015) / y[Cass]);
        const T fp = T(1) / (T(1) + T(0.15) / (CaMKb + y[CaMKt]));
        c.fp = fp;
        const T EK = RTF * clog(T(ko) / y[Ki]);
        const T ENa = RTF * clog(T(nao) / y[Nai]);
        const T EKs = RTF * clog(T(ko + 0.01833 * nao) / (y[Ki] + T(0.01833) * y[Nai]));
        const T ECl = T(-Rc * Tc / Fc) * clog(T(clo) / y[Cli]);
        const T EClss = T(-Rc * Tc / Fc) * clog(T(clo) / y[Clss]);
        const T dvk = Vv - EK;
        // INa, INaL
        c.INa = gs[0] * p[GNa] * (Vv - ENa) * cube(y[m]) *
                ((T(1) - fp) * y[h] * y[j] + fp * y[hp] * y[jp]);
        c.INaL = p[GNaL] * (Vv - ENa) * y[mL] * ((T(1) - fp) * y[hL] + fp * y[hLp]);
        // Ito
        const T AiS = T(1) - q.AiF;
        const T ii = q.AiF * y[iF] + AiS * y[iS];
        const T ip = q.AiF * y[iFp] + AiS * y[iSp];
        c.Ito = p[Gto] * dvk * ((T(1) - fp) * y[a] * ii + fp * y[ap] * ip);
        // chloride
        c.IClCaj = p[GClCa] / (T(1) + T(0.1) / y[Cass]) * (Vv - EClss);
        c.IClCasl = T(0) * p[GClCa] / (T(1) + T(0.1) / y[Cai]) * (Vv - ECl);   // 1 - Fjunc = 0
        c.IClb = p[GClb] * (Vv - ECl);
        // K currents
        const T aK1 = T(4.094) / (T(1) + cexp(T(0.1217) * (dvk - T(49.934))));
        const T bK1 = (T(15.72) * cexp(T(0.0674) * (dvk - T(3.257))) + cexp(T(0.0618) * (dvk - T(594.31)))) /
                      (T(1) + cexp(T(-0.1629) * (dvk + T(14.207))));
        c.IK1 = p[GK1] * (aK1 / (aK1 + bK1)) * dvk;                 // sqrt(ko / 5) = 1
        c.IKb = p[GKb] * q.xkb * dvk;
        const T KsCa = T(1) + T(0.6) / (T(1) + cpow(T(3.8e-05) / y[Cai], T(1.4)));
        c.IKs = gs[3] * p[GKs] * KsCa * y[xs1] * y[xs2] * (Vv - EKs);
        c.IKr = gs[1] * p[GKr] * y[O] * dvk;                        // sqrt(ko / 5) = 1
        c.Ikatp = p[fkatp] * p[gkatp] * T(1.0 / 65.0) * dvk;       // (ko/5)^0.24 = 1
        // activity coefficients of the GHK fluxes (:994-1000)
        const T Ii = T(0.5) * (y[Nai] + y[Ki] + y[Cli] + T(4) * y[Cai]) / T(1000);
        const T Iss = T(0.5) * (y[Nass] + y[Kss] + y[Clss] + T(4) * y[Cass]) / T(1000);
        const T si = csqrt(Ii), sss = csqrt(Iss);
        const T lgi = si / (T(1) + si) - T(0.3) * Ii, lgss = sss / (T(1) + sss) - T(0.3) * Iss;
        const T g1i = cexp(-p[constA] * lgi), g1ss = cexp(-p[constA] * lgss);
        const T g2i = cexp(-p[constA] * T(4) * lgi), g2ss = cexp(-p[constA] * T(4) * lgss);
        const T g2o = p[gcao], g1o = p[gko];
        // L-type channel; 0.8 of its flux faces the subspace
        const T f_ = T(0.6) * y[ff] + T(0.4) * y[fs];
        const T fca = q.Afcaf * y[fcaf] + (T(1) - q.Afcaf) * y[fcas];
        const T fcap = q.Afcaf * y[fcafp] + (T(1) - q.Afcaf) * y[fcas];
        const T fp_ = T(0.6) * y[ffp] + T(0.4) * y[fs];
        const T gi = f_ * (T(1) - y[nca_i]) + y[jca] * fca * y[nca_i];
        const T gpi = fp_ * (T(1) - y[nca_i]) + y[jca] * fcap * y[nca_i];
        const T gss = f_ * (T(1) - y[nca_ss]) + y[jca] * fca * y[nca_ss];
        const T gpss = fp_ * (T(1) - y[nca_ss]) + y[jca] * fcap * y[nca_ss];
        T q2, q1;   // V / (e^{2 V F/RT} - 1) and V / (e^{V F/RT} - 1), times F F/RT
        const T FF = T(Fc) * q.FRT;
        if constexpr (is_fp32<T>()) {
            q2 = FF * ghk_quot(Vv, T(2) * q.FRT, q.e2);
            q1 = FF * ghk_quot(Vv, q.FRT, q.e1);
        } else {
            q2 = q.vffrt / (q.e2 - T(1));
            q1 = q.vffrt / (q.e1 - T(1));
        }
        const T PhiCaLi = T(4) * (g2i * y[Cai] * q.e2 - g2o * T(cao)) * q2;
        const T PhiCaLss = T(4) * (g2ss * y[Cass] * q.e2 - g2o * T(cao)) * q2;
        const T PhiCaKi = (g1i * y[Ki] * q.e1 - g1o * T(ko)) * q1;
        const T PhiCaKss = (g1ss * y[Kss] * q.e1 - g1o * T(ko)) * q1;
        const T PhiCaNai = (g1i * y[Nai] * q.e1 - g1o * T(nao)) * q1;
        const T PhiCaNass = (g1ss * y[Nass] * q.e1 - g1o * T(nao)) * q1;
        const T pca = gs[2] * p[PCa], pcap = T(1.1) * pca;
        const T dd = y[d], fi = T(1.0 - 0.8), fs_ = T(0.8);
        c.ICaLi = fi * ((T(1) - fp) * pca * PhiCaLi * dd * gi + fp * pcap * PhiCaLi * dd * gpi);
        c.ICaLss = fs_ * ((T(1) - fp) * pca * PhiCaLss * dd * gss + fp * pcap * PhiCaLss * dd * gpss);
        c.ICaKi = fi * ((T(1) - fp) * T(0.0003574) * pca * PhiCaKi * dd * gi +
                        fp * T(0.0003574) * pcap * PhiCaKi * dd * gpi);
        c.ICaKss = fs_ * ((T(1) - fp) * T(0.0003574) * pca * PhiCaKss * dd * gss +
                          fp * T(0.0003574) * pcap * PhiCaKss * dd * gpss);
        c.ICaNai = fi * ((T(1) - fp) * T(0.00125) * pca * PhiCaNai * dd * gi +
                         fp * T(0.00125) * pcap * PhiCaNai * dd * gpi);
        c.ICaNass = fs_ * ((T(1) - fp) * T(0.00125) * pca * PhiCaNass * dd * gss +
                           fp * T(0.00125) * pcap * PhiCaNass * dd * gpss);
        c.ICab = p[PCab] * T(4) * (g2i * y[Cai] * q.e2 - g2o * T(cao)) * q2;
        c.INab = p[PNab] * (y[Nai] * q.e1 - T(nao)) * q1;
        c.IpCa = p[GpCa] * y[Cai] / (T(0.0005) + y[Cai]);
        // exchanger (0.35 in the subspace) and pump
        c.INaCai = ncx(y[Nai], y[Cai], q, T(1.0 - 0.35) * p[Gncx]);
        c.INaCass = ncx(y[Nass], y[Cass], q, T(0.35) * p[Gncx]);
        {
            const T Kki = T(0.5), Kko = T(0.3582);
            const T P = T(4.2) / (T(1) + T(1e-07 / 1.698e-07) + y[Nai] / T(224) + y[Ki] / T(292));
            const T dn_i = cube(T(1) + y[Nai] / q.Knai) + sq(T(1) + y[Ki] / Kki) - T(1);
            const T dn_o = cube(T(1) + T(nao) / q.Knao) + sq(T(1) + T(ko) / Kko) - T(1);
            const T a1 = T(949.5) * cube(y[Nai] / q.Knai) / dn_i;
            const T b1 = T(182.4 * 0.05), a2 = T(687.2);
            const T b2 = T(39.4) * cube(T(nao) / q.Knao) / dn_o;
            const T a3 = T(1899.0) * sq(T(ko) / Kko) / dn_o;
            const T b3 = T(79300.0) * P * T(1e-07) / T(1.0 + 9.8 / 1.698e-07);
            const T a4 = T(639.0 * 9.8 / 1.698e-07 / (1.0 + 9.8 / 1.698e-07));
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
        return c;
    }

    // Rates at the state y (y[V] in mV) through an update policy; V has a rate only when
    // solve_v (-Iion - Istim).  Every state is read before the policy writes any.
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, int tissue, const T* gs,
                             bool solve_v, T istim, const VTerms<T>& q)
    {
        const Currents<T> c = currents(y[V], y, p, q, gs);
        const T Iion = c.total();
        const T fp = c.fp;
        const T CaMKb = T(0.05) * (T(1) - y[CaMKt]) / (T(1) + T(0.0015) / y[Cass]);
        const T dCaMKt = T(0.05) * CaMKb * (CaMKb + y[CaMKt]) - T(0.00068) * y[CaMKt];
        // fluxes
        const T Jdiff = (y[Cass] - y[Cai]) / T(0.2);
        const T JdiffCl = (y[Clss] - y[Cli]) / T(2);
        const T JdiffK = (y[Kss] - y[Ki]) / T(2);
        const T JdiffNa = (y[Nass] - y[Nai]) / T(2);
        const T Jtr = (y[Cansr] - y[Cajsr]) / T(60);
        const T Jleak = T(0.0048825) * y[Cansr] / T(15);
        const T Jupnp = p[upScale] * T(0.005425) * y[Cai] / (y[Cai] + T(0.00092));
        const T Jupp = p[upScale] * T(2.75 * 0.005425) * y[Cai] / (y[Cai] + T(0.00092 - 0.00017));
        const T Jup = p[Jup_b] * ((T(1) - fp) * Jupnp + fp * Jupp - Jleak);
        const T Jrel = p[Jrel_b] * ((T(1) - fp) * y[Jrelnp] + fp * y[Jrelp]);
        // buffers
        const T Bcajsr = T(1) / (T(1) + T(10.0 * 0.8) / sq(T(0.8) + y[Cajsr]));
        const T Bcass = T(1) / (T(1) + T(0.047 * 0.00087) / sq(T(0.00087) + y[Cass]) +
                                T(1.124 * 0.0087) / sq(T(0.0087) + y[Cass]));
        const T Bcai = T(1) / (T(1) + p[cmdnmax] * T(0.00238) / sq(T(0.00238) + y[Cai]) +
                               T(0.07 * 0.0005) / sq(T(0.0005) + y[Cai]));
        const T AF = T(Acap / Fc);
        const T dNai = -(c.INa + c.INaL + T(3) * c.INaCai + c.ICaNai + T(3) * c.INaK + c.INab) *
                           AF / T(vmyo) + JdiffNa * T(vss / vmyo);
        const T dNass = -(c.ICaNass + T(3) * c.INaCass) * AF / T(vss) - JdiffNa;
        const T dKi = -(c.Ito + c.IKr + c.IKs + c.IK1 + c.IKb + c.Ikatp - T(2) * c.INaK + c.ICaKi) *
                          AF / T(vmyo) + JdiffK * T(vss / vmyo);
        const T dKss = -c.ICaKss * AF / T(vss) - JdiffK;
        const T dCli = (c.IClb + c.IClCasl) * AF / T(vmyo) + JdiffCl * T(vss / vmyo);
        const T dClss = c.IClCaj * AF / T(vss) - JdiffCl;
        const T dCai = Bcai * (-(c.ICaLi + c.IpCa + c.ICab - T(2) * c.INaCai) * AF / T(2 * vmyo) -
                               Jup * T(vnsr / vmyo) + Jdiff * T(vss / vmyo));
        const T dCass = Bcass * (-(c.ICaLss - T(2) * c.INaCass) * AF / T(2 * vss) +
                                 Jrel * T(vjsr / vss) - Jdiff);
        const T dCansr = Jup - Jtr * T(vjsr / vnsr);
        const T dCajsr = Bcajsr * (Jtr - Jrel);
        // state-dependent gate targets at the step start
        const T km2n = y[jca];
        const T anca_i = T(1) / (T(500) / km2n + sq(sq(T(1) + T(0.002) / y[Cai])));
        const T anca_ss = T(1) / (T(500) / km2n + sq(sq(T(1) + T(0.002) / y[Cass])));
        const T btv = p[bt];
        T trel = btv / (T(1) + T(0.0123) / y[Cajsr]);
        trel = trel < T(0.001) ? T(0.001) : trel;
        T trelp = T(1.25) * btv / (T(1) + T(0.0123) / y[Cajsr]);
        trelp = trelp < T(0.001) ? T(0.001) : trelp;
        const T relf = (tissue == 2 ? T(1.7) : T(1)) / (T(1) + cpow(T(1.7) / y[Cajsr], T(8)));
        const T Jrel_inf = -(T(0.5) * btv) * c.ICaLss * relf;
        const T Jrel_infp = -(T(0.5 * 1.25) * btv) * c.ICaLss * relf;
        // IKr Markov chain: each state linear in itself, the others at the step start
        const T C1v = y[C1], C2v = y[C2], C3v = y[C3], Iv = y[I], Ov = y[O];
        const T dC3 = q.be * C2v - q.al * C3v;
        const T aC2 = q.al * C3v + T(0.1911) * C1v, bC2 = q.be + T(0.154375);
        const T aC1 = T(0.154375) * C2v + q.be2 * Ov + q.beIC * Iv,
                bC1 = T(0.1911) + q.al2 + q.alCI;
        const T aO = q.al2 * C1v + q.bei * Iv, bO = q.be2 + q.ali;
        const T aI = q.alCI * C1v + q.ali * Ov, bI = q.beIC + q.bei;
        const T dep = tissue == 1 ? q.dep : T(1);
        const T tiF = q.tiF * dep, tiS = q.tiS * dep;
        const T ncai = y[nca_i], ncass = y[nca_ss];

        // writes
        out.gate(m, q.mss, q.tm);
        out.gate(h, q.hss, q.th);
        out.gate(hp, q.hssp, q.th);
        out.gate(j, q.hss, q.tj);
        out.gate(jp, q.hss, T(1.46) * q.tj);
        out.gate(hL, q.hLss, T(200));
        out.gate(mL, q.mLss, q.tm);
        out.gate(hLp, q.hLssp, T(600));
        out.gate(a, q.ass, q.tau_a);
        out.gate(ap, q.assp, q.tau_a);
        out.gate(iF, q.iss, tiF);
        out.gate(iS, q.iss, tiS);
        out.gate(iFp, q.iss, q.dti * tiF);
        out.gate(iSp, q.iss, q.dti * tiS);
        out.gate(d, q.dss, q.td);
        out.gate(ff, q.fss, q.tff);
        out.gate(fs, q.fss, q.tfs);
        out.gate(jca, q.jcass, T(72.5));
        out.gate(fcaf, q.fss, q.tfcaf);
        out.gate(fcafp, q.fss, T(2.5) * q.tfcaf);
        out.gate(fcas, q.fss, q.tfcas);
        out.gate(ffp, q.fss, T(2.5) * q.tff);
        out.gate(xs1, q.xs1ss, q.txs1);
        out.gate(xs2, q.xs1ss, q.txs2);
        out.lin(nca_i, anca_i * T(500), km2n, anca_i * T(500) - ncai * km2n);
        out.lin(nca_ss, anca_ss * T(500), km2n, anca_ss * T(500) - ncass * km2n);
        out.lin(C3, q.be * C2v, q.al, dC3);
        out.lin(C2, aC2, bC2, aC2 - bC2 * C2v);
        out.lin(C1, aC1, bC1, aC1 - bC1 * C1v);
        out.lin(O, aO, bO, aO - bO * Ov);
        out.lin(I, aI, bI, aI - bI * Iv);
        out.gate(Jrelnp, Jrel_inf, trel);
        out.gate(Jrelp, Jrel_infp, trelp);
        out.rate(CaMKt, dCaMKt);
        out.rate(Nai, dNai);
        out.rate(Nass, dNass);
        out.rate(Ki, dKi);
        out.rate(Kss, dKss);
        out.rate(Cli, dCli);
        out.rate(Clss, dClss);
        out.rate(Cai, dCai);
        out.rate(Cass, dCass);
        out.rate(Cansr, dCansr);
        out.rate(Cajsr, dCajsr);
        if (solve_v) out.rate(V, -Iion - istim);
    }

    template <class T>
    CG_HD static void step(T* y, T dt, const T* p, int tissue, const T* gs, bool solve_v,
                            T istim, const VTerms<T>& q)
    {
        RlStep<T> out{y, dt};
        rates(y, out, p, tissue, gs, solve_v, istim, q);
    }

    template <class T>
    CG_HD static void deriv(const T* y, T* dy, const T* p, int tissue, const T* gs,
                             bool solve_v, T istim, const VTerms<T>& q)
    {
        Deriv<T> out{y, dy};
        dy[V] = T(0);
        rates(y, out, p, tissue, gs, solve_v, istim, q);
    }
};

} // namespace cg
