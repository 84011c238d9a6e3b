// ten Tusscher-Noble-Noble-Panfilov 2004 (17 states) on gfx950.
// Equations: src/ionicModels/TNNP/TNNP_2004.H:385-509 (rates), :309-383 (constants and
// initial state); wrapper TNNP.C:122-174 (V = 1000 Vm in mV, Im = Iion_cm).
// Gates m h j d f s r Xr1 Xr2 Xs are Rush-Larsen; fCa and g relax with fixed tau but
// are frozen while they would rise above -60 mV (TNNP_2004.H:422,459), which RL keeps by
// taking xinf = x; K_i Na_i Ca_i Ca_SR are forward Euler.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct TNNP {
    static constexpr int NS = 17;
    static constexpr int NP = 42;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, ISCALE = 1.0, TSCALE = 1000.0, VMUL = 1000.0;
    enum { V, Ki, Nai, Cai, Xr1, Xr2, Xs, m, h, j, d, f, fCa, s, r, CaSR, g };
    enum { R, Tk, F, Cm, Vc, Pkna, Ko, Nao, Cao, gK1, gKr, gKs, gNa, gbNa, gCaL, gbCa, gto,
           PNaK, Kmk, KmNa, KNaCa, Ksat, alpha, gamma, KmCa, KmNai, gpCa, KpCa, gpK, taug,
           arel, brel, crel, Kup, Vleak, Vmaxup, Bufc, Kbufc, Bufsr, Kbufsr, Vsr, taufCa };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue < 1 || tissue > 3) return false;
        const double tab[NP] = {
            8.314, 310, 96.485, 185, 16404, 0.03, 5.4, 140, 2, 5.405, 0.096,
            tissue == 2 ? 0.062 : 0.245, 14.838, 0.00029, 0.175, 0.000592,
            tissue == 3 ? 0.073 : 0.294, 1.362, 1, 40, 1000, 0.1, 2.5, 0.35, 1.38, 87.5,
            0.825, 0.0005, 0.0146, 2, 0.016464, 0.25, 0.008232, 0.00025, 8e-5, 0.000425,
            0.15, 0.001, 10, 0.3, 1094, 2.0};
        for (int i = 0; i < NP; ++i) p[i] = tab[i];
        const double y0[NS] = {-86.2, 138.3, 11.6, 0.0002, 0, 1, 0, 0, 0.75, 0.75, 0, 1, 1,
                               1, 0, 0.2, 1};
        for (int i = 0; i < NS; ++i) y[i] = y0[i];
        return true;
    }

    // factors that depend on V only (shared by the step and the end-of-step current)
    template <class T> struct VTerms {
        T nak, ncx_num, ncx_den, cal_e2, ipk;
    };

    template <class T> CG_HD static VTerms<T> vterms(T Vv, const T* p)
    {
        const T FRT = p[F] / (p[R] * p[Tk]);
        VTerms<T> q;
        q.nak = T(1) / (T(1) + T(0.1245) * cexp(T(-0.1) * Vv * FRT) + T(0.0353) * cexp(-Vv * FRT));
        q.ncx_num = cexp(p[gamma] * Vv * FRT);
        q.ncx_den = cexp((p[gamma] - T(1)) * Vv * FRT);
        q.cal_e2 = cexp(T(2) * Vv * FRT);
        q.ipk = T(1) / (T(1) + cexp((T(25) - Vv) / T(5.98)));
        return q;
    }

    template <class T> struct Currents {
        T INa, IbNa, INaK, INaCa, IK1, Ito, IKr, IKs, ICaL, IbCa, IpK, IpCa;
        CG_HD T total() const
        {
            return IK1 + Ito + IKr + IKs + ICaL + INaK + INa + IbNa + INaCa + IbCa + IpK + IpCa;
        }
    };

    template <class T>
    CG_HD static Currents<T> currents(T Vv, const T* y, const T* p, const VTerms<T>& q,
                                       const T* gs)
    {
        // gs: per-cell multipliers of g_Na, g_Kr, g_CaL, g_Ks (1 when unscaled)
        const T RTF = p[R] * p[Tk] / p[F];
        const T FRT = p[F] / (p[R] * p[Tk]);
        Currents<T> c;
        const T ENa = RTF * clog(p[Nao] / y[Nai]);
        const T EK = RTF * clog(p[Ko] / y[Ki]);
        const T EKs = RTF * clog((p[Ko] + p[Pkna] * p[Nao]) / (y[Ki] + p[Pkna] * y[Nai]));
        const T ECa = T(0.5) * RTF * clog(p[Cao] / y[Cai]);
        c.INaK = (p[PNaK] * p[Ko] / (p[Ko] + p[Kmk])) * y[Nai] / (y[Nai] + p[KmNa]) * q.nak;
        c.INa = gs[0] * p[gNa] * cube(y[m]) * y[h] * y[j] * (Vv - ENa);
        c.IbNa = p[gbNa] * (Vv - ENa);
        c.INaCa = p[KNaCa] * (q.ncx_num * cube(y[Nai]) * p[Cao] -
                              q.ncx_den * cube(p[Nao]) * y[Cai] * p[alpha]) /
                  ((cube(p[KmNai]) + cube(p[Nao])) * (p[KmCa] + p[Cao]) * (T(1) + p[Ksat] * q.ncx_den));
        const T dvk = Vv - EK;
        const T aK1 = T(0.1) / (T(1) + cexp(T(0.06) * (dvk - T(200))));
        const T bK1 = (T(3) * cexp(T(0.0002) * (dvk + T(100))) + cexp(T(0.1) * (dvk - T(10)))) /
                      (T(1) + cexp(T(-0.5) * dvk));
        const T sko = csqrt(p[Ko] / T(5.4));
        c.IK1 = p[gK1] * (aK1 / (aK1 + bK1)) * sko * dvk;
        c.Ito = p[gto] * y[r] * y[s] * dvk;
        c.IKr = gs[1] * p[gKr] * sko * y[Xr1] * y[Xr2] * dvk;
        c.IKs = gs[3] * p[gKs] * sq(y[Xs]) * (Vv - EKs);
        if constexpr (is_fp32<T>())
            c.ICaL = gs[2] * (p[gCaL] * y[d] * y[f] * y[fCa] * T(4) * p[F] * FRT) *
                     (y[Cai] * q.cal_e2 - T(0.341) * p[Cao]) * ghk_quot(Vv, T(2) * FRT, q.cal_e2);
        else
            c.ICaL = gs[2] * (p[gCaL] * y[d] * y[f] * y[fCa] * T(4) * Vv * p[F] * FRT) *
                     (y[Cai] * q.cal_e2 - T(0.341) * p[Cao]) / (q.cal_e2 - T(1));
        c.IbCa = p[gbCa] * (Vv - ECa);
        c.IpK = p[gpK] * dvk * q.ipk;
        c.IpCa = p[gpCa] * y[Cai] / (y[Cai] + p[KpCa]);
        return c;
    }

    // Rates at the state y (y[V] is the membrane potential, mV) through an update policy;
    // V has a rate only when solve_v (-Iion - Istim).
    template <class T, class U>
    CG_HD static void rates(const T* y, U& out, const T* p, int tissue, const T* gs,
                             bool solve_v, T istim, const VTerms<T>& q)
    {
        const T Vv = y[V];
        const Currents<T> c = currents(Vv, y, p, q, gs);
        const T Iion = c.total();
        // concentration rates at the step start (TNNP_2004.H:475-509)
        const T cmvf = p[Cm] / (p[Vc] * p[F]);
        const T dNai = -(c.INa + c.IbNa + T(3) * c.INaK + T(3) * c.INaCa) * cmvf;
        const T dKi = -((c.IK1 + c.Ito + c.IKr + c.IKs + c.IpK + istim) - T(2) * c.INaK) * cmvf;
        const T Irel = (p[arel] * sq(y[CaSR]) / (sq(p[brel]) + sq(y[CaSR])) + p[crel]) * y[d] * y[g];
        const T Iup = p[Vmaxup] / (T(1) + sq(p[Kup]) / sq(y[Cai]));
        const T Ileak = p[Vleak] * (y[CaSR] - y[Cai]);
        const T dCai = ((-((c.ICaL + c.IbCa + c.IpCa) - T(2) * c.INaCa) / (T(2) * p[Vc] * p[F])) *
                        p[Cm] + Ileak - Iup + Irel) /
                       (T(1) + p[Bufc] * p[Kbufc] / sq(y[Cai] + p[Kbufc]));
        const T dCasr = (p[Vc] / p[Vsr]) * (Iup - (Irel + Ileak)) /
                        (T(1) + p[Bufsr] * p[Kbufsr] / sq(y[CaSR] + p[Kbufsr]));

        // Ca-dependent gate targets use Ca_i at the step start
        const T cai = y[Cai];
        const T fcainf = (T(1) / (T(1) + cpow(cai / T(0.000325), T(8))) +
                          T(0.1) / (T(1) + cexp((cai - T(0.0005)) / T(0.0001))) +
                          T(0.2) / (T(1) + cexp((cai - T(0.00075)) / T(0.0008))) + T(0.23)) / T(1.46);
        const T ginf = cai < T(0.00035) ? T(1) / (T(1) + cpow(cai / T(0.00035), T(6)))
                                        : T(1) / (T(1) + cpow(cai / T(0.00035), T(16)));
        const bool above = Vv > T(-60);
        const T fca_t = (fcainf > y[fCa] && above) ? y[fCa] : fcainf;
        const T g_t = (ginf > y[g] && above) ? y[g] : ginf;

        // voltage-gated RL (TNNP_2004.H:388-447)
        {
            const T minf = T(1) / sq(T(1) + cexp((T(-56.86) - Vv) / T(9.03)));
            const T am = T(1) / (T(1) + cexp((T(-60) - Vv) / T(5)));
            const T bm = T(0.1) / (T(1) + cexp((Vv + T(35)) / T(5))) +
                         T(0.1) / (T(1) + cexp((Vv - T(50)) / T(200)));
            out.gate(m, minf, am * bm);
            const T hinf = T(1) / sq(T(1) + cexp((Vv + T(71.55)) / T(7.43)));
            T ah, bh, aj, bj;
            if (Vv < T(-40)) {
                ah = T(0.057) * cexp(-(Vv + T(80)) / T(6.8));
                bh = T(2.7) * cexp(T(0.079) * Vv) + T(310000) * cexp(T(0.3485) * Vv);
                aj = ((T(-25428) * cexp(T(0.2444) * Vv) - T(6.948e-06) * cexp(T(-0.04391) * Vv)) *
                      (Vv + T(37.78))) / (T(1) + cexp(T(0.311) * (Vv + T(79.23))));
                bj = T(0.02424) * cexp(T(-0.01052) * Vv) / (T(1) + cexp(T(-0.1378) * (Vv + T(40.14))));
            } else {
                ah = T(0);
                bh = T(0.77) / (T(0.13) * (T(1) + cexp((Vv + T(10.66)) / T(-11.1))));
                aj = T(0);
                bj = T(0.6) * cexp(T(0.057) * Vv) / (T(1) + cexp(T(-0.1) * (Vv + T(32))));
            }
            out.gate(h, hinf, T(1) / (ah + bh));
            out.gate(j, hinf, T(1) / (aj + bj));
            const T dinf = T(1) / (T(1) + cexp((T(-5) - Vv) / T(7.5)));
            const T ad = T(1.4) / (T(1) + cexp((T(-35) - Vv) / T(13))) + T(0.25);
            const T bd = T(1.4) / (T(1) + cexp((Vv + T(5)) / T(5)));
            const T gd = T(1) / (T(1) + cexp((T(50) - Vv) / T(20)));
            out.gate(d, dinf, ad * bd + gd);
            const T finf = T(1) / (T(1) + cexp((Vv + T(20)) / T(7)));
            const T tf = T(1125) * cexp(-sq(Vv + T(27)) / T(240)) + T(80) +
                         T(165) / (T(1) + cexp((T(25) - Vv) / T(10)));
            out.gate(f, finf, tf);
            T sinf, ts;
            if (tissue == 1) {
                sinf = T(1) / (T(1) + cexp((Vv + T(20)) / T(5)));
                ts = T(85) * cexp(-sq(Vv + T(45)) / T(320)) + T(5) / (T(1) + cexp((Vv - T(20)) / T(5))) + T(3);
            } else {
                sinf = T(1.1) / (T(1) + cexp((Vv + T(28)) / T(6)));
                ts = T(1000) * cexp(-sq(Vv + T(67)) / T(1000)) + T(8);
            }
            out.gate(s, sinf, ts);
            const T rinf = T(1) / (T(1) + cexp((T(20) - Vv) / T(6)));
            const T tr = T(9.5) * cexp(-sq(Vv + T(40)) / T(1800)) + T(0.8);
            out.gate(r, rinf, tr);
            const T xr1inf = T(1) / (T(1) + cexp((T(-26) - Vv) / T(7)));
            const T txr1 = (T(450) / (T(1) + cexp((T(-45) - Vv) / T(10)))) *
                           (T(6) / (T(1) + cexp((Vv + T(30)) / T(11.5))));
            out.gate(Xr1, xr1inf, txr1);
            const T xr2inf = T(1) / (T(1) + cexp((Vv + T(88)) / T(24)));
            const T txr2 = (T(3) / (T(1) + cexp((T(-60) - Vv) / T(20)))) *
                           (T(1.12) / (T(1) + cexp((Vv - T(60)) / T(20))));
            out.gate(Xr2, xr2inf, txr2);
            const T xsinf = T(1) / (T(1) + cexp((T(-5) - Vv) / T(14)));
            const T txs = (T(1100) / csqrt(T(1) + cexp((T(-10) - Vv) / T(6)))) *
                          (T(1) / (T(1) + cexp((Vv - T(60)) / T(20))));
            out.gate(Xs, xsinf, txs);
        }
        out.gate(fCa, fca_t, p[taufCa]);
        out.gate(g, g_t, p[taug]);
        out.rate(Nai, dNai);
        out.rate(Ki, dKi);
        out.rate(Cai, dCai);
        out.rate(CaSR, dCasr);
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
