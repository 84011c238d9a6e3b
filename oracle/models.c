/*
 * Oracle (test infrastructure): FP64 restatement of the four ionic models on the
 * tissue path.  Each model exposes one evaluation that returns the rates, the
 * Rush-Larsen classification (xinf, tau) of each state, the total ionic current
 * and the single-cell stimulus.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   every model  - the reference's own generated expressions: each model's initConsts and
 *                  the function its derivatives() calls, evaluated from the header text
 *                  at seeded states (tests/golden/make_model_vectors.py); initial states
 *                  exact, rates in tissue and single-cell mode to 1e-13
 *                  (tests/test_oracle_model_vectors.py).
 *   BuenoOrovio  - also the singleCell.reference golden (RKF45 mode) and the
 *                  NiedererEtAl2012.reference golden (tissue).
 *   AlievPanfilov- also the reference's trace AlievPanfilov_myocyte_S1_1000.txt.
 *   TNNP, ORd    - also invariants: rest stability, RL exactness, published
 *                  action-potential shape, charge/concentration conservation.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

/* Heaviside of heaviside.H:39-49: strictly positive argument -> 1 */
static inline double H(double x) { return x > 0.0 ? 1.0 : 0.0; }

double orc_stimulus(double t, const orc_stim_t* s)
{
    /* stimulusIO.C:238-286 */
    if (!s) return 0.0;
    double I = 0.0;
    if (s->period_s1 > 0.0 && s->nstim1 > 0) {
        double tp = t - s->start;
        if (tp >= 0 && tp <= s->period_s1 * s->nstim1) {
            double ph = tp - floor(tp / s->period_s1) * s->period_s1;
            if (ph >= 0 && ph <= s->duration) I = -s->amplitude;
        }
    }
    if (I == 0.0 && s->period_s2 > 0.0 && s->nstim2 > 0) {
        double t1 = s->start + s->period_s1 * s->nstim1;
        double tp = t - t1;
        if (tp >= 0 && tp <= s->period_s2 * s->nstim2) {
            double ph = tp - floor(tp / s->period_s2) * s->period_s2;
            if (ph >= 0 && ph <= s->duration) I = -s->amplitude;
        }
    }
    return I;
}

/* =============================================================== BuenoOrovio
 * BuenoOrovio_2008.H:116-225; constant order (bo_consts.h) follows
 * BuenoOrovio_2008Names.H. */
#include "bo_consts.h"

static void bo_init(int tf, double* c, double* y)
{
    /* tissue flag: 1 epi, 2 M, 3 endo (anything else takes the endo column) */
    int e = tf == 1, m = tf == 2;
    c[BO_uo] = 0.0;
    c[BO_uu] = e ? 1.55 : m ? 1.61 : 1.56;
    c[BO_thv] = 0.3;
    c[BO_thw] = 0.13;
    c[BO_thvm] = e ? 0.006 : m ? 0.1 : 0.2;
    c[BO_tho] = m ? 0.005 : 0.006;
    c[BO_tv1m] = e ? 60 : m ? 80 : 75;
    c[BO_tv2m] = e ? 1150 : m ? 1.4506 : 10;
    c[BO_tvp] = 1.4506;
    c[BO_tw1m] = e ? 60 : m ? 70 : 6;
    c[BO_tw2m] = e ? 15 : m ? 8 : 140;
    c[BO_kwm] = e ? 65 : 200;
    c[BO_uwm] = e ? 0.03 : 0.016;
    c[BO_twp] = e ? 200 : 280;
    c[BO_tfi] = e ? 0.11 : m ? 0.078 : 0.1;
    c[BO_to1] = e ? 400 : m ? 410 : 470;
    c[BO_to2] = m ? 7 : 6;
    c[BO_tso1] = e ? 30.0181 : m ? 91 : 40;
    c[BO_tso2] = e ? 0.9957 : m ? 0.8 : 1.2;
    c[BO_kso] = e ? 2.0458 : m ? 2.1 : 2;
    c[BO_uso] = m ? 0.6 : 0.65;
    c[BO_ts1] = 2.7342;
    c[BO_ts2] = e ? 16 : m ? 4 : 2;
    c[BO_ks] = 2.0994;
    c[BO_us] = 0.9087;
    c[BO_tsi] = e ? 1.8875 : m ? 3.3849 : 2.9013;
    c[BO_twinf] = e ? 0.07 : m ? 0.01 : 0.0273;
    c[BO_winfstar] = e ? 0.94 : m ? 0.5 : 0.78;
    y[0] = 0.0; y[1] = 1.0; y[2] = 1.0; y[3] = 0.0;
}

static void bo_eval(double t, const double* c, const double* y, int solve_vm,
                    const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                    double* iion, double* istim)
{
    const double u = y[0], v = y[1], w = y[2], s = y[3];
    const double Hv = H(u - c[BO_thv]), Hw = H(u - c[BO_thw]);
    const double Ho = H(u - c[BO_tho]), Hvm = H(u - c[BO_thvm]);
    const double tso = c[BO_tso1] + (c[BO_tso2] - c[BO_tso1]) *
                       (1.0 + tanh(c[BO_kso] * (u - c[BO_uso]))) / 2.0;
    const double to = (1.0 - Ho) * c[BO_to1] + Ho * c[BO_to2];
    const double Jfi = -v * Hv * (u - c[BO_thv]) * (c[BO_uu] - u) / c[BO_tfi];
    const double Jso = (u - c[BO_uo]) * (1.0 - Hw) / to + Hw / tso;
    const double Jsi = -Hw * w * s / c[BO_tsi];
    const double Jion = Jfi + Jso + Jsi;
    const double tvm = (1.0 - Hvm) * c[BO_tv1m] + Hvm * c[BO_tv2m];
    const double twm = c[BO_tw1m] + (c[BO_tw2m] - c[BO_tw1m]) *
                       (1.0 + tanh(c[BO_kwm] * (u - c[BO_uwm]))) / 2.0;
    const double tS = (1.0 - Hw) * c[BO_ts1] + Hw * c[BO_ts2];
    const double vinf = H(c[BO_thvm] - u);
    const double winf = (1.0 - Ho) * (1.0 - u / c[BO_twinf]) + Ho * c[BO_winfstar];
    const double sinf = (1.0 + tanh(c[BO_ks] * (u - c[BO_us]))) / 2.0;

    r[1] = (1.0 - Hv) * (vinf - v) / tvm - Hv * v / c[BO_tvp];
    r[2] = (1.0 - Hw) * (winf - w) / twm - Hw * w / c[BO_twp];
    r[3] = (sinf - s) / tS;
    /* the same rates written as (xinf - x)/tau for Rush-Larsen */
    g[0] = 0;
    g[1] = 1; xi[1] = Hv > 0 ? 0.0 : vinf; ta[1] = Hv > 0 ? c[BO_tvp] : tvm;
    g[2] = 1; xi[2] = Hw > 0 ? 0.0 : winf; ta[2] = Hw > 0 ? c[BO_twp] : twm;
    g[3] = 1; xi[3] = sinf; ta[3] = tS;

    const double Ist = orc_stimulus(t, st);
    r[0] = solve_vm ? -Jion - Ist : 0.0;
    *iion = Jion; *istim = Ist;
}

/* ===================================================================== TNNP
 * ten Tusscher-Noble-Noble-Panfilov 2004, 17 states: TNNP_2004.H:309-509.
 * State order of TNNP_2004Names.H: V K_i Na_i Ca_i Xr1 Xr2 Xs m h j d f fCa s r
 * Ca_SR g. */
enum { TN_V, TN_Ki, TN_Nai, TN_Cai, TN_Xr1, TN_Xr2, TN_Xs, TN_m, TN_h, TN_j, TN_d,
       TN_f, TN_fCa, TN_s, TN_r, TN_CaSR, TN_g, TN_NS };
enum { TC_R, TC_T, TC_F, TC_Cm, TC_Vc, TC_Pkna, TC_Ko, TC_Nao, TC_Cao, TC_gK1, TC_gKr,
       TC_gKs, TC_gNa, TC_gbNa, TC_gCaL, TC_gbCa, TC_gto, TC_PNaK, TC_Kmk, TC_KmNa,
       TC_KNaCa, TC_Ksat, TC_alpha, TC_gamma, TC_KmCa, TC_KmNai, TC_gpCa, TC_KpCa,
       TC_gpK, TC_taug, TC_arel, TC_brel, TC_crel, TC_Kup, TC_Vleak, TC_Vmaxup,
       TC_Bufc, TC_Kbufc, TC_Bufsr, TC_Kbufsr, TC_Vsr, TC_taufCa, TC_NC };

static void tnnp_init(int tf, double* c, double* y)
{
    c[TC_R] = 8.314; c[TC_T] = 310; c[TC_F] = 96.485; c[TC_Cm] = 185; c[TC_Vc] = 16404;
    c[TC_Pkna] = 0.03; c[TC_Ko] = 5.4; c[TC_Nao] = 140; c[TC_Cao] = 2;
    c[TC_gK1] = 5.405; c[TC_gKr] = 0.096;
    c[TC_gKs] = tf == 2 ? 0.062 : 0.245;
    c[TC_gNa] = 14.838; c[TC_gbNa] = 0.00029; c[TC_gCaL] = 0.175; c[TC_gbCa] = 0.000592;
    c[TC_gto] = tf == 3 ? 0.073 : 0.294;
    c[TC_PNaK] = 1.362; c[TC_Kmk] = 1; c[TC_KmNa] = 40; c[TC_KNaCa] = 1000;
    c[TC_Ksat] = 0.1; c[TC_alpha] = 2.5; c[TC_gamma] = 0.35; c[TC_KmCa] = 1.38;
    c[TC_KmNai] = 87.5; c[TC_gpCa] = 0.825; c[TC_KpCa] = 0.0005; c[TC_gpK] = 0.0146;
    c[TC_taug] = 2; c[TC_arel] = 0.016464; c[TC_brel] = 0.25; c[TC_crel] = 0.008232;
    c[TC_Kup] = 0.00025; c[TC_Vleak] = 8e-5; c[TC_Vmaxup] = 0.000425;
    c[TC_Bufc] = 0.15; c[TC_Kbufc] = 0.001; c[TC_Bufsr] = 10; c[TC_Kbufsr] = 0.3;
    c[TC_Vsr] = 1094; c[TC_taufCa] = 2.0;

    y[TN_V] = -86.2; y[TN_Ki] = 138.3; y[TN_Nai] = 11.6; y[TN_Cai] = 0.0002;
    y[TN_Xr1] = 0; y[TN_Xr2] = 1; y[TN_Xs] = 0; y[TN_m] = 0; y[TN_h] = 0.75;
    y[TN_j] = 0.75; y[TN_d] = 0; y[TN_f] = 1; y[TN_fCa] = 1; y[TN_s] = 1; y[TN_r] = 0;
    y[TN_CaSR] = 0.2; y[TN_g] = 1;
}

static void tnnp_eval(int tf, double t, const double* c, const double* y, int solve_vm,
                      const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                      double* iion, double* istim)
{
    const double V = y[TN_V], Cai = y[TN_Cai], Nai = y[TN_Nai], Ki = y[TN_Ki];
    const double RTF = c[TC_R] * c[TC_T] / c[TC_F];
    const double FRT = c[TC_F] / (c[TC_R] * c[TC_T]);
    for (int i = 0; i < TN_NS; ++i) g[i] = 0;

    /* INa gates */
    double minf = 1.0 / pow(1.0 + exp((-56.86 - V) / 9.03), 2.0);
    double am = 1.0 / (1.0 + exp((-60.0 - V) / 5.0));
    double bm = 0.1 / (1.0 + exp((V + 35.0) / 5.0)) + 0.1 / (1.0 + exp((V - 50.0) / 200.0));
    double tm = 1.0 * am * bm;
    double hinf = 1.0 / pow(1.0 + exp((V + 71.55) / 7.43), 2.0);
    double ah = V < -40.0 ? 0.057 * exp(-(V + 80.0) / 6.8) : 0.0;
    double bh = V < -40.0 ? 2.7 * exp(0.079 * V) + 310000.0 * exp(0.3485 * V)
                          : 0.77 / (0.13 * (1.0 + exp((V + 10.66) / -11.1)));
    double th = 1.0 / (ah + bh);
    double jinf = hinf;
    double aj = V < -40.0 ? ((-25428.0 * exp(0.2444 * V) - 6.948e-06 * exp(-0.04391 * V)) *
                             (V + 37.78) / 1.0) / (1.0 + exp(0.311 * (V + 79.23)))
                          : 0.0;
    double bj = V < -40.0 ? 0.02424 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)))
                          : 0.6 * exp(0.057 * V) / (1.0 + exp(-0.1 * (V + 32.0)));
    double tj = 1.0 / (aj + bj);
    /* ICaL gates */
    double dinf = 1.0 / (1.0 + exp((-5.0 - V) / 7.5));
    double ad = 1.4 / (1.0 + exp((-35.0 - V) / 13.0)) + 0.25;
    double bd_ = 1.4 / (1.0 + exp((V + 5.0) / 5.0));
    double gd = 1.0 / (1.0 + exp((50.0 - V) / 20.0));
    double td = 1.0 * ad * bd_ + gd;
    double finf = 1.0 / (1.0 + exp((V + 20.0) / 7.0));
    double tf_ = 1125.0 * exp(-pow(V + 27.0, 2.0) / 240.0) + 80.0 +
                 165.0 / (1.0 + exp((25.0 - V) / 10.0));
    double afca = 1.0 / (1.0 + pow(Cai / 0.000325, 8.0));
    double bfca = 0.1 / (1.0 + exp((Cai - 0.0005) / 0.0001));
    double cfca = 0.2 / (1.0 + exp((Cai - 0.00075) / 0.0008));
    double fcainf = (afca + bfca + cfca + 0.23) / 1.46;
    /* Ito gates (tissue-conditional) */
    double sinf, ts;
    if (tf == 1) {
        sinf = 1.0 / (1.0 + exp((V + 20.0) / 5.0));
        ts = 85.0 * exp(-pow(V + 45.0, 2.0) / 320.0) + 5.0 / (1.0 + exp((V - 20.0) / 5.0)) + 3.0;
    } else {
        sinf = 1.1 / (1.0 + exp((V + 28.0) / 6.0));
        ts = 1000.0 * exp(-pow(V + 67.0, 2.0) / 1000.0) + 8.0;
    }
    double rinf = 1.0 / (1.0 + exp((20.0 - V) / 6.0));
    double tr = 9.5 * exp(-pow(V + 40.0, 2.0) / 1800.0) + 0.8;
    /* IKr / IKs gates */
    double xr1inf = 1.0 / (1.0 + exp((-26.0 - V) / 7.0));
    double axr1 = 450.0 / (1.0 + exp((-45.0 - V) / 10.0));
    double bxr1 = 6.0 / (1.0 + exp((V + 30.0) / 11.5));
    double txr1 = 1.0 * axr1 * bxr1;
    double xr2inf = 1.0 / (1.0 + exp((V + 88.0) / 24.0));
    double axr2 = 3.0 / (1.0 + exp((-60.0 - V) / 20.0));
    double bxr2 = 1.12 / (1.0 + exp((V - 60.0) / 20.0));
    double txr2 = 1.0 * axr2 * bxr2;
    double xsinf = 1.0 / (1.0 + exp((-5.0 - V) / 14.0));
    double axs = 1100.0 / pow(1.0 + exp((-10.0 - V) / 6.0), 0.5);
    double bxs = 1.0 / (1.0 + exp((V - 60.0) / 20.0));
    double txs = 1.0 * axs * bxs;
    /* RyR gate g */
    double ginf = Cai < 0.00035 ? 1.0 / (1.0 + pow(Cai / 0.00035, 6.0))
                                : 1.0 / (1.0 + pow(Cai / 0.00035, 16.0));

#define TN_GATE(idx, inf, tau) do { r[idx] = ((inf) - y[idx]) / (tau); g[idx] = 1; \
                                    xi[idx] = (inf); ta[idx] = (tau); } while (0)
    TN_GATE(TN_m, minf, tm);
    TN_GATE(TN_h, hinf, th);
    TN_GATE(TN_j, jinf, tj);
    TN_GATE(TN_d, dinf, td);
    TN_GATE(TN_f, finf, tf_);
    TN_GATE(TN_s, sinf, ts);
    TN_GATE(TN_r, rinf, tr);
    TN_GATE(TN_Xr1, xr1inf, txr1);
    TN_GATE(TN_Xr2, xr2inf, txr2);
    TN_GATE(TN_Xs, xsinf, txs);
#undef TN_GATE
    /* fCa and g: the rate is frozen at 0 while the gate would open above -60 mV
     * (TNNP_2004.H:422,459); otherwise first-order relaxation.  RL keeps the frozen
     * value by setting xinf = current value. */
    {
        int frz = fcainf > y[TN_fCa] && V > -60.0;
        r[TN_fCa] = frz ? 0.0 : (fcainf - y[TN_fCa]) / c[TC_taufCa];
        g[TN_fCa] = 1; xi[TN_fCa] = frz ? y[TN_fCa] : fcainf; ta[TN_fCa] = c[TC_taufCa];
        frz = ginf > y[TN_g] && V > -60.0;
        r[TN_g] = frz ? 0.0 : (ginf - y[TN_g]) / c[TC_taug];
        g[TN_g] = 1; xi[TN_g] = frz ? y[TN_g] : ginf; ta[TN_g] = c[TC_taug];
    }

    /* currents */
    double INaK = ((c[TC_PNaK] * c[TC_Ko] / (c[TC_Ko] + c[TC_Kmk])) * Nai / (Nai + c[TC_KmNa])) /
                  (1.0 + 0.1245 * exp(-0.1 * V * FRT) + 0.0353 * exp(-V * FRT));
    double ENa = RTF * log(c[TC_Nao] / Nai);
    double INa = c[TC_gNa] * pow(y[TN_m], 3.0) * y[TN_h] * y[TN_j] * (V - ENa);
    double IbNa = c[TC_gbNa] * (V - ENa);
    double INaCa = c[TC_KNaCa] *
                   (exp(c[TC_gamma] * V * FRT) * pow(Nai, 3.0) * c[TC_Cao] -
                    exp((c[TC_gamma] - 1.0) * V * FRT) * pow(c[TC_Nao], 3.0) * Cai * c[TC_alpha]) /
                   ((pow(c[TC_KmNai], 3.0) + pow(c[TC_Nao], 3.0)) * (c[TC_KmCa] + c[TC_Cao]) *
                    (1.0 + c[TC_Ksat] * exp((c[TC_gamma] - 1.0) * V * FRT)));
    double EK = RTF * log(c[TC_Ko] / Ki);
    double aK1 = 0.1 / (1.0 + exp(0.06 * ((V - EK) - 200.0)));
    double bK1 = (3.0 * exp(0.0002 * ((V - EK) + 100.0)) + 1.0 * exp(0.1 * ((V - EK) - 10.0))) /
                 (1.0 + exp(-0.5 * (V - EK)));
    double xK1 = aK1 / (aK1 + bK1);
    double IK1 = c[TC_gK1] * xK1 * pow(c[TC_Ko] / 5.4, 0.5) * (V - EK);
    double Ito = c[TC_gto] * y[TN_r] * y[TN_s] * (V - EK);
    double IKr = c[TC_gKr] * pow(c[TC_Ko] / 5.4, 0.5) * y[TN_Xr1] * y[TN_Xr2] * (V - EK);
    double EKs = RTF * log((c[TC_Ko] + c[TC_Pkna] * c[TC_Nao]) / (Ki + c[TC_Pkna] * Nai));
    double IKs = c[TC_gKs] * pow(y[TN_Xs], 2.0) * (V - EKs);
    double e2 = exp(2.0 * V * FRT);
    double ICaL = (c[TC_gCaL] * y[TN_d] * y[TN_f] * y[TN_fCa] * 4.0 * V * pow(c[TC_F], 2.0) /
                   (c[TC_R] * c[TC_T])) * (Cai * e2 - 0.341 * c[TC_Cao]) / (e2 - 1.0);
    double ECa = 0.5 * RTF * log(c[TC_Cao] / Cai);
    double IbCa = c[TC_gbCa] * (V - ECa);
    double IpK = c[TC_gpK] * (V - EK) / (1.0 + exp((25.0 - V) / 5.98));
    double IpCa = c[TC_gpCa] * Cai / (Cai + c[TC_KpCa]);
    double Iion = IK1 + Ito + IKr + IKs + ICaL + INaK + INa + IbNa + INaCa + IbCa + IpK + IpCa;
    double Ist = orc_stimulus(t, st);

    r[TN_V] = solve_vm ? -Iion - Ist : 0.0;
    r[TN_Nai] = -(INa + IbNa + 3.0 * INaK + 3.0 * INaCa) * c[TC_Cm] / (c[TC_Vc] * c[TC_F]);
    /* TNNP_2004.H:495: the potassium balance includes the stimulus current */
    r[TN_Ki] = -((IK1 + Ito + IKr + IKs + IpK + Ist) - 2.0 * INaK) * c[TC_Cm] / (c[TC_Vc] * c[TC_F]);
    double Irel = (c[TC_arel] * pow(y[TN_CaSR], 2.0) / (pow(c[TC_brel], 2.0) + pow(y[TN_CaSR], 2.0)) +
                   c[TC_crel]) * y[TN_d] * y[TN_g];
    double Iup = c[TC_Vmaxup] / (1.0 + pow(c[TC_Kup], 2.0) / pow(Cai, 2.0));
    double Ileak = c[TC_Vleak] * (y[TN_CaSR] - Cai);
    double dCai = ((-((ICaL + IbCa + IpCa) - 2.0 * INaCa) / (2.0 * c[TC_Vc] * c[TC_F])) * c[TC_Cm] +
                   Ileak - Iup) + Irel;
    double fCai = 1.0 / (1.0 + c[TC_Bufc] * c[TC_Kbufc] / pow(Cai + c[TC_Kbufc], 2.0));
    r[TN_Cai] = dCai * fCai;
    double dCasr = (c[TC_Vc] / c[TC_Vsr]) * (Iup - (Irel + Ileak));
    double fCasr = 1.0 / (1.0 + c[TC_Bufsr] * c[TC_Kbufsr] / pow(y[TN_CaSR] + c[TC_Kbufsr], 2.0));
    r[TN_CaSR] = dCasr * fCasr;
    *iion = Iion; *istim = Ist;
}

/* ====================================================================== ORd
 * O'Hara-Rudy 2011, 41 states: ORd_2011.H:493-989, state order of ORd_2011Names.H.
 *
 * Deviation: ORd_2011.H's ORdinitConsts never assigns 27 constants of INaK and SERCA
 * (Knai0, Knao0, delta, Kki, Kko, MgADP, MgATP, Kmgatp, H, eP, Khp, Knap, Kxkur,
 * k1p..k4m, a2, a4, b1, Pnak, upScale; CONSTANTS_ is zero-filled, ORd.C:54), which
 * makes INaK = 0*NaN = NaN in the reference.  The published ORd-2011 values
 * (O'Hara et al., PLoS Comput Biol 2011, endocardial column) are used here.  The
 * K_i rate omits the stimulus, as the reference does inside derivatives()
 * (ALGEBRAIC[Istim] is read before it is assigned, ORd_2011.H:962 vs :982, on a
 * zeroed scratch, ORd.C:165). */
enum { OR_V, OR_CaMKt, OR_Nai, OR_Nass, OR_Ki, OR_Kss, OR_Cass, OR_Cansr, OR_Cajsr, OR_Cai,
       OR_m, OR_hf, OR_hs, OR_j, OR_hsp, OR_jp, OR_mL, OR_hL, OR_hLp, OR_a, OR_iF, OR_iS,
       OR_ap, OR_iFp, OR_iSp, OR_d, OR_ff, OR_fs, OR_fcaf, OR_fcas, OR_jca, OR_ffp,
       OR_fcafp, OR_nca, OR_xrf, OR_xrs, OR_xs1, OR_xs2, OR_xk1, OR_Jrelnp, OR_Jrelp,
       OR_NS };
enum { OC_GNa, OC_GNaL, OC_Gto, OC_PCa, OC_GKr, OC_GKs, OC_GK1, OC_GKb, OC_Gncx, OC_Pnak,
       OC_upScale, OC_cmdnmax, OC_PCab, OC_PNab, OC_GpCa, OC_NC };

static void ord_init(int tf, double* c, double* y)
{
    (void)tf; /* ORd accepts "myocyte" only (ORd.C:93-96): endocardial column */
    c[OC_GNa] = 75.0; c[OC_GNaL] = 0.0075; c[OC_Gto] = 0.02; c[OC_PCa] = 0.0001;
    c[OC_GKr] = 0.046; c[OC_GKs] = 0.0034; c[OC_GK1] = 0.1908; c[OC_GKb] = 0.003;
    c[OC_Gncx] = 0.0008; c[OC_Pnak] = 30.0; c[OC_upScale] = 1.0; c[OC_cmdnmax] = 0.05;
    c[OC_PCab] = 2.5e-8; c[OC_PNab] = 3.75e-10; c[OC_GpCa] = 0.0005;
    static const double y0[OR_NS] = {
        -87.0, 0.0, 7.0, 7.0, 145.0, 145.0, 1e-4, 1.2, 1.2, 1e-4,
        0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0,
        0.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0,
        1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0 };
    memcpy(y, y0, sizeof y0);
}

static void ord_eval(double t, const double* c, const double* y, int solve_vm,
                     const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                     double* iion, double* istim)
{
    /* fixed constants of ORdinitConsts (ORd_2011.H:497-692) */
    const double L = 0.01, rad = 0.0011;
    const double Ageo = 2.0 * 3.14 * rad * rad + 2.0 * 3.14 * rad * L;
    const double vcell = 1000.0 * 3.14 * rad * rad * L;
    const double Acap = 2.0 * Ageo;
    const double vjsr = 0.0048 * vcell, vmyo = 0.68 * vcell, vnsr = 0.0552 * vcell,
                 vss = 0.02 * vcell;
    const double cao = 1.8, ko = 5.4, nao = 140.0;
    const double F = 96485.0, R = 8314.0, T = 310.0;
    const double KmCaMK = 0.15, aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015;
    const double KmCap = 0.0005;
    const double cm = 1.0;

    const double V = y[OR_V], Cai = y[OR_Cai], Cass = y[OR_Cass], Nai = y[OR_Nai],
                 Nass = y[OR_Nass], Ki = y[OR_Ki], Kss = y[OR_Kss];
    for (int i = 0; i < OR_NS; ++i) g[i] = 0;

    /* CaMK */
    double CaMKb = CaMKo * (1.0 - y[OR_CaMKt]) / (1.0 + KmCaM / Cass);
    double CaMKa = CaMKb + y[OR_CaMKt];
    r[OR_CaMKt] = aCaMK * CaMKb * (CaMKb + y[OR_CaMKt]) - bCaMK * y[OR_CaMKt];
    double IpCa = c[OC_GpCa] * Cai / (KmCap + Cai);
    double Jdiff = (Cass - Cai) / 0.2, JdiffK = (Kss - Ki) / 2.0, JdiffNa = (Nass - Nai) / 2.0;
    double Jtr = (y[OR_Cansr] - y[OR_Cajsr]) / 100.0;

    /* INaCa (i and ss compartments) */
    const double KmCaAct = 0.00015, kasymm = 12.5, kcaoff = 5000.0, kcaon = 1500000.0,
                 kna1 = 15.0, kna2 = 5.0, kna3 = 88.12, qca = 0.167, qna = 0.5224,
                 wca = 60000.0, wna = 60000.0, wnaca = 5000.0;
    const double h10 = kasymm + 1.0 + nao / kna1 * (1.0 + nao / kna2);
    const double h11 = nao * nao / (h10 * kna1 * kna2);
    const double h12 = 1.0 / h10;
    const double k1 = h12 * cao * kcaon, k2 = kcaoff, k5 = kcaoff;
    double hca = exp(qca * V * F / (R * T));
    double hna = exp(qna * V * F / (R * T));
    double INaCa[2];
    for (int comp = 0; comp < 2; ++comp) {
        double na = comp == 0 ? Nai : Nass, ca = comp == 0 ? Cai : Cass;
        double allo = 1.0 / (1.0 + pow(KmCaAct / ca, 2.0));
        double h4 = 1.0 + na / kna1 * (1.0 + na / kna2);
        double h1 = 1.0 + na / kna3 * (1.0 + hna);
        double h5 = na * na / (h4 * kna1 * kna2);
        double h6 = 1.0 / h4;
        double h7 = 1.0 + nao / kna3 * (1.0 + 1.0 / hna);
        double h2 = na * hna / (kna3 * h1);
        double h3 = 1.0 / h1;
        double h8 = nao / (kna3 * hna * h7);
        double h9 = 1.0 / h7;
        double k6 = h6 * ca * kcaon;
        double k3p = h9 * wca, k3pp = h8 * wnaca, k4p = h3 * wca / hca, k4pp = h2 * wnaca;
        double k7 = h5 * h2 * wna, k8 = h8 * h11 * wna;
        double k3 = k3p + k3pp, k4 = k4p + k4pp;
        double x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
        double x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
        double x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
        double x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
        double E1 = x1 / (x1 + x2 + x3 + x4), E2 = x2 / (x1 + x2 + x3 + x4),
               E3 = x3 / (x1 + x2 + x3 + x4), E4 = x4 / (x1 + x2 + x3 + x4);
        double JCa = E2 * k2 - E1 * k1;
        double JNa = 3.0 * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
        INaCa[comp] = (comp == 0 ? 0.8 : 0.2) * c[OC_Gncx] * allo * (1.0 * JNa + 2.0 * JCa);
    }

    /* INaK (published constants) */
    const double k1p = 949.5, k1m = 182.4, k2p = 687.2, k2m = 39.4, k3p_ = 1899.0,
                 k3m = 79300.0, k4p_ = 639.0, k4m = 40.0, Knai0 = 9.073, Knao0 = 27.78,
                 delta = -0.155, Kki = 0.5, Kko = 0.3582, MgADP = 0.05, MgATP = 9.8,
                 Kmgatp = 1.698e-7, Hp = 1.0e-7, eP = 4.2, Khp = 1.698e-7, Knap = 224.0,
                 Kxkur = 292.0;
    double Knai = Knai0 * exp(delta * V * F / (3.0 * R * T));
    double Knao = Knao0 * exp((1.0 - delta) * V * F / (3.0 * R * T));
    double P = eP / (1.0 + Hp / Khp + Nai / Knap + Ki / Kxkur);
    double a1 = k1p * pow(Nai / Knai, 3.0) /
                (pow(1.0 + Nai / Knai, 3.0) + pow(1.0 + Ki / Kki, 2.0) - 1.0);
    double b1 = k1m * MgADP, a2 = k2p;
    double b2 = k2m * pow(nao / Knao, 3.0) /
                (pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / Kko, 2.0) - 1.0);
    double a3 = k3p_ * pow(ko / Kko, 2.0) /
                (pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / Kko, 2.0) - 1.0);
    double b3 = k3m * P * Hp / (1.0 + MgATP / Kmgatp);
    double a4 = k4p_ * MgATP / Kmgatp / (1.0 + MgATP / Kmgatp);
    double b4 = k4m * pow(Ki / Kki, 2.0) /
                (pow(1.0 + Nai / Knai, 3.0) + pow(1.0 + Ki / Kki, 2.0) - 1.0);
    double x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
    double x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
    double x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
    double x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
    double sx = x1 + x2 + x3 + x4;
    double E1 = x1 / sx, E2 = x2 / sx, E3 = x3 / sx, E4 = x4 / sx;
    double JnakK = 2.0 * (E4 * b1 - E3 * a1);
    double JnakNa = 3.0 * (E1 * a3 - E2 * b3);
    double INaK = c[OC_Pnak] * (1.0 * JnakNa + 1.0 * JnakK);

    /* SERCA */
    double Jleak = 0.0039375 * y[OR_Cansr] / 15.0;
    double fJupp = 1.0 / (1.0 + KmCaMK / CaMKa);
    double Jupnp = c[OC_upScale] * 0.004375 * Cai / (Cai + 0.00092);
    double Jupp = c[OC_upScale] * 2.75 * 0.004375 * Cai / (Cai + 0.00092 - 0.00017);
    double Jup = (1.0 - fJupp) * Jupnp + fJupp * Jupp - Jleak;

    /* reversal potentials */
    const double PKNa = 0.01833;
    double EK = R * T / F * log(ko / Ki);
    double ENa = R * T / F * log(nao / Nai);
    double EKs = R * T / F * log((ko + PKNa * nao) / (Ki + PKNa * Nai));

#define OR_GATE(idx, inf, tau) do { double _i = (inf), _t = (tau); \
        r[idx] = (_i - y[idx]) / _t; g[idx] = 1; xi[idx] = _i; ta[idx] = _t; } while (0)
    /* IK1 */
    double rk1 = 1.0 / (1.0 + exp((V + 105.8 - 2.6 * ko) / 9.493));
    double txk1 = 122.2 / (exp(-(V + 127.2) / 20.36) + exp((V + 236.8) / 69.33));
    double xk1ss = 1.0 / (1.0 + exp(-(V + 2.5538 * ko + 144.59) / (1.5692 * ko + 3.8115)));
    OR_GATE(OR_xk1, xk1ss, txk1);
    double IK1 = c[OC_GK1] * sqrt(ko) * rk1 * y[OR_xk1] * (V - EK);
    /* IKb */
    double xkb = 1.0 / (1.0 + exp(-(V - 14.48) / 18.34));
    double IKb = c[OC_GKb] * xkb * (V - EK);
    /* IKr */
    double Axrf = 1.0 / (1.0 + exp((V + 54.81) / 38.21));
    double rkr = 1.0 / (1.0 + exp((V + 55.0) / 75.0)) * 1.0 / (1.0 + exp((V - 10.0) / 30.0));
    double txrf = 12.98 + 1.0 / (0.3652 * exp((V - 31.66) / 3.869) + 4.123e-05 * exp(-(V - 47.78) / 20.38));
    double txrs = 1.865 + 1.0 / (0.06629 * exp((V - 34.7) / 7.355) + 1.128e-05 * exp(-(V - 29.74) / 25.94));
    double xrss = 1.0 / (1.0 + exp(-(V + 8.337) / 6.789));
    double Axrs = 1.0 - Axrf;
    OR_GATE(OR_xrf, xrss, txrf);
    OR_GATE(OR_xrs, xrss, txrs);
    double xr = Axrf * y[OR_xrf] + Axrs * y[OR_xrs];
    double IKr = c[OC_GKr] * sqrt(ko / 5.4) * xr * rkr * (V - EK);
    /* IKs */
    double KsCa = 1.0 + 0.6 / (1.0 + pow(3.8e-05 / Cai, 1.4));
    double txs1 = 817.3 + 1.0 / (0.0002326 * exp((V + 48.28) / 17.8) + 0.001292 * exp(-(V + 210.0) / 230.0));
    double txs2 = 1.0 / (0.01 * exp((V - 50.0) / 20.0) + 0.0193 * exp(-(V + 66.54) / 31.0));
    double xs1ss = 1.0 / (1.0 + exp(-(V + 11.6) / 8.932));
    OR_GATE(OR_xs1, xs1ss, txs1);
    OR_GATE(OR_xs2, xs1ss, txs2);
    double IKs = c[OC_GKs] * KsCa * y[OR_xs1] * y[OR_xs2] * (V - EKs);
    /* INa */
    double fINap = 1.0 / (1.0 + KmCaMK / CaMKa);
    double hssp = 1.0 / (1.0 + exp((V + 89.1) / 6.086));
    double thf = 1.0 / (1.432e-05 * exp(-(V + 1.196) / 6.285) + 6.149 * exp((V + 0.5096) / 20.27));
    double ths = 1.0 / (0.009794 * exp(-(V + 17.95) / 28.05) + 0.3343 * exp((V + 5.73) / 56.66));
    double tj = 2.038 + 1.0 / (0.02136 * exp(-(V + 100.6) / 8.281) + 0.3052 * exp((V + 0.9941) / 38.45));
    double hss = 1.0 / (1.0 + exp((V + 82.9) / 6.086));
    double mss = 1.0 / (1.0 + exp(-(V + 39.57) / 9.871));
    double tm = 1.0 / (6.765 * exp((V + 11.64) / 34.77) + 8.552 * exp(-(V + 77.42) / 5.955));
    const double Ahf = 0.99, Ahs = 1.0 - Ahf;
    double h = Ahf * y[OR_hf] + Ahs * y[OR_hs];
    double hp = Ahf * y[OR_hf] + Ahs * y[OR_hsp];
    OR_GATE(OR_hf, hss, thf);
    OR_GATE(OR_hs, hss, ths);
    OR_GATE(OR_hsp, hssp, 3.0 * ths);
    OR_GATE(OR_m, mss, tm);
    double INa = c[OC_GNa] * (V - ENa) * pow(y[OR_m], 3.0) *
                 ((1.0 - fINap) * h * y[OR_j] + fINap * hp * y[OR_jp]);
    OR_GATE(OR_j, hss, tj);
    OR_GATE(OR_jp, hss, 1.46 * tj);
    /* Ito */
    double AiF = 1.0 / (1.0 + exp((V - 213.6) / 151.2));
    double ass = 1.0 / (1.0 + exp(-(V - 14.34) / 14.82));
    double assp = 1.0 / (1.0 + exp(-(V - 24.34) / 14.82));
    double dti_dev = 1.354 + 0.0001 / (exp((V - 167.4) / 15.89) + exp(-(V - 12.23) / 0.2154));
    double dti_rec = 1.0 - 0.5 / (1.0 + exp((V + 70.0) / 20.0));
    double fItop = 1.0 / (1.0 + KmCaMK / CaMKa);
    double iss = 1.0 / (1.0 + exp((V + 43.94) / 5.711));
    double ta_ = 1.0515 / (1.0 / (1.2089 * (1.0 + exp(-(V - 18.4099) / 29.3814))) +
                           3.5 / (1.0 + exp((V + 100.0) / 29.3814)));
    double tiF = 4.562 + 1.0 / (0.3933 * exp(-(V + 100.0) / 100.0) + 0.08004 * exp((V + 50.0) / 16.59));
    double tiS = 23.62 + 1.0 / (0.001416 * exp(-(V + 96.52) / 59.05) + 1.78e-08 * exp((V + 114.1) / 8.079));
    double AiS = 1.0 - AiF;
    /* delta_epi = 1 outside the epicardial column */
    OR_GATE(OR_a, ass, ta_);
    OR_GATE(OR_ap, assp, ta_);
    double ii = AiF * y[OR_iF] + AiS * y[OR_iS];
    double ip = AiF * y[OR_iFp] + AiS * y[OR_iSp];
    double tiFp = dti_dev * dti_rec * tiF, tiSp = dti_dev * dti_rec * tiS;
    OR_GATE(OR_iF, iss, tiF);
    OR_GATE(OR_iS, iss, tiS);
    double Ito = c[OC_Gto] * (V - EK) * ((1.0 - fItop) * y[OR_a] * ii + fItop * y[OR_ap] * ip);
    OR_GATE(OR_iFp, iss, tiFp);
    OR_GATE(OR_iSp, iss, tiSp);
    /* INaL */
    double fINaLp = 1.0 / (1.0 + KmCaMK / CaMKa);
    double hLss = 1.0 / (1.0 + exp((V + 87.61) / 7.488));
    double hLssp = 1.0 / (1.0 + exp((V + 93.81) / 7.488));
    double mLss = 1.0 / (1.0 + exp(-(V + 42.85) / 5.264));
    const double thL = 200.0, thLp = 3.0 * thL;
    OR_GATE(OR_hL, hLss, thL);
    OR_GATE(OR_mL, mLss, tm);
    double INaL = c[OC_GNaL] * (V - ENa) * y[OR_mL] * ((1.0 - fINaLp) * y[OR_hL] + fINaLp * y[OR_hLp]);
    OR_GATE(OR_hLp, hLssp, thLp);
    /* ICaL */
    double Afcaf = 0.3 + 0.6 / (1.0 + exp((V - 10.0) / 10.0));
    double dss = 1.0 / (1.0 + exp(-(V + 3.94) / 4.23));
    double fICaLp = 1.0 / (1.0 + KmCaMK / CaMKa);
    double fss = 1.0 / (1.0 + exp((V + 19.58) / 3.696));
    double km2n = y[OR_jca] * 1.0;
    double td = 0.6 + 1.0 / (exp(-0.05 * (V + 6.0)) + exp(0.09 * (V + 14.0)));
    double tfcaf = 7.0 + 1.0 / (0.04 * exp(-(V - 4.0) / 7.0) + 0.04 * exp((V - 4.0) / 7.0));
    double tfcas = 100.0 + 1.0 / (0.00012 * exp(-V / 3.0) + 0.00012 * exp(V / 7.0));
    double tff = 7.0 + 1.0 / (0.0045 * exp(-(V + 20.0) / 10.0) + 0.0045 * exp((V + 20.0) / 10.0));
    double tfs = 1000.0 + 1.0 / (3.5e-05 * exp(-(V + 5.0) / 4.0) + 3.5e-05 * exp((V + 5.0) / 6.0));
    double Afcas = 1.0 - Afcaf;
    const double k2n = 1000.0, Kmn = 0.002, tjca = 75.0, Aff = 0.6, Afs = 1.0 - Aff;
    double anca = 1.0 / (k2n / km2n + pow(1.0 + Kmn / Cass, 4.0));
    double fcass = fss;
    OR_GATE(OR_d, dss, td);
    OR_GATE(OR_ff, fss, tff);
    OR_GATE(OR_fs, fss, tfs);
    double f = Aff * y[OR_ff] + Afs * y[OR_fs];
    double fca = Afcaf * y[OR_fcaf] + Afcas * y[OR_fcas];
    double fcap = Afcaf * y[OR_fcafp] + Afcas * y[OR_fcas];
    double fp = Aff * y[OR_ffp] + Afs * y[OR_fs];
    OR_GATE(OR_fcaf, fcass, tfcaf);
    OR_GATE(OR_fcafp, fcass, 2.5 * tfcaf);
    OR_GATE(OR_fcas, fcass, tfcas);
    OR_GATE(OR_ffp, fss, 2.5 * tff);
    OR_GATE(OR_jca, fcass, tjca);
    /* nca: dn/dt = anca*k2n - n*km2n is linear in n -> xinf = anca*k2n/km2n, tau = 1/km2n */
    r[OR_nca] = anca * k2n - y[OR_nca] * km2n;
    g[OR_nca] = 1; xi[OR_nca] = anca * k2n / km2n; ta[OR_nca] = 1.0 / km2n;
    /* buffers */
    const double BSRmax = 0.047, KmBSR = 0.00087, BSLmax = 1.124, KmBSL = 0.0087,
                 csqnmax = 10.0, kmcsqn = 0.8, kmcmdn = 0.00238, trpnmax = 0.07, kmtrpn = 0.0005;
    r[OR_Cansr] = Jup - Jtr * vjsr / vnsr;
    double Bcajsr = 1.0 / (1.0 + csqnmax * kmcsqn / pow(kmcsqn + y[OR_Cajsr], 2.0));
    double Bcass = 1.0 / (1.0 + BSRmax * KmBSR / pow(KmBSR + Cass, 2.0) +
                          BSLmax * KmBSL / pow(KmBSL + Cass, 2.0));
    double Bcai = 1.0 / (1.0 + c[OC_cmdnmax] * kmcmdn / pow(kmcmdn + Cai, 2.0) +
                         trpnmax * kmtrpn / pow(kmtrpn + Cai, 2.0));
    double vffrt = V * F * F / (R * T), vfrt = V * F / (R * T);
    /* RyR */
    double fJrelp = 1.0 / (1.0 + KmCaMK / CaMKa);
    double Jrel = (1.0 - fJrelp) * y[OR_Jrelnp] + fJrelp * y[OR_Jrelp];
    const double bt = 4.75, a_rel = 0.5 * bt, btp = 1.25 * bt, a_relp = 0.5 * btp;
    double trel = bt / (1.0 + 0.0123 / y[OR_Cajsr]);
    if (trel < 0.001) trel = 0.001;
    double trelp = btp / (1.0 + 0.0123 / y[OR_Cajsr]);
    if (trelp < 0.001) trelp = 0.001;
    /* GHK fluxes */
    double e1 = exp(1.0 * vfrt), e2 = exp(2.0 * vfrt);
    double PhiCaK = 1.0 * vffrt * (0.75 * Kss * e1 - 0.75 * ko) / (e1 - 1.0);
    double PhiCaL = 4.0 * vffrt * (Cass * e2 - 0.341 * cao) / (e2 - 1.0);
    double PhiCaNa = 1.0 * vffrt * (0.75 * Nass * e1 - 0.75 * nao) / (e1 - 1.0);
    double ICab = c[OC_PCab] * 4.0 * vffrt * (Cai * e2 - 0.341 * cao) / (e2 - 1.0);
    double INab = c[OC_PNab] * vffrt * (Nai * exp(vfrt) - nao) / (exp(vfrt) - 1.0);
    r[OR_Cajsr] = Bcajsr * (Jtr - Jrel);
    r[OR_Ki] = -(Ito + IKr + IKs + IK1 + IKb - 2.0 * INaK) * cm * Acap / (F * vmyo) +
               JdiffK * vss / vmyo;
    const double PCa = c[OC_PCa], PCap = 1.1 * PCa, PCaK = 0.0003574 * PCa,
                 PCaNa = 0.00125 * PCa, PCaKp = 0.0003574 * PCap, PCaNap = 0.00125 * PCap;
    double gnp = f * (1.0 - y[OR_nca]) + y[OR_jca] * fca * y[OR_nca];
    double gp = fp * (1.0 - y[OR_nca]) + y[OR_jca] * fcap * y[OR_nca];
    double ICaK = (1.0 - fICaLp) * PCaK * PhiCaK * y[OR_d] * gnp + fICaLp * PCaKp * PhiCaK * y[OR_d] * gp;
    double ICaL = (1.0 - fICaLp) * PCa * PhiCaL * y[OR_d] * gnp + fICaLp * PCap * PhiCaL * y[OR_d] * gp;
    double ICaNa = (1.0 - fICaLp) * PCaNa * PhiCaNa * y[OR_d] * gnp + fICaLp * PCaNap * PhiCaNa * y[OR_d] * gp;
    r[OR_Cai] = Bcai * (-(IpCa + ICab - 2.0 * INaCa[0]) * cm * Acap / (2.0 * F * vmyo) -
                        Jup * vnsr / vmyo + Jdiff * vss / vmyo);
    r[OR_Nai] = -(INa + INaL + 3.0 * INaCa[0] + 3.0 * INaK + INab) * Acap * cm / (F * vmyo) +
                JdiffNa * vss / vmyo;
    r[OR_Cass] = Bcass * (-(ICaL - 2.0 * INaCa[1]) * cm * Acap / (2.0 * F * vss) +
                          Jrel * vjsr / vss - Jdiff);
    r[OR_Kss] = -ICaK * cm * Acap / (F * vss) - JdiffK;
    r[OR_Nass] = -(ICaNa + 3.0 * INaCa[1]) * cm * Acap / (F * vss) - JdiffNa;
    double Jrel_inf = a_rel * -ICaL / (1.0 + 1.0 * pow(1.5 / y[OR_Cajsr], 8.0));
    double Jrel_infp = a_relp * -ICaL / (1.0 + pow(1.5 / y[OR_Cajsr], 8.0));
    OR_GATE(OR_Jrelnp, Jrel_inf, trel);
    OR_GATE(OR_Jrelp, Jrel_infp, trelp);
#undef OR_GATE

    double Iion = INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa[0] +
                  INaCa[1] + INaK + INab + IKb + IpCa + ICab;
    double Ist = orc_stimulus(t, st);
    r[OR_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion; *istim = Ist;
}

/* ============================================================ AlievPanfilov
 * AlievPanfilov_1996.H:72-120 (time in units of 12.9 ms, u dimensionless). */
static void ap_init(double* c, double* y)
{
    c[0] = 20.0;   /* Vp */
    c[1] = -80.0;  /* Vr */
    c[2] = -70.0;  /* Vth */
    c[3] = 0.01;   /* epsilon */
    c[4] = 8.0;    /* k */
    c[5] = 0.2;    /* mu1 */
    c[6] = 0.3;    /* mu2 */
    c[7] = (c[2] - c[1]) / (c[0] - c[1]); /* a */
    y[0] = 0.05; y[1] = 0.0;
}

static void ap_eval(double t, const double* c, const double* y, int solve_vm,
                    const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                    double* iion, double* istim)
{
    (void)xi; (void)ta;
    const double u = y[0], rr = y[1], k = c[4], a = c[7];
    double eps = c[3] + c[5] * rr / (c[6] + u);
    r[1] = eps * (-rr - k * u * (u - (a + 1.0)));
    double Iion = k * u * (u - a) * (u - 1.0) + rr * u;
    double Ist = orc_stimulus(t, st);
    r[0] = solve_vm ? -Iion - Ist : 0.0;
    g[0] = 0; g[1] = 0;
    *iion = Iion; *istim = Ist;
}

/* ========================================= monodomainFDAManufactured (FDA)
 * monodomainFDAManufactured_2014.H:40-80 (constants per dimension, rates, Iion) and
 * monodomainFDAManufactured.C:100-147 (time in s, V = Vm, Im = Iion / Cm).  The states
 * start from the manufactured fields of manufacturedFDAReference.H, set by the caller.
 * tissue carries the dimension: ORC_DIM1/2/3. */
static void fda_init(int dim, double* c, double* y)
{
    c[0] = 2.0;                                               /* Cm */
    c[1] = dim == ORC_DIM1 ? -1.1 : dim == ORC_DIM2 ? -5.9 : -8.6;  /* Beta */
    c[2] = 3.0;                                               /* Chi */
    y[0] = y[1] = y[2] = y[3] = 0.0;
}

static void fda_eval(const double* c, const double* y, double* r, int* g, double* iion)
{
    const double V = y[0], u1 = y[1], u2 = y[2], u3 = y[3];
    const double a = u1 + u3 - V;
    r[0] = 0.0;
    r[1] = pow(a, 2.0) * pow(u2, 2.0) + 0.5 * a * pow(u2, 2.0) * (V - u3);
    r[2] = -a * pow(u2, 3.0);
    r[3] = 0.0;
    g[0] = g[1] = g[2] = g[3] = 0;
    *iion = -0.5 * c[0] * a * pow(u2, 2.0) * (V - u3) + (c[1] * (V - u3)) / c[2];
}

/* ============================================================== Courtemanche
 * Courtemanche-Ramirez-Nattel 1998 human atrial model, 21 states: constants and initial
 * state of Courtemanche_1998.H:337-455, right-hand side of :458-658, state order of
 * Courtemanche_1998Names.H.  The reference accepts tissue "myocyte" only
 * (Courtemanche.C:92-95).  The K_i balance reads ALGEBRAIC[Istim] before the call
 * assigns it (Courtemanche_1998.H:644 vs :652) from the zeroed scratch of derivatives()
 * (Courtemanche.C:161), so it omits the stimulus, as for ORd.
 * Parity pins (tests/test_oracle_models.py): charge conservation of the concentration
 * balances against Cm dV/dt, the published action potential at 1 Hz, and the
 * removable singularities of the rate expressions checked against their limits. */
enum { CR_V, CR_Nai, CR_Ki, CR_Cai, CR_CaUp, CR_CaRel, CR_m, CR_h, CR_j, CR_oa, CR_oi,
       CR_ua, CR_ui, CR_xr, CR_xs, CR_d, CR_f, CR_fCa, CR_u, CR_v, CR_w, CR_NS };
enum { CC_CMDN_max, CC_CSQN_max, CC_Km_CMDN, CC_Km_CSQN, CC_Km_TRPN, CC_TRPN_max,
       CC_Ca_up_max, CC_I_up_max, CC_K_up, CC_Cao, CC_Ko, CC_Nao, CC_Cm, CC_V_cell, CC_V_i,
       CC_V_rel, CC_V_up, CC_ECaL, CC_gCaL, CC_fCa_tau, CC_IpCa_max, CC_tau_tr, CC_F,
       CC_R, CC_T, CC_RTF, CC_FRT, CC_KQ10, CC_INaCa_max, CC_KmCa, CC_KmNa, CC_gam,
       CC_ksat, CC_INaK_max, CC_KmKo, CC_KmNai, CC_sigma, CC_K_rel, CC_c1, CC_c2,
       CC_u_tau, CC_gbCa, CC_gbNa, CC_gK1, CC_gKr, CC_gKs, CC_gKur_base, CC_gNa, CC_gto,
       CC_NC };

static void crn_init(double* c, double* y)
{
    c[CC_CMDN_max] = 0.05; c[CC_CSQN_max] = 10.0; c[CC_Km_CMDN] = 0.00238;
    c[CC_Km_CSQN] = 0.8; c[CC_Km_TRPN] = 0.0005; c[CC_TRPN_max] = 0.07;
    c[CC_Ca_up_max] = 15.0; c[CC_I_up_max] = 0.005; c[CC_K_up] = 0.00092;
    c[CC_Cao] = 1.8; c[CC_Ko] = 5.4; c[CC_Nao] = 140.0;
    c[CC_Cm] = 100.0; c[CC_V_cell] = 20100.0;
    c[CC_V_i] = 0.68 * c[CC_V_cell]; c[CC_V_rel] = 0.0048 * c[CC_V_cell];
    c[CC_V_up] = 0.0552 * c[CC_V_cell];
    c[CC_ECaL] = 65.0; c[CC_gCaL] = 0.12375; c[CC_fCa_tau] = 2.0; c[CC_IpCa_max] = 0.275;
    c[CC_tau_tr] = 180.0;
    c[CC_F] = 96.4867; c[CC_R] = 8.3143; c[CC_T] = 310.0;
    c[CC_RTF] = c[CC_R] * c[CC_T] / c[CC_F]; c[CC_FRT] = 1.0 / c[CC_RTF];
    c[CC_KQ10] = 3.0;
    c[CC_INaCa_max] = 1600.0; c[CC_KmCa] = 1.38; c[CC_KmNa] = 87.5; c[CC_gam] = 0.35;
    c[CC_ksat] = 0.1;
    c[CC_INaK_max] = 0.59933874; c[CC_KmKo] = 1.5; c[CC_KmNai] = 10.0;
    c[CC_sigma] = (exp(c[CC_Nao] / 67.3) - 1.0) / 7.0;
    c[CC_K_rel] = 30.0; c[CC_c1] = 3.4175e-13; c[CC_c2] = 1.367e-15; c[CC_u_tau] = 8.0;
    c[CC_gbCa] = 0.001131; c[CC_gbNa] = 0.0006744375; c[CC_gK1] = 0.09;
    c[CC_gKr] = 0.029411765; c[CC_gKs] = 0.12941176; c[CC_gKur_base] = 0.005;
    c[CC_gNa] = 7.8; c[CC_gto] = 0.1652;
    static const double y0[CR_NS] = {
        -81.9463303822041098, 13.8169746305367962, 136.355229902154434,
        1.23092247890489894e-04, 1.54668119199095355, 1.07650740580354909,
        2.56385228666526068e-03, 0.970298907063270155, 0.981123905023234988,
        2.91755626557170314e-02, 0.999342865333055497, 4.58838038240151104e-03,
        0.991468962753066063, 8.33819909884048389e-04, 1.86683180787284714e-02,
        1.24231529593716656e-04, 0.951907788168154578, 0.739682838459564729,
        -1.97647749727073971e-40, 1.0, 0.999233799248152699 };
    memcpy(y, y0, sizeof y0);
}

/* x / (1 - exp(-x/s)) with the generated code's guard |x| < eps -> s */
static double crn_xe(double x, double s, double eps)
{
    return fabs(x) < eps ? s : x / (1.0 - exp(-x / s));
}

static void crn_eval(double t, const double* c, const double* y, int solve_vm,
                     const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                     double* iion, double* istim)
{
    const double V = y[CR_V], Cai = y[CR_Cai], Nai = y[CR_Nai], Ki = y[CR_Ki];
    const double CaUp = y[CR_CaUp], CaRel = y[CR_CaRel];
    const double FRT = c[CC_FRT], RTF = c[CC_RTF];
    for (int i = 0; i < CR_NS; ++i) g[i] = 0;
#define CR_GATE(idx, inf, tau) do { r[idx] = ((inf) - y[idx]) / (tau); g[idx] = 1; \
                                    xi[idx] = (inf); ta[idx] = (tau); } while (0)
    /* reversal potentials */
    const double ECa = 0.5 * RTF * log(c[CC_Cao] / Cai);
    const double EK = RTF * log(c[CC_Ko] / Ki);
    const double ENa = RTF * log(c[CC_Nao] / Nai);

    /* INa: m, h, j (alpha/beta form; h and j alphas vanish above -40 mV) */
    const double am = V == -47.13 ? 3.2 : 0.32 * (V + 47.13) / (1.0 - exp(-0.1 * (V + 47.13)));
    const double bm = 0.08 * exp(-V / 11.0);
    CR_GATE(CR_m, am / (am + bm), 1.0 / (am + bm));
    const int lo = V < -40.0;
    const double ah = lo ? 0.135 * exp((V + 80.0) / -6.8) : 0.0;
    const double bh = lo ? 3.56 * exp(0.079 * V) + 310000.0 * exp(0.35 * V)
                         : 1.0 / (0.13 * (1.0 + exp((V + 10.66) / -11.1)));
    CR_GATE(CR_h, ah / (ah + bh), 1.0 / (ah + bh));
    const double aj = lo ? (-127140.0 * exp(0.2444 * V) - 3.474e-05 * exp(-0.04391 * V)) *
                               (V + 37.78) / (1.0 + exp(0.311 * (V + 79.23)))
                         : 0.0;
    const double bj = lo ? 0.1212 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)))
                         : 0.3 * exp(-2.535e-07 * V) / (1.0 + exp(-0.1 * (V + 32.0)));
    CR_GATE(CR_j, aj / (aj + bj), 1.0 / (aj + bj));
    const double INa = c[CC_gNa] * pow(y[CR_m], 3.0) * y[CR_h] * y[CR_j] * (V - ENa);

    /* Ito: oa, oi (temperature factor KQ10 on both time constants) */
    {
        const double a = 0.65 / (exp((V + 10.0) / -8.5) + exp((V - 30.0) / -59.0));
        const double b = 0.65 / (2.5 + exp((V + 82.0) / 17.0));
        CR_GATE(CR_oa, 1.0 / (1.0 + exp((V + 20.47) / -17.54)), 1.0 / (a + b) / c[CC_KQ10]);
        const double a2 = 1.0 / (18.53 + exp((V + 113.7) / 10.95));
        const double b2 = 1.0 / (35.56 + exp((V + 1.26) / -7.44));
        CR_GATE(CR_oi, 1.0 / (1.0 + exp((V + 43.1) / 5.3)), 1.0 / (a2 + b2) / c[CC_KQ10]);
    }
    const double Ito = c[CC_gto] * pow(y[CR_oa], 3.0) * y[CR_oi] * (V - EK);

    /* IKur: ua, ui; the conductance rises sigmoidally with V */
    {
        const double a = 0.65 / (exp((V + 10.0) / -8.5) + exp((V - 30.0) / -59.0));
        const double b = 0.65 / (2.5 + exp((V + 82.0) / 17.0));
        CR_GATE(CR_ua, 1.0 / (1.0 + exp((V + 30.3) / -9.6)), 1.0 / (a + b) / c[CC_KQ10]);
        const double a2 = 1.0 / (21.0 + exp((V - 185.0) / -28.0));
        const double b2 = 1.0 / exp((V - 158.0) / -16.0);
        CR_GATE(CR_ui, 1.0 / (1.0 + exp((V - 99.45) / 27.48)), 1.0 / (a2 + b2) / c[CC_KQ10]);
    }
    const double gKur = c[CC_gKur_base] * (1.0 + 10.0 / (1.0 + exp((V - 15.0) / -13.0)));
    const double IKur = gKur * pow(y[CR_ua], 3.0) * y[CR_ui] * (V - EK);

    /* IKr: xr, with inward rectification */
    {
        const double a = 0.0003 * crn_xe(V + 14.1, 5.0, 1e-6);
        const double xb = V - 3.3328;
        const double b = 7.3898e-05 * (fabs(xb) < 1e-07 ? 5.1237 : xb / (exp(xb / 5.1237) - 1.0));
        CR_GATE(CR_xr, 1.0 / (1.0 + exp((V + 14.1) / -6.5)), 1.0 / (a + b));
    }
    const double IKr = c[CC_gKr] * y[CR_xr] * (V - EK) / (1.0 + exp((V + 15.0) / 22.4));

    /* IKs: xs, squared */
    {
        const double a = 4e-05 * crn_xe(V - 19.9, 17.0, 1e-6);
        const double xb = V - 19.9;
        const double b = 3.5e-05 * (fabs(xb) < 1e-06 ? 9.0 : xb / (exp(xb / 9.0) - 1.0));
        CR_GATE(CR_xs, 1.0 / sqrt(1.0 + exp((V - 19.9) / -12.7)), 0.5 / (a + b));
    }
    const double IKs = c[CC_gKs] * pow(y[CR_xs], 2.0) * (V - EK);

    /* ICaL: d, f, fCa; fixed reversal potential ECaL */
    {
        const double xd = V + 10.0;
        const double td = fabs(xd) < 1e-06 ? 1.0 / (6.24 * 2.0 * 0.035)
                          : (1.0 - exp(xd / -6.24)) / (0.035 * xd * (1.0 + exp(xd / -6.24)));
        CR_GATE(CR_d, 1.0 / (1.0 + exp(xd / -8.0)), td);
        CR_GATE(CR_f, 1.0 / (1.0 + exp((V + 28.0) / 6.9)),
                9.0 / (0.0197 * exp(-pow(0.0337, 2.0) * pow(V + 10.0, 2.0)) + 0.02));
        CR_GATE(CR_fCa, 1.0 / (1.0 + Cai / 0.00035), c[CC_fCa_tau]);
    }
    const double ICaL = c[CC_gCaL] * y[CR_d] * y[CR_f] * y[CR_fCa] * (V - c[CC_ECaL]);

    /* pumps, exchanger and background currents */
    const double eg = exp(c[CC_gam] * V * FRT), eg1 = exp((c[CC_gam] - 1.0) * V * FRT);
    const double INaCa = c[CC_INaCa_max] *
                         (eg * pow(Nai, 3.0) * c[CC_Cao] - eg1 * pow(c[CC_Nao], 3.0) * Cai) /
                         ((pow(c[CC_KmNa], 3.0) + pow(c[CC_Nao], 3.0)) * (c[CC_KmCa] + c[CC_Cao]) *
                          (1.0 + c[CC_ksat] * eg1));
    const double fNaK = 1.0 / (1.0 + 0.1245 * exp(-0.1 * V * FRT) +
                               0.0365 * c[CC_sigma] * exp(-V * FRT));
    const double INaK = c[CC_INaK_max] * fNaK * c[CC_Ko] / (c[CC_Ko] + c[CC_KmKo]) /
                        (1.0 + pow(c[CC_KmNai] / Nai, 1.5));
    const double IpCa = c[CC_IpCa_max] * Cai / (0.0005 + Cai);
    const double IbCa = c[CC_gbCa] * (V - ECa);
    const double IbNa = c[CC_gbNa] * (V - ENa);
    const double IK1 = c[CC_gK1] * (V - EK) / (1.0 + exp(0.07 * (V + 80.0)));

    /* SR release (u, v, w) driven by the flux signal Fn */
    const double Irel = c[CC_K_rel] * pow(y[CR_u], 2.0) * y[CR_v] * y[CR_w] * (CaRel - Cai);
    {
        const double xw = V - 7.9;
        const double tw = 6.0 * (fabs(xw) < 1e-06 ? 2.0 / 13.0
                                 : (1.0 - exp(-xw / 5.0)) / ((1.0 + 0.3 * exp(-xw / 5.0)) * xw));
        CR_GATE(CR_w, 1.0 - 1.0 / (1.0 + exp(-(V - 40.0) / 17.0)), tw);
        const double Fn = 1e-12 * c[CC_V_rel] * Irel -
                          5e-13 / c[CC_F] * (0.5 * ICaL - 0.2 * INaCa) * c[CC_Cm];
        const double eu = exp(-(Fn - c[CC_c1]) / c[CC_c2]);
        CR_GATE(CR_u, 1.0 / (1.0 + eu), c[CC_u_tau]);
        CR_GATE(CR_v, 1.0 - 1.0 / (1.0 + exp(-(Fn - 0.2 * c[CC_c1]) / c[CC_c2])),
                1.91 + 2.09 / (1.0 + eu));
    }
#undef CR_GATE

    /* calcium handling */
    const double Iup = c[CC_I_up_max] / (1.0 + c[CC_K_up] / Cai);
    const double Ileak = c[CC_I_up_max] * CaUp / c[CC_Ca_up_max];
    const double Itr = (CaUp - CaRel) / c[CC_tau_tr];
    r[CR_CaRel] = (Itr - Irel) /
                  (1.0 + c[CC_CSQN_max] * c[CC_Km_CSQN] / pow(CaRel + c[CC_Km_CSQN], 2.0));
    r[CR_CaUp] = Iup - (Ileak + Itr * c[CC_V_rel] / c[CC_V_up]);
    const double B1 = (2.0 * INaCa - (IpCa + ICaL + IbCa)) * c[CC_Cm] / (2.0 * c[CC_V_i] * c[CC_F]) +
                      (c[CC_V_up] * (Ileak - Iup) + Irel * c[CC_V_rel]) / c[CC_V_i];
    const double B2 = 1.0 + c[CC_TRPN_max] * c[CC_Km_TRPN] / pow(Cai + c[CC_Km_TRPN], 2.0) +
                      c[CC_CMDN_max] * c[CC_Km_CMDN] / pow(Cai + c[CC_Km_CMDN], 2.0);
    r[CR_Cai] = B1 / B2;

    const double Iion = INa + IK1 + Ito + IKur + IKr + IKs + ICaL + IpCa + INaK + INaCa + IbNa + IbCa;
    const double cvf = c[CC_Cm] / (c[CC_V_i] * c[CC_F]);
    r[CR_Ki] = (2.0 * INaK - (IK1 + Ito + IKur + IKr + IKs)) * cvf;
    r[CR_Nai] = (-3.0 * INaK - (3.0 * INaCa + IbNa + INa)) * cvf;
    const double Ist = orc_stimulus(t, st);
    r[CR_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion; *istim = Ist;
}

/* ========================================================= PerisYague_2022
 * Peris-Yague et al. 2022 human atrial model with chloride, 22 states: constants and
 * initial state of PerisYague_2022.H:211-294, right-hand side of :296-445, state order
 * of its STATES_NAMES table (:24-47).  Selected by the key "PerisYague_2022"
 * (PerisYague.H:124); tissue "myocyte" only (PerisYague.C:92-95).  Differences from
 * Courtemanche that the restatement keeps as the reference has them: the h-gate beta
 * above -40 mV is 0.13/(1 + exp(-(V + 10.66)/11.1)) (:333); the d-gate time constant's
 * guard at V = -5 mV returns 1.0, not the expression's limit (:368); the IKr and IKs
 * alpha/beta quotients have no guard (0/0 at V = -14.1, 3.3328, 18.80816 mV exactly);
 * IbK (gbK = 0) enters neither Iion nor the K_i balance; the K_i balance has no stimulus.
 * Parity pins (tests/test_oracle_models.py): charge conservation including the chloride
 * balance, continuity at the guarded singularities of m and w, and the rate limits of the
 * unguarded quotients approached from both sides. */
enum { PY_V, PY_Nai, PY_Ki, PY_Cli, PY_Cai, PY_CaUp, PY_CaRel, PY_m, PY_h, PY_j, PY_ua,
       PY_uif, PY_uis, PY_xr, PY_xs, PY_d, PY_f, PY_fCa, PY_qCa, PY_u, PY_v, PY_w, PY_NS };
enum { PC_R, PC_T, PC_F, PC_Cm, PC_V_i, PC_V_up, PC_V_rel, PC_Ko, PC_Nao, PC_Cao, PC_Clo,
       PC_gNa, PC_gK1, PC_gClCa, PC_gKur_amp, PC_gKr, PC_gKs, PC_gCaL, PC_gbCa, PC_gbNa,
       PC_gbK, PC_INaK_max, PC_INaCa_max, PC_IpCa_max, PC_I_up_max, PC_KQ10, PC_gam,
       PC_KmNai, PC_KmKo, PC_KmNa, PC_KmCa, PC_ksat, PC_krel, PC_K_up, PC_Ca_up_max,
       PC_CMDN_max, PC_TRPN_max, PC_CSQN_max, PC_Km_CMDN, PC_Km_TRPN, PC_Km_CSQN,
       PC_tau_tr, PC_c1, PC_c2, PC_u_tau, PC_fCa_tau, PC_RTF, PC_FRT, PC_sigma, PC_NC };

static void pys_init(double* c, double* y)
{
    c[PC_R] = 8.3143; c[PC_T] = 310.0; c[PC_F] = 96.4867; c[PC_Cm] = 100.0;
    c[PC_V_i] = 13668.0; c[PC_V_up] = 1109.52; c[PC_V_rel] = 96.48;
    c[PC_Ko] = 5.4; c[PC_Nao] = 140.0; c[PC_Cao] = 1.8; c[PC_Clo] = 132.0;
    c[PC_gNa] = 13.99; c[PC_gK1] = 0.08218; c[PC_gClCa] = 0.15731; c[PC_gKur_amp] = 0.45539;
    c[PC_gKr] = 0.01730; c[PC_gKs] = 0.0594; c[PC_gCaL] = 0.06574; c[PC_gbCa] = 0.00113;
    c[PC_gbNa] = 0.000674; c[PC_gbK] = 0.0;
    c[PC_INaK_max] = 0.94935; c[PC_INaCa_max] = 2304.0; c[PC_IpCa_max] = 0.275;
    c[PC_I_up_max] = 0.005; c[PC_KQ10] = 3.0; c[PC_gam] = 0.35; c[PC_KmNai] = 10.0;
    c[PC_KmKo] = 1.5; c[PC_KmNa] = 87.5; c[PC_KmCa] = 1.38; c[PC_ksat] = 0.1;
    c[PC_krel] = 30.0; c[PC_K_up] = 0.00092; c[PC_Ca_up_max] = 15.0;
    c[PC_CMDN_max] = 0.05; c[PC_TRPN_max] = 0.07; c[PC_CSQN_max] = 10.0;
    c[PC_Km_CMDN] = 0.00238; c[PC_Km_TRPN] = 0.0005; c[PC_Km_CSQN] = 0.8;
    c[PC_tau_tr] = 180.0; c[PC_c1] = 1.367e-13; c[PC_c2] = 13.67e-16; c[PC_u_tau] = 8.0;
    c[PC_fCa_tau] = 2.0;
    c[PC_RTF] = c[PC_R] * c[PC_T] / c[PC_F]; c[PC_FRT] = 1.0 / c[PC_RTF];
    c[PC_sigma] = (exp(c[PC_Nao] / 67.3) - 1.0) / 7.0;
    static const double y0[PY_NS] = {
        -76.0, 11.17, 139.0, 30.0, 0.0001013, 1.488, 1.488, 2.91e-3, 0.965, 0.978, 4.96e-3,
        1.0, 1.0, 3.29e-5, 1.87e-2, 1.37e-4, 0.999, 0.775, 0.0, 0.0235, 1.0, 0.999 };
    memcpy(y, y0, sizeof y0);
}

static void pys_eval(double t, const double* c, const double* y, int solve_vm,
                     const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                     double* iion, double* istim)
{
    const double V = y[PY_V], Cai = y[PY_Cai], Nai = y[PY_Nai], Ki = y[PY_Ki];
    const double Cli = y[PY_Cli], CaRel = y[PY_CaRel], CaUp = y[PY_CaUp];
    const double RTF = c[PC_RTF], FRT = c[PC_FRT];
    for (int i = 0; i < PY_NS; ++i) g[i] = 0;
#define PY_GATE(idx, inf, tau) do { r[idx] = ((inf) - y[idx]) / (tau); g[idx] = 1; \
                                    xi[idx] = (inf); ta[idx] = (tau); } while (0)
    const double ENa = RTF * log(c[PC_Nao] / Nai);
    const double EK = RTF * log(c[PC_Ko] / Ki);
    const double ECa = 0.5 * RTF * log(c[PC_Cao] / Cai);
    const double ECl = -RTF * log(c[PC_Clo] / Cli);

    /* INa (time constants scaled by 1.7, 2, 2) */
    const double am = fabs(V + 47.13) < 1e-6 ? 3.2 : 0.32 * (V + 47.13) / (1.0 - exp(-0.1 * (V + 47.13)));
    const double bm = 0.08 * exp(-V / 11.0);
    PY_GATE(PY_m, am / (am + bm), 1.7 / (am + bm));
    const int lo = V < -40.0;
    const double ah = lo ? 0.135 * exp(-(V + 80.0) / 6.8) : 0.0;
    const double bh = lo ? 3.56 * exp(0.079 * V) + 3.1e5 * exp(0.35 * V)
                         : 0.13 / (1.0 + exp(-(V + 10.66) / 11.1));
    PY_GATE(PY_h, ah / (ah + bh), 2.0 / (ah + bh));
    const double aj = lo ? (-127140.0 * exp(0.2444 * V) - 3.474e-5 * exp(-0.04391 * V)) *
                               (V + 37.78) / (1.0 + exp(0.311 * (V + 79.23)))
                         : 0.0;
    const double bj = lo ? 0.1212 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)))
                         : 0.3 * exp(-2.535e-7 * V) / (1.0 + exp(-0.1 * (V + 32.0)));
    PY_GATE(PY_j, aj / (aj + bj), 2.0 / (aj + bj));
    const double INa = c[PC_gNa] * pow(y[PY_m], 3.0) * y[PY_h] * y[PY_j] * (V - ENa);

    const double IK1 = c[PC_gK1] * (V - EK - 5.0) / (1.0 + exp(0.063 * (V + 70.0)));

    /* IKr: xr (steady state from alpha/beta) */
    const double axr = 0.0003 * (V + 14.1) / (1.0 - exp(-(V + 14.1) / 5.0));
    const double bxr = 7.3898e-5 * (V - 3.3328) / (exp((V - 3.3328) / 5.1237) - 1.0);
    PY_GATE(PY_xr, axr / (axr + bxr), 1.0 / (axr + bxr));
    const double IKr = c[PC_gKr] * y[PY_xr] * (V - EK) / (1.0 + exp((V + 15.0) / 22.4));

    /* IKs: xs, rates around 18.80816 mV, steady state around 19.9 mV */
    const double axs = 4e-5 * (V - 18.80816) / (1.0 - exp(-(V - 18.80816) / 17.0));
    const double bxs = 3.5e-5 * (V - 18.80816) / (exp((V - 18.80816) / 9.0) - 1.0);
    PY_GATE(PY_xs, 1.0 / sqrt(1.0 + exp(-(V - 19.9) / 12.7)), 0.5 / (axs + bxs));
    const double IKs = c[PC_gKs] * pow(y[PY_xs], 2.0) * (V - EK);

    /* IKur: ua, fast and slow inactivation */
    const double gKur = c[PC_gKur_amp] * (0.005 + 0.05 / (1.0 + exp(-(V - 15.0) / 13.0)));
    const double aua = 0.65 / (exp(-(V + 10.0) / 8.5) + exp(-(V - 30.0) / 59.0));
    const double bua = 0.65 / (2.5 + exp((V + 82.0) / 17.0));
    PY_GATE(PY_ua, 1.0 / (1.0 + exp(-(V + 30.3) / 9.6)), 1.0 / (c[PC_KQ10] * (aua + bua)));
    const double uinf = 1.0 / (1.0 + exp((V + 17.5) / 5.85));
    PY_GATE(PY_uif, uinf, 400.0 + 1068.0 * exp(-pow(V / 50.0, 2.0)));
    PY_GATE(PY_uis, uinf, 2000.0 + 60000.0 * exp(-pow((V + 39.3) / 30.0, 2.0)));
    const double IKur = gKur * pow(y[PY_ua], 3.0) * (0.25 * y[PY_uif] + 0.75 * y[PY_uis]) * (V - EK);

    /* ICaL */
    const double td = fabs(V + 5.0) < 1e-6 ? 1.0
                      : (1.0 - exp(-(V + 5.0) / 6.24)) /
                            (0.035 * (V + 5.0) * (1.0 + exp(-(V + 5.0) / 6.24)));
    PY_GATE(PY_d, 1.0 / (1.0 + exp(-(V + 5.0) / 8.0)), td);
    PY_GATE(PY_f, 1.0 / (1.0 + exp((V + 28.0) / 6.9)),
            9.0 / (0.0197 * exp(-0.03372 * pow(V + 5.0, 2.0)) + 0.02));
    PY_GATE(PY_fCa, 1.0 / (1.0 + Cai / 0.00035), c[PC_fCa_tau]);
    const double ICaL = c[PC_gCaL] * y[PY_d] * y[PY_f] * y[PY_fCa] * (V - 65.0);

    /* pumps, exchanger, background */
    const double fNaK = 1.0 / (1.0 + 0.1245 * exp(-0.1 * V * FRT) + 0.0365 * c[PC_sigma] * exp(-V * FRT));
    const double INaK = c[PC_INaK_max] * fNaK * c[PC_Ko] / (c[PC_Ko] + c[PC_KmKo]) /
                        (1.0 + pow(c[PC_KmNai] / Nai, 1.5));
    const double eg1 = exp((c[PC_gam] - 1.0) * V * FRT);
    const double INaCa = c[PC_INaCa_max] *
                         (exp(c[PC_gam] * V * FRT) * pow(Nai, 3.0) * c[PC_Cao] - eg1 * pow(c[PC_Nao], 3.0) * Cai) /
                         ((pow(c[PC_KmNa], 3.0) + pow(c[PC_Nao], 3.0)) * (c[PC_KmCa] + c[PC_Cao]) *
                          (1.0 + c[PC_ksat] * eg1));
    const double IbNa = c[PC_gbNa] * (V - ENa);
    const double IbCa = c[PC_gbCa] * (V - ECa);
    const double IpCa = c[PC_IpCa_max] * Cai / (0.0005 + Cai);

    /* SR release and the Ca-activated chloride current, both driven by Fn */
    const double Irel = c[PC_krel] * pow(y[PY_u], 2.0) * y[PY_v] * y[PY_w] * (CaRel - Cai);
    const double Itr = (CaUp - CaRel) / c[PC_tau_tr];
    const double Iup = c[PC_I_up_max] / (1.0 + c[PC_K_up] / Cai);
    const double Ileak = c[PC_I_up_max] * CaUp / c[PC_Ca_up_max];
    const double Fn = 1e-12 * c[PC_V_rel] * Irel - 5e-13 / c[PC_F] * (0.5 * ICaL - 0.2 * INaCa) * c[PC_Cm];
    PY_GATE(PY_qCa, 1.0 - pow(1.0 / (1.0 + Fn / 1.1e-10), 3.0), 2.0);
    const double IClCa = c[PC_gClCa] * y[PY_qCa] * (V - ECl);
    {
        const double xw = V - 7.9;
        const double tw = 6.0 * (fabs(xw) < 1e-6 ? 2.0 / 13.0
                                 : (1.0 - exp(-xw / 5.0)) / ((1.0 + 0.3 * exp(-xw / 5.0)) * xw));
        PY_GATE(PY_w, 1.0 - 1.0 / (1.0 + exp(-(V - 40.0) / 17.0)), tw);
        const double eu = exp(-(Fn - c[PC_c1]) / c[PC_c2]);
        PY_GATE(PY_u, 1.0 / (1.0 + eu), c[PC_u_tau]);
        PY_GATE(PY_v, 1.0 - 1.0 / (1.0 + exp(-(Fn - 0.2 * c[PC_c1]) / c[PC_c2])),
                1.91 + 2.09 / (1.0 + eu));
    }
#undef PY_GATE

    const double B1 = (2.0 * INaCa - (IpCa + ICaL + IbCa)) * c[PC_Cm] / (2.0 * c[PC_V_i] * c[PC_F]) +
                      (c[PC_V_up] * (Ileak - Iup) + Irel * c[PC_V_rel]) / c[PC_V_i];
    const double B2 = 1.0 + c[PC_TRPN_max] * c[PC_Km_TRPN] / pow(Cai + c[PC_Km_TRPN], 2.0) +
                      c[PC_CMDN_max] * c[PC_Km_CMDN] / pow(Cai + c[PC_Km_CMDN], 2.0);
    r[PY_Cai] = B1 / B2;
    r[PY_CaRel] = (Itr - Irel) / (1.0 + c[PC_CSQN_max] * c[PC_Km_CSQN] / pow(CaRel + c[PC_Km_CSQN], 2.0));
    r[PY_CaUp] = Iup - (Ileak + Itr * c[PC_V_rel] / c[PC_V_up]);
    const double cvf = c[PC_Cm] / (c[PC_V_i] * c[PC_F]);
    r[PY_Nai] = (-3.0 * INaK - (3.0 * INaCa + IbNa + INa)) * cvf;
    r[PY_Ki] = (2.0 * INaK - (IK1 + IKur + IKr + IKs)) * cvf;
    r[PY_Cli] = IClCa * cvf;

    const double Iion = INa + IK1 + IKur + IKr + IKs + ICaL + IpCa + INaK + INaCa + IbNa + IbCa + IClCa;
    const double Ist = orc_stimulus(t, st);
    r[PY_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion; *istim = Ist;
}

/* ================================================================= Stewart
 * Stewart et al. 2009 human Purkinje-fibre model, 20 states: constants and initial state
 * of Stewart_2009.H:262-377, right-hand side of :380-530, state order of its
 * STATES_NAMES table.  Tissue "myocyte" only (Stewart.C:93-96).  In tissue mode the
 * reference leaves RATES[V] unwritten (:524-528); here dV/dt = 0 there, as for the other
 * models.  R_prime (dR/dt = k4 - (k2 Ca_ss + k4) R) is integrated as a linear gate.
 * Parity pins (tests/test_oracle_models.py): charge conservation over the cytosolic,
 * subspace and SR compartments; the GHK quotient of ICaL continuous through V = 15 mV
 * from both sides; Rush-Larsen and RKF45 agreeing on the automatic rhythm from the
 * reference's initial state; the reference's generated expressions at seeded states
 * (tests/test_oracle_model_vectors.py). */
enum { SW_V, SW_y, SW_Xr1, SW_Xr2, SW_Xs, SW_m, SW_h, SW_j, SW_d, SW_f, SW_f2, SW_fCass,
       SW_s, SW_r, SW_Cai, SW_CaSR, SW_Cass, SW_Rp, SW_Nai, SW_Ki, SW_NS };
enum { SC_KpCa, SC_gpCa, SC_gCaL, SC_gbca, SC_Bufc, SC_Bufsr, SC_Bufss, SC_Cao, SC_EC,
       SC_Kbufc, SC_Kbufsr, SC_Kbufss, SC_Kup, SC_Vleak, SC_Vrel, SC_Vsr, SC_Vss, SC_Vxfer,
       SC_Vmaxup, SC_k1p, SC_k2p, SC_k3, SC_k4, SC_maxsr, SC_minsr, SC_gNa, SC_gfK, SC_gfNa,
       SC_gK1, SC_Cm, SC_F, SC_R, SC_T, SC_Vc, SC_Ko, SC_gpK, SC_gKr, SC_Pkna, SC_gKs,
       SC_gbna, SC_KNaCa, SC_Ksat, SC_KmCa, SC_KmNai, SC_alpha, SC_gam, SC_Nao, SC_KmNa,
       SC_Kmk, SC_PNaK, SC_gsus, SC_gto, SC_NC };

static void stw_init(double* c, double* y)
{
    c[SC_KpCa] = 0.0005; c[SC_gpCa] = 0.1238; c[SC_gCaL] = 3.98e-05; c[SC_gbca] = 0.000592;
    c[SC_Bufc] = 0.2; c[SC_Bufsr] = 10.0; c[SC_Bufss] = 0.4; c[SC_Cao] = 2.0; c[SC_EC] = 1.5;
    c[SC_Kbufc] = 0.001; c[SC_Kbufsr] = 0.3; c[SC_Kbufss] = 0.00025; c[SC_Kup] = 0.00025;
    c[SC_Vleak] = 0.00036; c[SC_Vrel] = 0.102; c[SC_Vsr] = 0.001094; c[SC_Vss] = 5.468e-05;
    c[SC_Vxfer] = 0.0038; c[SC_Vmaxup] = 0.006375; c[SC_k1p] = 0.15; c[SC_k2p] = 0.045;
    c[SC_k3] = 0.06; c[SC_k4] = 0.005; c[SC_maxsr] = 2.5; c[SC_minsr] = 1.0;
    c[SC_gNa] = 130.5744; c[SC_gfK] = 0.0234346; c[SC_gfNa] = 0.0145654; c[SC_gK1] = 0.065;
    c[SC_Cm] = 0.185; c[SC_F] = 96485.3415; c[SC_R] = 8314.472; c[SC_T] = 310.0;
    c[SC_Vc] = 0.016404; c[SC_Ko] = 5.4; c[SC_gpK] = 0.0146; c[SC_gKr] = 0.0918;
    c[SC_Pkna] = 0.03; c[SC_gKs] = 0.2352; c[SC_gbna] = 0.00029; c[SC_KNaCa] = 1000.0;
    c[SC_Ksat] = 0.1; c[SC_KmCa] = 1.38; c[SC_KmNai] = 87.5; c[SC_alpha] = 2.5;
    c[SC_gam] = 0.35; c[SC_Nao] = 140.0; c[SC_KmNa] = 40.0; c[SC_Kmk] = 1.0;
    c[SC_PNaK] = 2.724; c[SC_gsus] = 0.0227; c[SC_gto] = 0.08184;
    static const double y0[SW_NS] = {
        -69.1370441635924067, 4.57562667986601973e-02, 5.50281999719088016e-03,
        0.313213286437995009, 9.53708522974788982e-03, 4.17391656294996971e-02,
        0.190678733735144990, 0.238219836154029002, 2.87906256206415015e-04,
        0.989328560287986991, 0.995474890442184956, 0.999955429598213041,
        0.963861017995010005, 1.03618091196912006e-03, 1.01878186157052002e-04,
        3.10836886659417022, 4.46818714055411022e-04, 0.991580051907845039,
        8.80420286531673035, 136.781894160227012 };
    memcpy(y, y0, sizeof y0);
}

static void stw_eval(double t, const double* c, const double* y, int solve_vm,
                     const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                     double* iion, double* istim)
{
    const double V = y[SW_V];
    const double RTF = c[SC_R] * c[SC_T] / c[SC_F], FRT = c[SC_F] / (c[SC_R] * c[SC_T]);
    for (int i = 0; i < SW_NS; ++i) g[i] = 0;
#define SW_GATE(idx, inf, tau) do { r[idx] = ((inf) - y[idx]) / (tau); g[idx] = 1; \
                                    xi[idx] = (inf); ta[idx] = (tau); } while (0)
    /* ICaL gates d, f, f2, fCass */
    {
        const double ad = 1.4 / (1.0 + exp((-35.0 - V) / 13.0)) + 0.25;
        const double bd = 1.4 / (1.0 + exp((V + 5.0) / 5.0));
        const double gd = 1.0 / (1.0 + exp((50.0 - V) / 20.0));
        SW_GATE(SW_d, 1.0 / (1.0 + exp((-8.0 - V) / 7.5)), ad * bd + gd);
        SW_GATE(SW_f2, 0.67 / (1.0 + exp((V + 35.0) / 7.0)) + 0.33,
                562.0 * exp(-pow(V + 27.0, 2.0) / 240.0) + 31.0 / (1.0 + exp((25.0 - V) / 10.0)) +
                    80.0 / (1.0 + exp((V + 30.0) / 10.0)));
        const double cq = 1.0 + pow(y[SW_Cass] / 0.05, 2.0);
        SW_GATE(SW_fCass, 0.6 / cq + 0.4, 80.0 / cq + 2.0);
        SW_GATE(SW_f, 1.0 / (1.0 + exp((V + 20.0) / 7.0)),
                1102.5 * exp(-pow(V + 27.0, 2.0) / 225.0) + 200.0 / (1.0 + exp((13.0 - V) / 10.0)) +
                    180.0 / (1.0 + exp((V + 30.0) / 10.0)) + 20.0);
    }
    /* INa gates m, h, j (TNNP form) */
    {
        const int lo = V < -40.0;
        const double ah = lo ? 0.057 * exp(-(V + 80.0) / 6.8) : 0.0;
        const double bh = lo ? 2.7 * exp(0.079 * V) + 310000.0 * exp(0.3485 * V)
                             : 0.77 / (0.13 * (1.0 + exp((V + 10.66) / -11.1)));
        const double hinf = 1.0 / pow(1.0 + exp((V + 71.55) / 7.43), 2.0);
        SW_GATE(SW_h, hinf, 1.0 / (ah + bh));
        const double aj = lo ? (-25428.0 * exp(0.2444 * V) - 6.948e-06 * exp(-0.04391 * V)) *
                                   (V + 37.78) / (1.0 + exp(0.311 * (V + 79.23)))
                             : 0.0;
        const double bj = lo ? 0.02424 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)))
                             : 0.6 * exp(0.057 * V) / (1.0 + exp(-0.1 * (V + 32.0)));
        SW_GATE(SW_j, hinf, 1.0 / (aj + bj));
        const double am = 1.0 / (1.0 + exp((-60.0 - V) / 5.0));
        const double bm = 0.1 / (1.0 + exp((V + 35.0) / 5.0)) + 0.1 / (1.0 + exp((V - 50.0) / 200.0));
        SW_GATE(SW_m, 1.0 / pow(1.0 + exp((-56.86 - V) / 9.03), 2.0), am * bm);
    }
    /* If gate y, IKr Xr1 Xr2, IKs Xs, Ito r s */
    SW_GATE(SW_y, 1.0 / (1.0 + exp((V + 80.6) / 6.8)),
            4000.0 / (exp(-2.9 - 0.04 * V) + exp(3.6 + 0.11 * V)));
    SW_GATE(SW_Xr1, 1.0 / (1.0 + exp((-26.0 - V) / 7.0)),
            450.0 / (1.0 + exp((-45.0 - V) / 10.0)) * (6.0 / (1.0 + exp((V + 30.0) / 11.5))));
    SW_GATE(SW_Xr2, 1.0 / (1.0 + exp((V + 88.0) / 24.0)),
            3.0 / (1.0 + exp((-60.0 - V) / 20.0)) * (1.12 / (1.0 + exp((V - 60.0) / 20.0))));
    SW_GATE(SW_Xs, 1.0 / (1.0 + exp((-5.0 - V) / 14.0)),
            1400.0 / sqrt(1.0 + exp((5.0 - V) / 6.0)) * (1.0 / (1.0 + exp((V - 35.0) / 15.0))) + 80.0);
    SW_GATE(SW_r, 1.0 / (1.0 + exp((20.0 - V) / 13.0)),
            10.45 * exp(-pow(V + 40.0, 2.0) / 1800.0) + 7.3);
    SW_GATE(SW_s, 1.0 / (1.0 + exp((V + 27.0) / 13.0)),
            85.0 * exp(-pow(V + 25.0, 2.0) / 320.0) + 5.0 / (1.0 + exp((V - 40.0) / 5.0)) + 42.0);

    /* calcium handling: subspace release through the RyR state R' */
    const double Cai = y[SW_Cai], CaSR = y[SW_CaSR], Cass = y[SW_Cass];
    const double bufc = 1.0 / (1.0 + c[SC_Bufc] * c[SC_Kbufc] / pow(Cai + c[SC_Kbufc], 2.0));
    const double bufsr = 1.0 / (1.0 + c[SC_Bufsr] * c[SC_Kbufsr] / pow(CaSR + c[SC_Kbufsr], 2.0));
    const double bufss = 1.0 / (1.0 + c[SC_Bufss] * c[SC_Kbufss] / pow(Cass + c[SC_Kbufss], 2.0));
    const double Ileak = c[SC_Vleak] * (CaSR - Cai);
    const double Iup = c[SC_Vmaxup] / (1.0 + pow(c[SC_Kup], 2.0) / pow(Cai, 2.0));
    const double Ixfer = c[SC_Vxfer] * (Cass - Cai);
    const double kcasr = c[SC_maxsr] - (c[SC_maxsr] - c[SC_minsr]) / (1.0 + pow(c[SC_EC] / CaSR, 2.0));
    const double k1 = c[SC_k1p] / kcasr, k2 = c[SC_k2p] * kcasr;
    const double O = k1 * pow(Cass, 2.0) * y[SW_Rp] / (c[SC_k3] + k1 * pow(Cass, 2.0));
    /* dR'/dt = k4 - (k2 Ca_ss + k4) R' */
    {
        const double b = k2 * Cass + c[SC_k4];
        SW_GATE(SW_Rp, c[SC_k4] / b, 1.0 / b);
    }
    const double Irel = c[SC_Vrel] * O * (CaSR - Cass);
    r[SW_CaSR] = bufsr * (Iup - (Irel + Ileak));

    /* currents */
    const double EK = RTF * log(c[SC_Ko] / y[SW_Ki]);
    const double ENa = RTF * log(c[SC_Nao] / y[SW_Nai]);
    const double ECa = 0.5 * RTF * log(c[SC_Cao] / Cai);
    const double EKs = RTF * log((c[SC_Ko] + c[SC_Pkna] * c[SC_Nao]) / (y[SW_Ki] + c[SC_Pkna] * y[SW_Nai]));
    const double INaK = c[SC_PNaK] * c[SC_Ko] / (c[SC_Ko] + c[SC_Kmk]) * y[SW_Nai] /
                        (y[SW_Nai] + c[SC_KmNa]) /
                        (1.0 + 0.1245 * exp(-0.1 * V * FRT) + 0.0353 * exp(-V * FRT));
    const double Isus = c[SC_gsus] * (1.0 / (1.0 + exp((5.0 - V) / 17.0))) * (V - EK);
    const double Ito = c[SC_gto] * y[SW_r] * y[SW_s] * (V - EK);
    const double e2 = exp(2.0 * (V - 15.0) * FRT);
    const double ICaL = c[SC_gCaL] * y[SW_d] * y[SW_f] * y[SW_f2] * y[SW_fCass] * 4.0 * (V - 15.0) *
                        pow(c[SC_F], 2.0) / (c[SC_R] * c[SC_T]) * (0.25 * Cass * e2 - c[SC_Cao]) / (e2 - 1.0);
    const double IbCa = c[SC_gbca] * (V - ECa);
    const double IfK = y[SW_y] * c[SC_gfK] * (V - EK);
    const double IfNa = y[SW_y] * c[SC_gfNa] * (V - ENa);
    const double IK1 = c[SC_gK1] * (1.0 / (1.0 + exp(0.1 * (V + 75.44)))) * (V - 8.0 - EK);
    const double IpK = c[SC_gpK] * (V - EK) / (1.0 + exp((25.0 - V) / 5.98));
    const double IKr = c[SC_gKr] * sqrt(c[SC_Ko] / 5.4) * y[SW_Xr1] * y[SW_Xr2] * (V - EK);
    const double IKs = c[SC_gKs] * pow(y[SW_Xs], 2.0) * (V - EKs);
    const double IpCa = c[SC_gpCa] * Cai / (Cai + c[SC_KpCa]);
    const double INaCa = c[SC_KNaCa] *
                         (exp(c[SC_gam] * V * FRT) * pow(y[SW_Nai], 3.0) * c[SC_Cao] -
                          exp((c[SC_gam] - 1.0) * V * FRT) * pow(c[SC_Nao], 3.0) * Cai * c[SC_alpha]) /
                         ((pow(c[SC_KmNai], 3.0) + pow(c[SC_Nao], 3.0)) * (c[SC_KmCa] + c[SC_Cao]) *
                          (1.0 + c[SC_Ksat] * exp((c[SC_gam] - 1.0) * V * FRT)));
    const double INa = c[SC_gNa] * pow(y[SW_m], 3.0) * y[SW_h] * y[SW_j] * (V - ENa);
    const double IbNa = c[SC_gbna] * (V - ENa);
#undef SW_GATE

    r[SW_Cai] = bufc * ((Ileak - Iup) * c[SC_Vsr] / c[SC_Vc] + Ixfer -
                        (IbCa + IpCa - 2.0 * INaCa) * c[SC_Cm] / (2.0 * c[SC_Vc] * c[SC_F]));
    r[SW_Cass] = bufss * (-ICaL * c[SC_Cm] / (2.0 * c[SC_Vss] * c[SC_F]) +
                          Irel * c[SC_Vsr] / c[SC_Vss] - Ixfer * c[SC_Vc] / c[SC_Vss]);
    const double cvf = c[SC_Cm] / (c[SC_Vc] * c[SC_F]);
    r[SW_Ki] = -(IK1 + Ito + IfK + Isus + IKr + IKs + IpK - 2.0 * INaK) * cvf;
    r[SW_Nai] = -(INa + IbNa + IfNa + 3.0 * INaK + 3.0 * INaCa) * cvf;
    const double Iion = IK1 + Ito + Isus + IKr + IKs + ICaL + INaK + INa + IbNa + INaCa + IbCa +
                        IpK + IpCa + IfNa + IfK;
    const double Ist = orc_stimulus(t, st);
    r[SW_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion; *istim = Ist;
}

/* ================================================================== Grandi
 * Grandi et al. 2011 human atrial myocyte (41 states): constants and initial state of
 * Grandi_2011.H:414-657 (GrandiinitConsts, with AF = ISO = RA = 0 as it sets them), right-
 * hand side of :660-931, state order of its STATES_INDEX enum.  Tissue "myocyte" only
 * (Grandi.C:93-96); V in mV (Grandi.C:121-124), Im = Iion_cm.  K_i is constant (its rate
 * is 0).  In tissue mode dV/dt is 0.
 * Rush-Larsen classification: the Hodgkin-Huxley gates (d f xr xs ikur_r ikur_s h j m hl
 * x y) as gates, the first-order kinetics linear in their own state (ml, fCaB_jn/sl, the
 * Ca and Na buffers, Csqn) as linear gates with the other states frozen, Ca_sr, Ca_jn and
 * Ca_sl linearised in themselves (below), the rest forward Euler.
 * Parity pins: the reference's own expressions at seeded states
 * (tests/test_oracle_model_vectors.py). */
enum { GR_V, GR_Cai, GR_Cajn, GR_Casl, GR_Casr, GR_Csqn, GR_d, GR_f, GR_fCaBjn, GR_fCaBsl,
       GR_xr, GR_xs, GR_kur_r, GR_kur_s, GR_h, GR_j, GR_m, GR_hl, GR_ml, GR_x, GR_y, GR_Ki,
       GR_ri, GR_ro, GR_rr, GR_Nai, GR_Najn, GR_Nasl, GR_CaM, GR_Myoc, GR_Myom, GR_SLHjn,
       GR_SLHsl, GR_SLLjn, GR_SLLsl, GR_SRB, GR_TnCHc, GR_TnCHm, GR_TnCL, GR_NaBjn, GR_NaBsl,
       GR_NS };
/* runtime constants (scalable per cell: gNa, gKr, pCa, gKs) and the geometry and
 * buffer constants the rates use */
enum { GC_gNa, GC_gKr, GC_pCa, GC_gKs, GC_gK1, GC_gto, GC_gKur, GC_gCaB, GC_gNaB, GC_gClB,
       GC_GClCa, GC_IbarNaK, GC_IbarNCX, GC_IbarSLCaP, GC_VmaxSRCaP, GC_ks, GC_pK, GC_pNa,
       GC_gKp, GC_gNaL, GC_Vjn, GC_Vsl, GC_Vmyo, GC_Vsr, GC_BmaxSLHj, GC_BmaxSLHsl,
       GC_BmaxSLLj, GC_BmaxSLLsl, GC_BmaxCsqn, GC_sigma, GC_ECl, GC_FRT, GC_NC };

static void grd_init(double* c, double* y)
{
    const double pi = 3.14159265358979312, len = 100.0, rad = 10.25;
    const double Vcell = pi * pow(rad, 2.0) * len * 1e-15;
    const double F = 96485.0, R = 8314.0, T = 310.0;
    c[GC_Vjn] = 0.0539 * 0.01 * Vcell;
    c[GC_Vmyo] = 0.65 * Vcell;
    c[GC_Vsl] = 0.02 * Vcell;
    c[GC_Vsr] = 0.035 * Vcell;
    c[GC_FRT] = F / R / T;
    c[GC_BmaxSLLj] = 0.0046 * c[GC_Vmyo] / c[GC_Vjn] * 0.1;
    c[GC_BmaxSLLsl] = c[GC_Vmyo] / c[GC_Vsl] * 0.0374;
    c[GC_BmaxSLHj] = 0.00165 * c[GC_Vmyo] / c[GC_Vjn] * 0.1;
    c[GC_BmaxSLHsl] = c[GC_Vmyo] / c[GC_Vsl] * 0.0134;
    c[GC_BmaxCsqn] = c[GC_Vmyo] / c[GC_Vsr] * 0.14;
    c[GC_sigma] = (exp(140.0 / 67.3) - 1.0) / 7.0;
    c[GC_ECl] = 1.0 / c[GC_FRT] * log(15.0 / 150.0);
    c[GC_gNa] = 23.0;
    c[GC_gKr] = 0.035 * sqrt(5.4 / 5.4);
    c[GC_pCa] = 0.00027;
    c[GC_gKs] = 0.0035;
    c[GC_gK1] = sqrt(5.4 / 5.4) * 0.0525;
    c[GC_gto] = 0.165;
    c[GC_gKur] = 0.045;
    c[GC_gCaB] = 6.06430000000000033e-04;
    c[GC_gNaB] = 0.000597;
    c[GC_gClB] = 0.009;
    c[GC_GClCa] = 0.0548;
    c[GC_IbarNaK] = 1.26;
    c[GC_IbarNCX] = 3.15;
    c[GC_IbarSLCaP] = 0.0471;
    c[GC_VmaxSRCaP] = 0.0053114;
    c[GC_ks] = 25.0;
    c[GC_pK] = 1.35e-07;
    c[GC_pNa] = 7.5e-09;
    c[GC_gKp] = 0.002;
    c[GC_gNaL] = 0.0;                       /* gNaL_max * AF, AF = 0 */
    static const double y0[GR_NS] = {
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
    for (int i = 0; i < GR_NS; ++i) y[i] = y0[i];
}

static void grd_eval(double t, const double* c, const double* y, int solve_vm,
                     const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                     double* iion, double* istim)
{
    const double V = y[GR_V], FRT = c[GC_FRT], F = 96485.0, Cm = 1.1e-10;
    const double Fjn = 0.11, Fsl = 1.0 - 0.11, FjnL = 0.9, FslL = 1.0 - 0.9;
    const double Ko = 5.4, Nao = 140.0, Cao = 1.8, Mgi = 1.0;
    const double Vjn = c[GC_Vjn], Vsl = c[GC_Vsl], Vmyo = c[GC_Vmyo], Vsr = c[GC_Vsr];
    for (int i = 0; i < GR_NS; ++i) g[i] = 0;
#define GR_GATE(idx, inf, tau) do { const double i_ = (inf), t_ = (tau); \
        r[idx] = (i_ - y[idx]) / t_; g[idx] = 1; xi[idx] = i_; ta[idx] = t_; } while (0)
    /* linear in its own state: dx/dt = a - b x; the rate itself as the reference writes it */
#define GR_LIN(idx, a, b, rate) do { const double a_ = (a), b_ = (b); \
        r[idx] = (rate); g[idx] = 1; xi[idx] = a_ / b_; ta[idx] = 1.0 / b_; } while (0)

    /* voltage-gated kinetics */
    GR_GATE(GR_xr, 1.0 / (1.0 + exp(-(V + 10.0) / 5.0)),
            550.0 / (1.0 + exp((-22.0 - V) / 9.0)) * 6.0 / (1.0 + exp((V + 11.0) / 9.0)) +
                230.0 / (1.0 + exp((V + 40.0) / 20.0)));
    GR_GATE(GR_kur_r, 1.0 / (1.0 + exp((V + 6.0) / -8.6)),
            9.0 / (1.0 + exp((V + 5.0) / 12.0)) + 0.5);
    GR_GATE(GR_kur_s, 1.0 / (1.0 + exp((V + 7.5) / 10.0)),
            590.0 / (1.0 + exp((V + 60.0) / 10.0)) + 3050.0);
    const double hjinf = 1.0 / pow(1.0 + exp((V + 71.55) / 7.43), 2.0);
    const double ah = V >= -40.0 ? 0.0 : 0.057 * exp(-(V + 80.0) / 6.8);
    const double bh = V >= -40.0 ? 0.77 / (0.13 * (1.0 + exp(-(V + 10.66) / 11.1)))
                                 : 2.7 * exp(0.079 * V) + 3.1 * pow(10.0, 5.0) * exp(0.3485 * V);
    GR_GATE(GR_h, hjinf, 1.0 / (ah + bh));
    const double aj = V >= -40.0 ? 0.0
        : (-2.5428 * pow(10.0, 4.0) * exp(0.2444 * V) - 6.948 * pow(10.0, -6.0) * exp(-0.04391 * V)) *
              (V + 37.78) / (1.0 + exp(0.311 * (V + 79.23)));
    const double bj = V >= -40.0 ? 0.6 * exp(0.057 * V) / (1.0 + exp(-0.1 * (V + 32.0)))
                                 : 0.02424 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)));
    GR_GATE(GR_j, hjinf, 1.0 / (aj + bj));
    GR_GATE(GR_m, 1.0 / pow(1.0 + exp(-(56.86 + V) / 9.03), 2.0),
            0.1292 * exp(-pow((V + 45.79) / 15.54, 2.0)) + 0.06487 * exp(-pow((V - 4.823) / 51.12, 2.0)));
    GR_GATE(GR_hl, 1.0 / (1.0 + exp((V + 91.0) / 6.1)), 600.0);
    const double aml = 0.32 * (V + 47.13) / (1.0 - exp(-0.1 * (V + 47.13)));
    const double bml = 0.08 * exp(-V / 11.0);
    GR_LIN(GR_ml, aml, aml + bml, aml * (1.0 - y[GR_ml]) - bml * y[GR_ml]);
    GR_GATE(GR_x, 1.0 / (1.0 + exp(-(V + 1.0) / 11.0)), 3.5 * exp(-pow(V / 30.0, 2.0)) + 1.5);
    GR_GATE(GR_y, 1.0 / (1.0 + exp((V + 40.5) / 11.5)),
            25.635 * exp(-pow((V + 52.45) / 15.8827, 2.0)) + 24.14);
    const double dinf = 1.0 / (1.0 + exp(-(V + 9.0) / 6.0));
    GR_GATE(GR_d, dinf, dinf * (1.0 - exp(-(V + 9.0) / 6.0)) / (0.035 * (V + 9.0)));
    GR_GATE(GR_f, 1.0 / (1.0 + exp((V + 30.0) / 7.0)) + 0.2 / (1.0 + exp((50.0 - V) / 20.0)),
            1.0 / (0.0197 * exp(-pow(0.0337 * (V + 25.0), 2.0)) + 0.02));
    GR_GATE(GR_xs, 1.0 / (1.0 + exp(-(V + 3.8) / 14.25)),
            990.1 / (1.0 + exp(-(V + 2.436) / 14.12)));
    GR_LIN(GR_fCaBjn, 1.7 * y[GR_Cajn], 1.7 * y[GR_Cajn] + 0.0119,
           (1.7 * y[GR_Cajn] * (1.0 - y[GR_fCaBjn]) - 0.0119 * y[GR_fCaBjn]) * 1.0);
    GR_LIN(GR_fCaBsl, 1.7 * y[GR_Casl], 1.7 * y[GR_Casl] + 0.0119,
           (1.7 * y[GR_Casl] * (1.0 - y[GR_fCaBsl]) - 0.0119 * y[GR_fCaBsl]) * 1.0);

    /* buffers: kon C (Bmax - B) - koff B, linear in B (a second bound form frozen) */
    const double dNaBjn = 0.0001 * y[GR_Najn] * (7.561 - y[GR_NaBjn]) - 0.001 * y[GR_NaBjn];
    const double dNaBsl = 0.0001 * y[GR_Nasl] * (1.65 - y[GR_NaBsl]) - 0.001 * y[GR_NaBsl];
    GR_LIN(GR_NaBjn, 0.0001 * y[GR_Najn] * 7.561, 0.0001 * y[GR_Najn] + 0.001, dNaBjn);
    GR_LIN(GR_NaBsl, 0.0001 * y[GR_Nasl] * 1.65, 0.0001 * y[GR_Nasl] + 0.001, dNaBsl);
    const double koff_tncl = 0.0196;
    const double dCaM = 34.0 * y[GR_Cai] * (0.024 - y[GR_CaM]) - 0.238 * y[GR_CaM];
    const double dMyoc = 13.8 * y[GR_Cai] * (0.14 - y[GR_Myoc] - y[GR_Myom]) - 0.00046 * y[GR_Myoc];
    const double dMyom = 0.0157 * Mgi * (0.14 - y[GR_Myoc] - y[GR_Myom]) - 5.7e-05 * y[GR_Myom];
    const double dSLHjn = 100.0 * y[GR_Cajn] * (c[GC_BmaxSLHj] - y[GR_SLHjn]) - 0.03 * y[GR_SLHjn];
    const double dSLHsl = 100.0 * y[GR_Casl] * (c[GC_BmaxSLHsl] - y[GR_SLHsl]) - 0.03 * y[GR_SLHsl];
    const double dSLLjn = 100.0 * y[GR_Cajn] * (c[GC_BmaxSLLj] - y[GR_SLLjn]) - 1.3 * y[GR_SLLjn];
    const double dSLLsl = 100.0 * y[GR_Casl] * (c[GC_BmaxSLLsl] - y[GR_SLLsl]) - 1.3 * y[GR_SLLsl];
    const double dSRB = 100.0 * y[GR_Cai] * (0.0171 - y[GR_SRB]) - 0.06 * y[GR_SRB];
    const double dTnCHc = 2.37 * y[GR_Cai] * (0.14 - y[GR_TnCHc] - y[GR_TnCHm]) - 3.2e-05 * y[GR_TnCHc];
    const double dTnCHm = 0.003 * Mgi * (0.14 - y[GR_TnCHc] - y[GR_TnCHm]) - 0.00333 * y[GR_TnCHm];
    const double dTnCL = 32.7 * y[GR_Cai] * (0.07 - y[GR_TnCL]) - koff_tncl * y[GR_TnCL];
    GR_LIN(GR_CaM, 34.0 * y[GR_Cai] * 0.024, 34.0 * y[GR_Cai] + 0.238, dCaM);
    GR_LIN(GR_Myoc, 13.8 * y[GR_Cai] * (0.14 - y[GR_Myom]), 13.8 * y[GR_Cai] + 0.00046, dMyoc);
    GR_LIN(GR_Myom, 0.0157 * Mgi * (0.14 - y[GR_Myoc]), 0.0157 * Mgi + 5.7e-05, dMyom);
    GR_LIN(GR_SLHjn, 100.0 * y[GR_Cajn] * c[GC_BmaxSLHj], 100.0 * y[GR_Cajn] + 0.03, dSLHjn);
    GR_LIN(GR_SLHsl, 100.0 * y[GR_Casl] * c[GC_BmaxSLHsl], 100.0 * y[GR_Casl] + 0.03, dSLHsl);
    GR_LIN(GR_SLLjn, 100.0 * y[GR_Cajn] * c[GC_BmaxSLLj], 100.0 * y[GR_Cajn] + 1.3, dSLLjn);
    GR_LIN(GR_SLLsl, 100.0 * y[GR_Casl] * c[GC_BmaxSLLsl], 100.0 * y[GR_Casl] + 1.3, dSLLsl);
    GR_LIN(GR_SRB, 100.0 * y[GR_Cai] * 0.0171, 100.0 * y[GR_Cai] + 0.06, dSRB);
    GR_LIN(GR_TnCHc, 2.37 * y[GR_Cai] * (0.14 - y[GR_TnCHm]), 2.37 * y[GR_Cai] + 3.2e-05, dTnCHc);
    GR_LIN(GR_TnCHm, 0.003 * Mgi * (0.14 - y[GR_TnCHc]), 0.003 * Mgi + 0.00333, dTnCHm);
    GR_LIN(GR_TnCL, 32.7 * y[GR_Cai] * 0.07, 32.7 * y[GR_Cai] + koff_tncl, dTnCL);
    const double JCaBcyt = dTnCL + dTnCHc + dTnCHm + dCaM + dMyoc + dMyom + dSRB;
    const double JCaBjn = dSLLjn + dSLHjn;
    const double JCaBsl = dSLLsl + dSLHsl;

    /* reversal potentials */
    const double ECajn = 1.0 / FRT / 2.0 * log(Cao / y[GR_Cajn]);
    const double ECasl = 1.0 / FRT / 2.0 * log(Cao / y[GR_Casl]);
    const double EK = 1.0 / FRT * log(Ko / y[GR_Ki]);
    const double ENajn = 1.0 / FRT * log(Nao / y[GR_Najn]);
    const double ENasl = 1.0 / FRT * log(Nao / y[GR_Nasl]);
    const double EKs = 1.0 / FRT * log((Ko + 0.01833 * Nao) / (y[GR_Ki] + 0.01833 * y[GR_Nai]));

    /* INaK */
    const double fnak = 1.0 / (1.0 + 0.1245 * exp(-0.1 * V * FRT) + 0.0365 * c[GC_sigma] * exp(-V * FRT));
    const double KmNaip = 11.0;
    const double INaKjn = Fjn * c[GC_IbarNaK] * fnak * Ko / (1.0 + pow(KmNaip / y[GR_Najn], 4.0)) / (Ko + 1.5);
    const double INaKsl = Fsl * c[GC_IbarNaK] * fnak * Ko / (1.0 + pow(KmNaip / y[GR_Nasl], 4.0)) / (Ko + 1.5);
    const double INaK = INaKjn + INaKsl;

    /* RyR and SR fluxes */
    const double RI = 1.0 - y[GR_rr] - y[GR_ro] - y[GR_ri];
    const double JSRCarel = c[GC_ks] * y[GR_ro] * (y[GR_Casr] - y[GR_Cajn]);
    const double JSRleak = (y[GR_Casr] - y[GR_Cajn]) * 5.348e-06;
    const double Kmf = 2.5 * 0.000246, Kmr = 1.7, hill = 1.787;
    const double fi = pow(y[GR_Cai] / Kmf, hill), fr = pow(y[GR_Casr] / Kmr, hill);
    const double Jserca = pow(2.6, 0.0) * c[GC_VmaxSRCaP] * (fi - fr) / (1.0 + fi + fr);
    const double kCaSR = 15.0 - (15.0 - 1.0) / (1.0 + pow(0.45 / y[GR_Casr], 2.5));
    const double kiSRCa = 0.5 * kCaSR, koSRCa = 10.0 / kCaSR;
    const double cj2 = pow(y[GR_Cajn], 2.0);
    r[GR_ri] = kiSRCa * y[GR_Cajn] * y[GR_ro] - 0.005 * y[GR_ri] - (0.06 * y[GR_ri] - koSRCa * cj2 * RI);
    r[GR_ro] = koSRCa * cj2 * y[GR_rr] - 0.06 * y[GR_ro] - (kiSRCa * y[GR_Cajn] * y[GR_ro] - 0.005 * y[GR_ri]);
    r[GR_rr] = 0.005 * RI - kiSRCa * y[GR_Cajn] * y[GR_rr] - (koSRCa * cj2 * y[GR_rr] - 0.06 * y[GR_ro]);

    /* background and L-type currents */
    const double ICaBjn = Fjn * c[GC_gCaB] * (V - ECajn), ICaBsl = Fsl * c[GC_gCaB] * (V - ECasl);
    const double e2 = exp(2.0 * V * FRT), e1 = exp(V * FRT);
    const double ibarcajn = c[GC_pCa] * 4.0 * (V * F * FRT) * (0.341 * y[GR_Cajn] * e2 - 0.341 * Cao) / (e2 - 1.0);
    const double ibarcasl = c[GC_pCa] * 4.0 * (V * F * FRT) * (0.341 * y[GR_Casl] * e2 - 0.341 * Cao) / (e2 - 1.0);
    const double ibark = c[GC_pK] * V * F * FRT * (0.75 * y[GR_Ki] * e1 - 0.75 * Ko) / (e1 - 1.0);
    const double ibarnajn = c[GC_pNa] * V * F * FRT * (0.75 * y[GR_Najn] * e1 - 0.75 * Nao) / (e1 - 1.0);
    const double ibarnasl = c[GC_pNa] * V * F * FRT * (0.75 * y[GR_Nasl] * e1 - 0.75 * Nao) / (e1 - 1.0);
    const double q10 = pow(1.8, 0.0) * 0.45, df = y[GR_d] * y[GR_f];
    const double ICaLCajn = FjnL * ibarcajn * df * (1.0 - y[GR_fCaBjn]) * q10;
    const double ICaLCasl = FslL * ibarcasl * df * (1.0 - y[GR_fCaBsl]) * q10;
    const double ICaLK = ibark * df * (FjnL * (1.0 - y[GR_fCaBjn]) + FslL * (1.0 - y[GR_fCaBsl])) * q10;
    const double ICaLNajn = FjnL * ibarnajn * df * (1.0 - y[GR_fCaBjn]) * q10;
    const double ICaLNasl = FslL * ibarnasl * df * (1.0 - y[GR_fCaBsl]) * q10;

    /* chloride */
    const double IClB = c[GC_gClB] * (V - c[GC_ECl]);
    const double IClCa = Fjn * c[GC_GClCa] / (1.0 + 0.1 / y[GR_Cajn]) * (V - c[GC_ECl]) +
                         Fsl * c[GC_GClCa] / (1.0 + 0.1 / y[GR_Casl]) * (V - c[GC_ECl]);

    /* potassium */
    const double k1a = 1.02 / (1.0 + exp(0.2385 * (V - EK - 59.215)));
    const double k1b = (0.49124 * exp(0.08032 * (V + 5.476 - EK)) + exp(0.06175 * (V - EK - 594.31))) /
                       (1.0 + exp(-0.5143 * (V - EK + 4.753)));
    const double IK1 = c[GC_gK1] * (k1a / (k1a + k1b)) * (V - EK);
    const double kp = 1.0 / (1.0 + exp(7.488 - V / 5.98));
    const double IKp = Fjn * c[GC_gKp] * kp * (V - EK) + Fsl * c[GC_gKp] * kp * (V - EK);
    const double IKr = c[GC_gKr] * y[GR_xr] * (1.0 / (1.0 + exp((V + 74.0) / 24.0))) * (V - EK);
    const double xs2 = pow(y[GR_xs], 2.0);
    const double IKs = Fjn * c[GC_gKs] * xs2 * (V - EKs) + Fsl * c[GC_gKs] * xs2 * (V - EKs);
    const double IKur = c[GC_gKur] * y[GR_kur_r] * y[GR_kur_s] * (V - EK);
    const double Ito = c[GC_gto] * y[GR_x] * y[GR_y] * (V - EK);

    /* sodium */
    const double m3hj = pow(y[GR_m], 3.0) * y[GR_h] * y[GR_j];
    const double INajn = Fjn * c[GC_gNa] * m3hj * (V - ENajn);
    const double INasl = Fsl * c[GC_gNa] * m3hj * (V - ENasl);
    const double INaBjn = Fjn * c[GC_gNaB] * (V - ENajn), INaBsl = Fsl * c[GC_gNaB] * (V - ENasl);
    const double ml3 = pow(y[GR_ml], 3.0);
    const double INaLjn = Fjn * c[GC_gNaL] * ml3 * y[GR_hl] * (V - ENajn);
    const double INaLsl = Fsl * c[GC_gNaL] * ml3 * y[GR_hl] * (V - ENasl);

    /* NCX and sarcolemmal Ca pump */
    const double nu = 0.35, KmCai = 0.00359, KmCao = 1.3, KmNai = 12.29, KmNao = 87.5;
    const double Kajn = 1.0 / (1.0 + pow(0.000384 / y[GR_Cajn], 2.0));
    const double Kasl = 1.0 / (1.0 + pow(0.000384 / y[GR_Casl], 2.0));
    const double en = exp(nu * V * FRT), en1 = exp((nu - 1.0) * V * FRT);
    const double Nao3 = pow(Nao, 3.0), KmNao3 = pow(KmNao, 3.0);
    const double nj3 = pow(y[GR_Najn], 3.0), ns3 = pow(y[GR_Nasl], 3.0);
    const double s1jn = en * nj3 * Cao, s2jn = en1 * Nao3 * y[GR_Cajn];
    const double s3jn = KmCai * Nao3 * (1.0 + pow(y[GR_Najn] / KmNai, 3.0)) +
                        KmNao3 * y[GR_Cajn] * (1.0 + y[GR_Cajn] / KmCai) + KmCao * nj3 +
                        nj3 * Cao + Nao3 * y[GR_Cajn];
    const double s1sl = en * ns3 * Cao, s2sl = en1 * Nao3 * y[GR_Casl];
    const double s3sl = KmCai * Nao3 * (1.0 + pow(y[GR_Nasl] / KmNai, 3.0)) +
                        KmNao3 * y[GR_Casl] * (1.0 + y[GR_Casl] / KmCai) + KmCao * ns3 +
                        ns3 * Cao + Nao3 * y[GR_Casl];
    const double ncx = c[GC_IbarNCX] * pow(1.57, 0.0);
    const double INaCajn = Fjn * ncx * Kajn * (s1jn - s2jn) / s3jn / (1.0 + 0.27 * en1);
    const double INaCasl = Fsl * ncx * Kasl * (s1sl - s2sl) / s3sl / (1.0 + 0.27 * en1);
    const double pb = pow(0.0005 * 1.0, 1.6);
    const double paj = pow(y[GR_Cajn] * 1.0, 1.6), pas = pow(y[GR_Casl] * 1.0, 1.6);
    const double IpCajn = Fjn * pow(2.35, 0.0) * c[GC_IbarSLCaP] * paj / (paj + pb);
    const double IpCasl = Fsl * pow(2.35, 0.0) * c[GC_IbarSLCaP] * pas / (pas + pb);

    /* totals */
    const double ICatjn = ICaLCajn + ICaBjn + IpCajn - 2.0 * INaCajn;
    const double ICatsl = ICaLCasl + ICaBsl + IpCasl - 2.0 * INaCasl;
    const double INatjn = INajn + INaBjn + 3.0 * INaCajn + 3.0 * INaKjn + ICaLNajn + INaLjn;
    const double INatsl = INasl + INaBsl + 3.0 * INaCasl + 3.0 * INaKsl + ICaLNasl + INaLsl;
    const double IKt = Ito + IKr + IKs + IK1 - 2.0 * INaK + ICaLK + IKp + IKur;
    const double Iion = (INatjn + INatsl) + (IClCa + IClB) + (ICatjn + ICatsl) + IKt;

    /* concentrations */
    const double dCsqn = 100.0 * y[GR_Casr] * (c[GC_BmaxCsqn] - y[GR_Csqn]) - 65.0 * y[GR_Csqn];
    GR_LIN(GR_Csqn, 100.0 * y[GR_Casr] * c[GC_BmaxCsqn], 100.0 * y[GR_Casr] + 65.0, dCsqn);
    const double JCaslmyo = 3.72425607984805052e-12, JCajnsl = 8.24130542277896849e-13;
    const double JNaslmyo = 1.63862792221979433e-12, JNajnsl = 1.83127823220607955e-14;
    r[GR_Cai] = -Jserca * Vsr / Vmyo - JCaBcyt + JCaslmyo / Vmyo * (y[GR_Casl] - y[GR_Cai]);
    r[GR_Nai] = JNaslmyo / Vmyo * (y[GR_Nasl] - y[GR_Nai]);
    /* SR, junctional and subsarcolemmal Ca: stiff at dt ~ 10 us (release, the small
     * junction volume and the buffer on-rates give -d(rate)/dx ~ 1e2..1e3 /ms), so the
     * Rush-Larsen step treats each as linear in itself with the diagonal coefficient b of
     * those terms frozen: x' = a - b x, a = rate + b x (the same rate) */
    const double rCasr = Jserca - (JSRleak * Vmyo / Vsr + JSRCarel) - dCsqn;
    const double bCasr = c[GC_ks] * y[GR_ro] + 5.348e-06 * Vmyo / Vsr +
                         100.0 * (c[GC_BmaxCsqn] - y[GR_Csqn]);
    GR_LIN(GR_Casr, rCasr + bCasr * y[GR_Casr], bCasr, rCasr);
    const double rCajn = -ICatjn * Cm / (Vjn * 2.0 * F) + JCajnsl / Vjn * (y[GR_Casl] - y[GR_Cajn]) -
                         JCaBjn + JSRCarel * Vsr / Vjn + JSRleak * Vmyo / Vjn;
    const double bCajn = c[GC_ks] * y[GR_ro] * Vsr / Vjn + 5.348e-06 * Vmyo / Vjn + JCajnsl / Vjn +
                         100.0 * ((c[GC_BmaxSLLj] - y[GR_SLLjn]) + (c[GC_BmaxSLHj] - y[GR_SLHjn]));
    GR_LIN(GR_Cajn, rCajn + bCajn * y[GR_Cajn], bCajn, rCajn);
    const double rCasl = -ICatsl * Cm / (Vsl * 2.0 * F) + JCajnsl / Vsl * (y[GR_Cajn] - y[GR_Casl]) +
                         JCaslmyo / Vsl * (y[GR_Cai] - y[GR_Casl]) - JCaBsl;
    const double bCasl = JCajnsl / Vsl + JCaslmyo / Vsl +
                         100.0 * ((c[GC_BmaxSLLsl] - y[GR_SLLsl]) + (c[GC_BmaxSLHsl] - y[GR_SLHsl]));
    GR_LIN(GR_Casl, rCasl + bCasl * y[GR_Casl], bCasl, rCasl);
    r[GR_Najn] = -INatjn * Cm / (Vjn * F) + JNajnsl / Vjn * (y[GR_Nasl] - y[GR_Najn]) - dNaBjn;
    r[GR_Nasl] = -INatsl * Cm / (Vsl * F) + JNajnsl / Vsl * (y[GR_Najn] - y[GR_Nasl]) +
                 JNaslmyo / Vsl * (y[GR_Nai] - y[GR_Nasl]) - dNaBsl;
    r[GR_Ki] = 0.0;
#undef GR_GATE
#undef GR_LIN
    const double Ist = orc_stimulus(t, st);
    r[GR_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion;
    *istim = Ist;
}

/* ============================================================ ToRORd_dynCl
 * ToR-ORd with dynamic chloride (Tomek et al. 2019, 2023 revision), 45 states:
 * constants and initial states of ToRORd_dynCl_2023.H:571-846 (ToRORd_dynClinitConsts,
 * three tissue columns), right-hand side of :848-1177, state order of its STATES_INDEX
 * enum.  Tissues endocardialCells, mCells, epicardialCells (ToRORd_dynCl.C:93-96).  V in
 * mV, Im = Iion_cm.  The K_i rate omits the stimulus, as the reference does inside
 * derivatives() (ALGEBRAIC[Istim] is read before it is assigned, :1119 vs :1172, on a
 * zeroed scratch).  I_katp is 0 (fkatp = 0 in initConsts).
 * Rush-Larsen classification: the Hodgkin-Huxley gates and the RyR release fluxes as
 * gates; nca_i, nca_ss and the five IKr Markov states as linear in their own state with
 * the others frozen over the step; the rest forward Euler.
 * Parity pins: the reference's own expressions at seeded states
 * (tests/test_oracle_model_vectors.py). */
enum { TR_V, TR_CaMKt, TR_Nai, TR_Nass, TR_Ki, TR_Kss, TR_Cass, TR_Cansr, TR_Cajsr, TR_Cai,
       TR_Cli, TR_Clss, TR_m, TR_h, TR_j, TR_hp, TR_jp, TR_mL, TR_hL, TR_hLp, TR_a, TR_iF,
       TR_iS, TR_ap, TR_iFp, TR_iSp, TR_d, TR_ff, TR_fs, TR_fcaf, TR_fcas, TR_jca, TR_ffp,
       TR_fcafp, TR_nca_ss, TR_nca_i, TR_C1, TR_C2, TR_C3, TR_I, TR_O, TR_xs1, TR_xs2,
       TR_Jrelnp, TR_Jrelp, TR_NS };
/* tissue-dependent and rescalable constants (the tissue flag itself, for delta_epi of Ito
 * and the M-cell release factor, comes with each evaluation) */
enum { RC_GNa, RC_GKr, RC_PCa, RC_GKs, RC_GK1, RC_GKb, RC_GNaL, RC_Gto, RC_Gncx, RC_Pnak,
       RC_cmdnmax, RC_upScale, RC_GpCa, RC_GClCa, RC_GClb, RC_PCab, RC_PNab, RC_Jrel_b,
       RC_Jup_b, RC_bt, RC_gkatp, RC_fkatp, RC_NC };

static void tor_init(int tf, double* c, double* y)
{
    /* initConsts' tissue selections (:618-769): flag 1 epi, 2 M, 3 endo */
    const int epi = tf == 1, mid = tf == 2;
    const double PCa_b = 8.37570000000000046e-05;
    c[RC_GNa] = 11.7802;
    c[RC_GKr] = epi ? 0.0321 * 1.3 : mid ? 0.0321 * 0.8 : 0.0321;
    c[RC_PCa] = epi ? PCa_b * 1.2 : mid ? PCa_b * 2.0 : PCa_b;
    c[RC_GKs] = epi ? 0.0011 * 1.4 : 0.0011;
    c[RC_GK1] = epi ? 0.6992 * 1.2 : mid ? 0.6992 * 1.3 : 0.6992;
    c[RC_GKb] = epi ? 0.0189 * 0.6 : 0.0189;
    c[RC_GNaL] = epi ? 0.0279 * 0.6 : 0.0279;
    c[RC_Gto] = (epi || mid) ? 0.16 * 2.0 : 0.16;
    c[RC_Gncx] = epi ? 0.0034 * 1.1 : mid ? 0.0034 * 1.4 : 0.0034;
    c[RC_Pnak] = epi ? 15.4509 * 0.9 : mid ? 15.4509 * 0.7 : 15.4509;
    c[RC_cmdnmax] = epi ? 0.05 * 1.3 : 0.05;
    c[RC_upScale] = epi ? 1.3 : 1.0;
    c[RC_GpCa] = 0.0005;
    c[RC_GClCa] = 0.2843;
    c[RC_GClb] = 0.00198;
    c[RC_PCab] = 5.91940000000000005e-08;
    c[RC_PNab] = 1.9239e-09;
    c[RC_Jrel_b] = 1.5378;
    c[RC_Jup_b] = 1.0;
    c[RC_bt] = 4.75;
    c[RC_gkatp] = 4.3195;
    c[RC_fkatp] = 0.0;
    /* initial states (:782-846), columns endo / M / epi */
    static const double y0[3][TR_NS] = {
        {-89.74808, 1.10e-02, 12.39736, 12.3977, 147.7115, 147.7114, 6.50e-05, 1.528001,
         1.525693, 7.45e-05, 29.20698, 29.20696, 6.52e-04, 0.8473267, 0.8471657, 0.7018454,
         0.8469014, 1.35e-04, 0.5566017, 0.3115491, 8.90e-04, 0.9996716, 0.5988908,
         4.53e-04, 0.9996716, 0.6620692, 1.59e-31, 1.0, 0.9401791, 1.0, 0.9999014,
         0.9999846, 1.0, 1.0, 4.90e-04, 8.33e-04, 0.9982511, 7.94e-04, 6.53e-04, 9.80e-06,
         2.92e-04, 0.243959, 1.59e-04, 1.81e-22, 4.36e-21},
        {-91.33918, 2.01882e-02, 15.94867, 15.94922, 156.7131, 156.7130, 6.64e-05, 2.012225,
         2.016415, 8.30e-05, 48.91277, 48.91274, 4.62e-04, 0.8739077, 0.8737841, 0.7478972,
         0.8735375, 9.99e-05, 0.5986118, 0.3339899, 7.99e-04, 0.9997514, 0.5702538,
         4.07e-04, 0.9997514, 0.6351927, -8.33e-30, 1.0, 0.9183587, 1.0, 0.9997540,
         0.9999743, 1.0, 1.0, 5.34e-04, 1.26e-03, 0.9983451, 7.09e-04, 5.91e-04, 1.06e-05,
         3.45e-04, 0.2642293, 1.33e-04, -1.30e-21, -7.61e-20},
        {-90.74563, 1.27e-02, 13.40062, 13.40094, 152.3639, 152.3638, 5.75e-05, 1.806794,
         1.805047, 6.62e-05, 34.31721, 34.31719, 5.25e-04, 0.8645148, 0.8644571, 0.7313656,
         0.8643527, 1.12e-04, 0.5916536, 0.3476812, 8.32e-04, 0.9997242, 0.9997235,
         4.24e-04, 0.9997242, 0.9997241, -2.49e-36, 1.0, 0.9510602, 1.0, 0.9999377,
         0.9999886, 1.0, 1.0, 3.05e-04, 5.27e-04, 0.9984733, 7.39e-04, 6.03e-04, 5.68e-06,
         1.79e-04, 0.2233584, 1.42e-04, 6.78e-25, -1.58e-23}};
    memcpy(y, y0[tf == 3 ? 0 : mid ? 1 : 2], sizeof y0[0]);
}

static void tor_eval(int tf, double t, const double* c, const double* y, int solve_vm,
                     const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                     double* iion, double* istim)
{
    /* fixed constants of initConsts (:573-770) */
    const double L = 0.01, rad = 0.0011;
    const double Ageo = 2.0 * 3.14 * rad * rad + 2.0 * 3.14 * rad * L;
    const double vcell = 1000.0 * 3.14 * rad * rad * L;
    const double Acap = 2.0 * Ageo;
    const double vjsr = 0.0048 * vcell, vmyo = 0.68 * vcell, vnsr = 0.0552 * vcell,
                 vss = 0.02 * vcell;
    const double cao = 1.8, clo = 150.0, ko = 5.0, nao = 140.0;
    const double F = 96485.0, R = 8314.0, T = 310.0;
    const double KmCaMK = 0.15, aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015;

    const double V = y[TR_V], Cai = y[TR_Cai], Cass = y[TR_Cass], Nai = y[TR_Nai],
                 Nass = y[TR_Nass], Ki = y[TR_Ki], Kss = y[TR_Kss], Cli = y[TR_Cli],
                 Clss = y[TR_Clss];
    for (int i = 0; i < TR_NS; ++i) g[i] = 0;
#define TR_GATE(idx, inf, tau) do { const double i_ = (inf), t_ = (tau); \
        r[idx] = (i_ - y[idx]) / t_; g[idx] = 1; xi[idx] = i_; ta[idx] = t_; } while (0)
    /* dx/dt = a - b x written as the reference writes it (rate) */
#define TR_LIN(idx, a, b, rate) do { const double a_ = (a), b_ = (b); \
        r[idx] = (rate); g[idx] = 1; xi[idx] = a_ / b_; ta[idx] = 1.0 / b_; } while (0)

    /* CaMK, IpCa, diffusion, transfer, SERCA (:850-877) */
    const double CaMKb = CaMKo * (1.0 - y[TR_CaMKt]) / (1.0 + KmCaM / Cass);
    const double CaMKa = CaMKb + y[TR_CaMKt];
    r[TR_CaMKt] = aCaMK * CaMKb * (CaMKb + y[TR_CaMKt]) - bCaMK * y[TR_CaMKt];
    const double fcamk = 1.0 / (1.0 + KmCaMK / CaMKa);
    const double IpCa = c[RC_GpCa] * Cai / (0.0005 + Cai);
    const double Jdiff = (Cass - Cai) / 0.2, JdiffCl = (Clss - Cli) / 2.0,
                 JdiffK = (Kss - Ki) / 2.0, JdiffNa = (Nass - Nai) / 2.0;
    const double Jtr = (y[TR_Cansr] - y[TR_Cajsr]) / 60.0;
    const double Jleak = 0.0048825 * y[TR_Cansr] / 15.0;
    const double Jupnp = c[RC_upScale] * 0.005425 * Cai / (Cai + 0.00092);
    const double Jupp = c[RC_upScale] * 2.75 * 0.005425 * Cai / (Cai + 0.00092 - 0.00017);
    const double Jup = c[RC_Jup_b] * ((1.0 - fcamk) * Jupnp + fcamk * Jupp - Jleak);

    /* reversal potentials (:879-884); z_Cl = -1 */
    const double RTF = R * T / F;
    const double ECl = R * T / (-1.0 * F) * log(clo / Cli);
    const double EClss = R * T / (-1.0 * F) * log(clo / Clss);
    const double EK = RTF * log(ko / Ki);
    const double ENa = RTF * log(nao / Nai);
    const double EKs = RTF * log((ko + 0.01833 * nao) / (Ki + 0.01833 * Nai));

    /* chloride (:886-891); Fjunc = 1 leaves the subsarcolemmal Ca-activated part at 0 */
    const double IClCa_junc = 1.0 * c[RC_GClCa] / (1.0 + 0.1 / Cass) * (V - EClss);
    const double IClCa_sl = (1.0 - 1.0) * c[RC_GClCa] / (1.0 + 0.1 / Cai) * (V - ECl);
    const double IClb = c[RC_GClb] * (V - ECl);
    const double IClCa = IClCa_junc + IClCa_sl;

    /* IK1, IKb, IKs (:893-906) */
    const double dvk = V - EK;
    const double aK1 = 4.094 / (1.0 + exp(0.1217 * (dvk - 49.934)));
    const double bK1 = (15.72 * exp(0.0674 * (dvk - 3.257)) + exp(0.0618 * (dvk - 594.31))) /
                       (1.0 + exp(-0.1629 * (dvk + 14.207)));
    const double IK1 = c[RC_GK1] * sqrt(ko / 5.0) * (aK1 / (aK1 + bK1)) * dvk;
    const double IKb = c[RC_GKb] * (1.0 / (1.0 + exp(-(V - 10.8968) / 23.9871))) * dvk;
    const double KsCa = 1.0 + 0.6 / (1.0 + pow(3.8e-05 / Cai, 1.4));
    const double xs1ss = 1.0 / (1.0 + exp(-(V + 11.6) / 8.932));
    TR_GATE(TR_xs1, xs1ss, 817.3 + 1.0 / (0.0002326 * exp((V + 48.28) / 17.8) +
                                          0.001292 * exp(-(V + 210.0) / 230.0)));
    TR_GATE(TR_xs2, xs1ss, 1.0 / (0.01 * exp((V - 50.0) / 20.0) + 0.0193 * exp(-(V + 66.54) / 31.0)));
    const double IKs = c[RC_GKs] * KsCa * y[TR_xs1] * y[TR_xs2] * (V - EKs);

    /* INa (:908-930) */
    const double ah = V >= -40.0 ? 0.0 : 0.057 * exp(-(V + 80.0) / 6.8);
    const double aj = V >= -40.0 ? 0.0
        : (-25428.0 * exp(0.2444 * V) - 6.948e-06 * exp(-0.04391 * V)) * (V + 37.78) /
              (1.0 + exp(0.311 * (V + 79.23)));
    const double bh = V >= -40.0 ? 0.77 / (0.13 * (1.0 + exp(-(V + 10.66) / 11.1)))
                                 : 2.7 * exp(0.079 * V) + 310000.0 * exp(0.3485 * V);
    const double bj = V >= -40.0 ? 0.6 * exp(0.057 * V) / (1.0 + exp(-0.1 * (V + 32.0)))
                                 : 0.02424 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)));
    const double hss = 1.0 / pow(1.0 + exp((V + 71.55) / 7.43), 2.0);
    const double hssp = 1.0 / pow(1.0 + exp((V + 77.55) / 7.43), 2.0);
    const double mss = 1.0 / pow(1.0 + exp(-(V + 56.86) / 9.03), 2.0);
    const double tm = 0.1292 * exp(-pow((V + 45.79) / 15.54, 2.0)) +
                      0.06487 * exp(-pow((V - 4.823) / 51.12, 2.0));
    const double INa = c[RC_GNa] * (V - ENa) * pow(y[TR_m], 3.0) *
                       ((1.0 - fcamk) * y[TR_h] * y[TR_j] + fcamk * y[TR_hp] * y[TR_jp]);
    const double th = 1.0 / (ah + bh), tj = 1.0 / (aj + bj);
    TR_GATE(TR_m, mss, tm);
    TR_GATE(TR_h, hss, th);
    TR_GATE(TR_hp, hssp, th);
    TR_GATE(TR_j, hss, tj);
    TR_GATE(TR_jp, hss, 1.46 * tj);

    /* INaL (:932-941) */
    const double INaL = c[RC_GNaL] * (V - ENa) * y[TR_mL] *
                        ((1.0 - fcamk) * y[TR_hL] + fcamk * y[TR_hLp]);
    TR_GATE(TR_hL, 1.0 / (1.0 + exp((V + 87.61) / 7.488)), 200.0);
    TR_GATE(TR_mL, 1.0 / (1.0 + exp(-(V + 42.85) / 5.264)), tm);
    TR_GATE(TR_hLp, 1.0 / (1.0 + exp((V + 93.81) / 7.488)), 3.0 * 200.0);

    /* I_katp (:943-944): akik = (ko/5)^0.24, bkik = 1/(1 + (2/0.25)^2) */
    const double Ikatp = c[RC_fkatp] * c[RC_gkatp] * pow(ko / 5.0, 0.24) *
                         (1.0 / (1.0 + pow(2.0 / 0.25, 2.0))) * dvk;

    /* Ito (:946-973); EKshift = 0 */
    const double AiF = 1.0 / (1.0 + exp((V - 213.6) / 151.2));
    const double delta_epi = tf == 1 ? 1.0 - 0.95 / (1.0 + exp((V + 70.0) / 5.0)) : 1.0;
    const double dti = (1.354 + 0.0001 / (exp((V - 167.4) / 15.89) + exp(-(V - 12.23) / 0.2154))) *
                       (1.0 - 0.5 / (1.0 + exp((V + 70.0) / 20.0)));
    const double iss = 1.0 / (1.0 + exp((V + 43.94) / 5.711));
    const double tau_a = 1.0515 / (1.0 / (1.2089 * (1.0 + exp(-(V - 18.4099) / 29.3814))) +
                                   3.5 / (1.0 + exp((V + 100.0) / 29.3814)));
    const double tiF = (4.562 + 1.0 / (0.3933 * exp(-(V + 100.0) / 100.0) +
                                       0.08004 * exp((V + 50.0) / 16.59))) * delta_epi;
    const double tiS = (23.62 + 1.0 / (0.001416 * exp(-(V + 96.52) / 59.05) +
                                       1.78e-08 * exp((V + 114.1) / 8.079))) * delta_epi;
    const double AiS = 1.0 - AiF;
    const double ii = AiF * y[TR_iF] + AiS * y[TR_iS];
    const double ip = AiF * y[TR_iFp] + AiS * y[TR_iSp];
    const double Ito = c[RC_Gto] * dvk * ((1.0 - fcamk) * y[TR_a] * ii + fcamk * y[TR_ap] * ip);
    TR_GATE(TR_a, 1.0 / (1.0 + exp(-(V - 14.34) / 14.82)), tau_a);
    TR_GATE(TR_ap, 1.0 / (1.0 + exp(-(V - 24.34) / 14.82)), tau_a);
    TR_GATE(TR_iF, iss, tiF);
    TR_GATE(TR_iS, iss, tiS);
    TR_GATE(TR_iFp, iss, dti * tiF);
    TR_GATE(TR_iSp, iss, dti * tiS);

    /* ICaL gates and activity coefficients (:975-1008) */
    const double Afcaf = 0.3 + 0.6 / (1.0 + exp((V - 10.0) / 10.0)), Afcas = 1.0 - Afcaf;
    const double Ii = 0.5 * (Nai + Ki + Cli + 4.0 * Cai) / 1000.0;
    const double Iss = 0.5 * (Nass + Kss + Clss + 4.0 * Cass) / 1000.0;
    const double dss = V >= 31.4978 ? 1.0 : 1.0763 * exp(-1.007 * exp(-0.0829 * V));
    const double fss = 1.0 / (1.0 + exp((V + 19.58) / 3.696));
    const double km2n = y[TR_jca] * 1.0;
    const double tfcaf = 7.0 + 1.0 / (0.04 * exp(-(V - 4.0) / 7.0) + 0.04 * exp((V - 4.0) / 7.0));
    const double tff = 7.0 + 1.0 / (0.0045 * exp(-(V + 20.0) / 10.0) + 0.0045 * exp((V + 20.0) / 10.0));
    const double k2n = 500.0, Kmn = 0.002, Aff = 0.6, Afs = 1.0 - Aff;
    const double anca_i = 1.0 / (k2n / km2n + pow(1.0 + Kmn / Cai, 4.0));
    const double anca_ss = 1.0 / (k2n / km2n + pow(1.0 + Kmn / Cass, 4.0));
    const double f = Aff * y[TR_ff] + Afs * y[TR_fs];
    const double fca = Afcaf * y[TR_fcaf] + Afcas * y[TR_fcas];
    const double fcap = Afcaf * y[TR_fcafp] + Afcas * y[TR_fcas];
    const double fp = Aff * y[TR_ffp] + Afs * y[TR_fs];
    const double constA = 1820000.0 * pow(74.0 * T, -1.5);
    const double Io = 0.5 * (nao + ko + clo + 4.0 * cao) / 1000.0;
    const double sIi = sqrt(Ii), sIss = sqrt(Iss), sIo = sqrt(Io);
    const double lgi = sIi / (1.0 + sIi) - 0.3 * Ii, lgss = sIss / (1.0 + sIss) - 0.3 * Iss,
                 lgo = sIo / (1.0 + sIo) - 0.3 * Io;
    const double gamma_cai = exp(-constA * 4.0 * lgi), gamma_cass = exp(-constA * 4.0 * lgss);
    const double gamma_ki = exp(-constA * 1.0 * lgi), gamma_kss = exp(-constA * 1.0 * lgss);
    const double gamma_cao = exp(-constA * 4.0 * lgo), gamma_o1 = exp(-constA * 1.0 * lgo);
    TR_GATE(TR_d, dss, 0.6 + 1.0 / (exp(-0.05 * (V + 6.0)) + exp(0.09 * (V + 14.0))));
    TR_GATE(TR_ff, fss, tff);
    TR_GATE(TR_fs, fss, 1000.0 + 1.0 / (3.5e-05 * exp(-(V + 5.0) / 4.0) + 3.5e-05 * exp((V + 5.0) / 6.0)));
    TR_GATE(TR_jca, 1.0 / (1.0 + exp((V + 18.08) / 2.7916)), 72.5);
    TR_GATE(TR_fcaf, fss, tfcaf);
    TR_GATE(TR_fcafp, fss, 2.5 * tfcaf);
    TR_GATE(TR_fcas, fss, 100.0 + 1.0 / (0.00012 * exp(-V / 3.0) + 0.00012 * exp(V / 7.0)));
    TR_GATE(TR_ffp, fss, 2.5 * tff);
    TR_LIN(TR_nca_i, anca_i * k2n, km2n, anca_i * k2n - y[TR_nca_i] * km2n);
    TR_LIN(TR_nca_ss, anca_ss * k2n, km2n, anca_ss * k2n - y[TR_nca_ss] * km2n);

    /* IKr Markov chain (:1011, :1063-1109) */
    const double vfrt = V * F / (R * T), vffrt = V * F * F / (R * T);
    const double IKr = c[RC_GKr] * sqrt(ko / 5.0) * y[TR_O] * dvk;
    const double al = 0.1161 * exp(0.299 * vfrt), be = 0.2442 * exp(-1.604 * vfrt);
    const double al2 = 0.0578 * exp(0.971 * vfrt), be2 = 0.000349 * exp(-1.062 * vfrt);
    const double ali = 0.2533 * exp(0.5953 * vfrt), bei = 0.06525 * exp(-0.8209 * vfrt);
    const double alCI = 5.2e-05 * exp(1.525 * vfrt);
    const double beIC = be2 * bei * alCI / (al2 * ali);
    const double al1 = 0.154375, be1 = 0.1911;
    const double C1 = y[TR_C1], C2 = y[TR_C2], C3 = y[TR_C3], Ik = y[TR_I], O = y[TR_O];
    TR_LIN(TR_C3, be * C2, al, be * C2 - al * C3);
    TR_LIN(TR_C2, al * C3 + be1 * C1, be + al1, al * C3 + be1 * C1 - (be + al1) * C2);
    TR_LIN(TR_C1, al1 * C2 + be2 * O + beIC * Ik, be1 + al2 + alCI,
           al1 * C2 + be2 * O + beIC * Ik - (be1 + al2 + alCI) * C1);
    TR_LIN(TR_O, al2 * C1 + bei * Ik, be2 + ali, al2 * C1 + bei * Ik - (be2 + ali) * O);
    TR_LIN(TR_I, alCI * C1 + ali * O, beIC + bei, alCI * C1 + ali * O - (beIC + bei) * Ik);

    /* INaCa, i and ss compartments (:1013-1023, :1065-1166) */
    const double kna1 = 15.0, kna2 = 5.0, kna3 = 88.12, kcaon = 1500000.0, kcaoff = 5000.0,
                 wca = 60000.0, wna = 60000.0, wnaca = 5000.0;
    const double h10 = 12.5 + 1.0 + nao / kna1 * (1.0 + nao / kna2);
    const double h11 = nao * nao / (h10 * kna1 * kna2), h12 = 1.0 / h10;
    const double k1 = h12 * cao * kcaon, k2 = kcaoff, k5 = kcaoff;
    const double hca = exp(0.167 * vfrt), hna = exp(0.5224 * vfrt);
    double INaCa[2];
    for (int comp = 0; comp < 2; ++comp) {
        const double na = comp == 0 ? Nai : Nass, ca = comp == 0 ? Cai : Cass;
        const double allo = 1.0 / (1.0 + pow(0.00015 / ca, 2.0));
        const double h4 = 1.0 + na / kna1 * (1.0 + na / kna2);
        const double h5 = na * na / (h4 * kna1 * kna2), h6 = 1.0 / h4;
        const double h1 = 1.0 + na / kna3 * (1.0 + hna);
        const double h7 = 1.0 + nao / kna3 * (1.0 + 1.0 / hna);
        const double h2 = na * hna / (kna3 * h1), h3 = 1.0 / h1;
        const double h8 = nao / (kna3 * hna * h7), h9 = 1.0 / h7;
        const double k6 = h6 * ca * kcaon;
        const double k3pp = h8 * wnaca, k4pp = h2 * wnaca;
        const double k3 = h9 * wca + k3pp, k4 = h3 * wca / hca + k4pp;
        const double k7 = h5 * h2 * wna, k8 = h8 * h11 * wna;
        const double x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
        const double x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
        const double x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
        const double x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
        const double sx = x1 + x2 + x3 + x4;
        const double E1 = x1 / sx, E2 = x2 / sx, E3 = x3 / sx, E4 = x4 / sx;
        const double JCa = E2 * k2 - E1 * k1;
        const double JNa = 3.0 * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
        INaCa[comp] = (comp == 0 ? 1.0 - 0.35 : 0.35) * c[RC_Gncx] * allo * (1.0 * JNa + 2.0 * JCa);
    }

    /* INaK (:1025-1027, :1081-1139) */
    const double Knai = 9.073 * exp(-0.155 * vfrt / 3.0);
    const double Knao = 27.78 * exp((1.0 - -0.155) * vfrt / 3.0);
    const double Pp = 4.2 / (1.0 + 1e-07 / 1.698e-07 + Nai / 224.0 + Ki / 292.0);
    const double dnk = pow(1.0 + Nai / Knai, 3.0) + pow(1.0 + Ki / 0.5, 2.0) - 1.0;
    const double dno = pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / 0.3582, 2.0) - 1.0;
    const double a1 = 949.5 * pow(Nai / Knai, 3.0) / dnk;
    const double b1 = 182.4 * 0.05, a2 = 687.2;
    const double b2 = 39.4 * pow(nao / Knao, 3.0) / dno;
    const double a3 = 1899.0 * pow(ko / 0.3582, 2.0) / dno;
    const double b3 = 79300.0 * Pp * 1e-07 / (1.0 + 9.8 / 1.698e-07);
    const double a4 = 639.0 * 9.8 / 1.698e-07 / (1.0 + 9.8 / 1.698e-07);
    const double b4 = 40.0 * pow(Ki / 0.5, 2.0) / dnk;
    const double x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
    const double x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
    const double x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
    const double x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
    const double sx = x1 + x2 + x3 + x4;
    const double JnakK = 2.0 * (x4 / sx * b1 - x3 / sx * a1);
    const double JnakNa = 3.0 * (x1 / sx * a3 - x2 / sx * b3);
    const double INaK = c[RC_Pnak] * (1.0 * JnakNa + 1.0 * JnakK);

    /* GHK fluxes and L-type currents (:1047-1062) */
    const double e1 = exp(1.0 * vfrt), e2 = exp(2.0 * vfrt);
    const double PhiCaK_i = 1.0 * vffrt * (gamma_ki * Ki * e1 - gamma_o1 * ko) / (e1 - 1.0);
    const double PhiCaK_ss = 1.0 * vffrt * (gamma_kss * Kss * e1 - gamma_o1 * ko) / (e1 - 1.0);
    const double PhiCaL_i = 4.0 * vffrt * (gamma_cai * Cai * e2 - gamma_cao * cao) / (e2 - 1.0);
    const double PhiCaL_ss = 4.0 * vffrt * (gamma_cass * Cass * e2 - gamma_cao * cao) / (e2 - 1.0);
    const double PhiCaNa_i = 1.0 * vffrt * (gamma_ki * Nai * e1 - gamma_o1 * nao) / (e1 - 1.0);
    const double PhiCaNa_ss = 1.0 * vffrt * (gamma_kss * Nass * e1 - gamma_o1 * nao) / (e1 - 1.0);
    const double ICab = c[RC_PCab] * 4.0 * vffrt * (gamma_cai * Cai * e2 - gamma_cao * cao) / (e2 - 1.0);
    const double INab = c[RC_PNab] * vffrt * (Nai * exp(vfrt) - nao) / (exp(vfrt) - 1.0);
    const double PCa = c[RC_PCa], PCap = 1.1 * PCa;
    const double gi = f * (1.0 - y[TR_nca_i]) + y[TR_jca] * fca * y[TR_nca_i];
    const double gpi = fp * (1.0 - y[TR_nca_i]) + y[TR_jca] * fcap * y[TR_nca_i];
    const double gs = f * (1.0 - y[TR_nca_ss]) + y[TR_jca] * fca * y[TR_nca_ss];
    const double gps = fp * (1.0 - y[TR_nca_ss]) + y[TR_jca] * fcap * y[TR_nca_ss];
    const double dd = y[TR_d];
    /* fractionSS = 0.8 of the L-type flux faces the subspace */
    const double fi = 1.0 - 0.8, fs = 0.8;
    const double ICaL_i = fi * ((1.0 - fcamk) * PCa * PhiCaL_i * dd * gi + fcamk * PCap * PhiCaL_i * dd * gpi);
    const double ICaL_ss = fs * ((1.0 - fcamk) * PCa * PhiCaL_ss * dd * gs + fcamk * PCap * PhiCaL_ss * dd * gps);
    const double ICaK_i = fi * ((1.0 - fcamk) * (0.0003574 * PCa) * PhiCaK_i * dd * gi +
                                fcamk * (0.0003574 * PCap) * PhiCaK_i * dd * gpi);
    const double ICaK_ss = fs * ((1.0 - fcamk) * (0.0003574 * PCa) * PhiCaK_ss * dd * gs +
                                 fcamk * (0.0003574 * PCap) * PhiCaK_ss * dd * gps);
    const double ICaNa_i = fi * ((1.0 - fcamk) * (0.00125 * PCa) * PhiCaNa_i * dd * gi +
                                 fcamk * (0.00125 * PCap) * PhiCaNa_i * dd * gpi);
    const double ICaNa_ss = fs * ((1.0 - fcamk) * (0.00125 * PCa) * PhiCaNa_ss * dd * gs +
                                  fcamk * (0.00125 * PCap) * PhiCaNa_ss * dd * gps);

    /* buffers and RyR (:1032-1045) */
    const double Bcajsr = 1.0 / (1.0 + 10.0 * 0.8 / pow(0.8 + y[TR_Cajsr], 2.0));
    const double Bcass = 1.0 / (1.0 + 0.047 * 0.00087 / pow(0.00087 + Cass, 2.0) +
                                1.124 * 0.0087 / pow(0.0087 + Cass, 2.0));
    const double Bcai = 1.0 / (1.0 + c[RC_cmdnmax] * 0.00238 / pow(0.00238 + Cai, 2.0) +
                               0.07 * 0.0005 / pow(0.0005 + Cai, 2.0));
    const double bt = c[RC_bt], btp = 1.25 * bt;
    const double a_rel = 0.5 * bt / 1.0, a_relp = 0.5 * btp / 1.0;
    const double Jrel = c[RC_Jrel_b] * ((1.0 - fcamk) * y[TR_Jrelnp] + fcamk * y[TR_Jrelp]);
    double trel = bt / (1.0 + 0.0123 / y[TR_Cajsr]);
    if (trel < 0.001) trel = 0.001;
    double trelp = btp / (1.0 + 0.0123 / y[TR_Cajsr]);
    if (trelp < 0.001) trelp = 0.001;
    const double p8 = 1.0 + pow(1.7 / y[TR_Cajsr], 8.0);
    double Jrel_inf = -a_rel * ICaL_ss / 1.0 / p8;
    double Jrel_infp = -a_relp * ICaL_ss / 1.0 / p8;
    if (tf == 2) {
        Jrel_inf *= 1.7;
        Jrel_infp *= 1.7;
    }
    TR_GATE(TR_Jrelnp, Jrel_inf, trel);
    TR_GATE(TR_Jrelp, Jrel_infp, trelp);
#undef TR_GATE
#undef TR_LIN

    /* concentrations (:1029-1031, :1119-1165) */
    r[TR_Cansr] = Jup - Jtr * vjsr / vnsr;
    r[TR_Cli] = (IClb + IClCa_sl) * Acap / (F * vmyo) + JdiffCl * vss / vmyo;
    r[TR_Clss] = IClCa_junc * Acap / (F * vss) - JdiffCl;
    r[TR_Cajsr] = Bcajsr * (Jtr - Jrel);
    r[TR_Ki] = -(Ito + IKr + IKs + IK1 + IKb + Ikatp - 2.0 * INaK + ICaK_i) * Acap / (F * vmyo) +
               JdiffK * vss / vmyo;
    r[TR_Kss] = -ICaK_ss * Acap / (F * vss) - JdiffK;
    r[TR_Cai] = Bcai * (-(ICaL_i + IpCa + ICab - 2.0 * INaCa[0]) * Acap / (2.0 * F * vmyo) -
                        Jup * vnsr / vmyo + Jdiff * vss / vmyo);
    r[TR_Cass] = Bcass * (-(ICaL_ss - 2.0 * INaCa[1]) * Acap / (2.0 * F * vss) +
                          Jrel * vjsr / vss - Jdiff);
    r[TR_Nai] = -(INa + INaL + 3.0 * INaCa[0] + ICaNa_i + 3.0 * INaK + INab) * Acap / (F * vmyo) +
                JdiffNa * vss / vmyo;
    r[TR_Nass] = -(ICaNa_ss + 3.0 * INaCa[1]) * Acap / (F * vss) - JdiffNa;

    const double Iion = INa + INaL + Ito + (ICaL_ss + ICaL_i) + (ICaNa_ss + ICaNa_i) +
                        (ICaK_ss + ICaK_i) + IKr + IKs + IK1 + INaCa[0] + INaCa[1] + INaK + INab +
                        IKb + IpCa + ICab + IClCa + IClb + Ikatp;
    const double Ist = orc_stimulus(t, st);
    r[TR_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion;
    *istim = Ist;
}

/* =================================================================== Gaur
 * Gaur et al. 2021 human ventricular myocyte with spontaneous SR release (29 states):
 * constants and initial state of Gaur_2021.H:460-647 (GaurinitConsts), right-hand side of
 * :649-921, state order of its STATES_INDEX enum.  Tissue "myocyte" only (Gaur.C:93-96).
 * V in mV (Gaur.C:121-126), Im = Iion_cm.  The model's own factors are kept as written
 * (e.g. Acap / (2 F vmyo) in the K_i and Na_i balances).
 * Rush-Larsen classification: the Hodgkin-Huxley gates and the two release fluxes as
 * gates; ITo_aa (linear in itself); the SOICR timer and its refractory variable relax
 * exponentially (tau 1 us / 0.1 ms) or grow at rate 1, which the step takes exactly; the
 * rest forward Euler.
 * Parity pins: the reference's own expressions at seeded states
 * (tests/test_oracle_model_vectors.py). */
enum { GA_V, GA_nai, GA_nass, GA_ki, GA_kss, GA_cai, GA_cai2, GA_cass, GA_cansr, GA_cajsr,
       GA_cacsr, GA_m, GA_h, GA_j, GA_ml, GA_hl, GA_d, GA_fca, GA_xr, GA_xs1, GA_xs2, GA_aa,
       GA_Jrel2, GA_Jrel1, GA_CaMKt, GA_tjsrol, GA_A, GA_fs, GA_ff, GA_NS };
/* rescalable conductances and permeabilities */
enum { GAC_GNa, GAC_GKr, GAC_PCa, GAC_GKs, GAC_GNaL, GAC_GK1, GAC_Gto, GAC_Gncx, GAC_Pnak,
       GAC_GKb, GAC_PNab, GAC_PCab, GAC_GpCa, GAC_NC };

static void gaur_init(double* c, double* y)
{
    c[GAC_GNa] = 14.838;
    c[GAC_GKr] = 0.5 * 0.015;
    c[GAC_PCa] = 0.0002;
    c[GAC_GKs] = 0.021;
    c[GAC_GNaL] = 0.0075;
    c[GAC_GK1] = 0.75 * 0.2;
    c[GAC_Gto] = 0.4 * 0.5;
    c[GAC_Gncx] = 1.0 * 0.0008;
    c[GAC_Pnak] = 1.0 * 30.0;
    c[GAC_GKb] = 0.003;
    c[GAC_PNab] = 3.75e-10;
    c[GAC_PCab] = 1.0 * 2.5e-08;
    c[GAC_GpCa] = 0.0575;
    static const double y0[GA_NS] = {
        -87.3, 6.43, 6.43, 140.76, 140.76, 6.75e-05, 6.81e-05, 6.75e-05, 1.37, 1.31, 1.3,
        0.0022, 0.631, 0.631, 0.001, 0.301, 4.42e-07, 0.988, 0.153, 0.054, 0.046, 0.995,
        8.59e-10, 9.71e-22, 8.23910999999999914e-03, 100.0, 100.0, 0.0, 0.0};
    memcpy(y, y0, sizeof y0);
}

/* Na/Ca exchanger of one compartment (:764-819), gfrac its share of Gncx */
static double gaur_ncx(double na, double ca, double hca, double hna, double gfrac)
{
    const double nao = 140.0, cao = 1.0 * 1.8, kna1 = 15.0, kna2 = 5.0, kna3 = 88.12,
                 kcaon = 1500000.0, kcaoff = 5000.0, wca = 60000.0, wna = 60000.0,
                 wnaca = 5000.0, KmCaAct = 0.00015;
    const double h10 = 12.5 + 1.0 + nao / kna1 * (1.0 + nao / kna2);
    const double h11 = nao * nao / (h10 * kna1 * kna2), h12 = 1.0 / h10;
    const double k1 = h12 * cao * kcaon, k2 = kcaoff, k5 = kcaoff;
    const double allo = 1.0 / (1.0 + KmCaAct / ca * KmCaAct / ca);
    const double h4 = 1.0 + na / kna1 * (1.0 + na / kna2);
    const double h1 = 1.0 + na / kna3 * (1.0 + hna);
    const double h5 = na * na / (h4 * kna1 * kna2), h6 = 1.0 / h4;
    const double h7 = 1.0 + nao / kna3 * (1.0 + 1.0 / hna);
    const double h2 = na * hna / (kna3 * h1), h3 = 1.0 / h1;
    const double h8 = nao / (kna3 * hna * h7), h9 = 1.0 / h7;
    const double k6 = h6 * ca * kcaon;
    const double k3pp = h8 * wnaca, k4pp = h2 * wnaca;
    const double k3 = h9 * wca + k3pp, k4 = h3 * wca / hca + k4pp;
    const double k7 = h5 * h2 * wna, k8 = h8 * h11 * wna;
    const double x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
    const double x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
    const double x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
    const double x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
    const double sx = x1 + x2 + x3 + x4;
    const double E1 = x1 / sx, E2 = x2 / sx, E3 = x3 / sx, E4 = x4 / sx;
    const double JCa = E2 * k2 - E1 * k1;
    const double JNa = 3.0 * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
    return gfrac * allo * (1.0 * JNa + 2.0 * JCa);
}

static void gaur_eval(double t, const double* c, const double* y, int solve_vm,
                      const orc_stim_t* st, double* r, int* g, double* xi, double* ta,
                      double* iion, double* istim)
{
    /* cell geometry and fixed constants (:463-642) */
    const double F = 96485.0, L = 0.017, R = 8314.0, T = 310.0, cao = 1.0 * 1.8, cli = 19.53,
                 clo = 100.0, ko = 5.4, nao = 140.0, pi = 3.14, rad = 0.0011, vmyo1frac = 0.4;
    const double Ageo = 2.0 * pi * rad * rad + 2.0 * pi * rad * L;
    const double vcell = 1000.0 * pi * rad * rad * L;
    const double Acap = 2.0 * Ageo;
    const double vjsr = 0.0048 * vcell / 2.0, vmyo = 0.68 * vcell, vnsr = 0.0552 * vcell,
                 vss = 0.02 * vcell / 2.0, vcsr = vjsr;
    const double vmyo1 = vmyo * vmyo1frac, vnsr1 = vnsr / 2.0, vnsr2 = vnsr / 2.0,
                 vmyo2 = vmyo - vmyo1;
    const double ECl = R * T / F * log(cli / clo);
    const double KmCaMK = 0.065, SOICR = 10.0;
    const double V = y[GA_V], nai = y[GA_nai], nass = y[GA_nass], ki = y[GA_ki],
                 kss = y[GA_kss], cai = y[GA_cai], cai2 = y[GA_cai2], cass = y[GA_cass];
    for (int i = 0; i < GA_NS; ++i) g[i] = 0;
#define GA_GATE(idx, inf, tau) do { const double i_ = (inf), t_ = (tau); \
        r[idx] = (i_ - y[idx]) / t_; g[idx] = 1; xi[idx] = i_; ta[idx] = t_; } while (0)

    const double Jdiff = (cass - cai) / 0.2, JdiffK = (kss - ki) / 2.0,
                 JdiffNa = (nass - nai) / 2.0;
    /* CaMK, reversal potentials (:656-669) */
    const double EK = R * T / F * log(ko / ki), ENa = R * T / F * log(nao / nai);
    const double vfrt = V * F / (R * T);
    const double CaMKb = 0.05 * (1.0 - y[GA_CaMKt]) / (1.0 + 0.0015 / cass);
    const double EKs = R * T / F * log((ko + 0.01833 * nao) / (ki + 0.01833 * nai));
    const double CaMKa = CaMKb + y[GA_CaMKt];
    r[GA_CaMKt] = 0.05 * CaMKb * (CaMKb + y[GA_CaMKt]) - 0.00068 * y[GA_CaMKt];
    const double CaMK_f = 1.0 / (1.0 + KmCaMK / CaMKa);

    /* I_Na (:671-689) */
    const double aa_h = V >= -40.0 ? 0.0 : 0.057 * exp(-(V + 80.0) / 6.8);
    const double aa_j = V >= -40.0 ? 0.0
        : (-25428.0 * exp(0.2444 * V) - 6.948e-06 * exp(-0.04391 * V)) * (V + 37.78) /
              (1.0 + exp(0.311 * (V + 79.23)));
    const double aa_m = 1.0 / (1.0 + exp((-60.0 - V) / 5.0));
    const double bb_h = V >= -40.0 ? 0.77 / (0.13 * (1.0 + exp(-(V + 10.66) / 11.1)))
                                   : 2.7 * exp(0.079 * V) + 310000.0 * exp(0.3485 * V);
    const double bb_j = V >= -40.0 ? 0.6 * exp(0.057 * V) / (1.0 + exp(-0.1 * (V + 32.0)))
                                   : 0.02424 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)));
    const double bb_m = 0.1 / (1.0 + exp((V + 35.0) / 5.0)) + 0.1 / (1.0 + exp((V - 50.0) / 200.0));
    const double eh = 1.0 + exp((V + 71.55) / 7.43), em = 1.0 + exp((-56.86 - V) / 9.03);
    const double h_inf = 1.0 / (eh * eh), m_inf = 1.0 / (em * em);
    const double INa = c[GAC_GNa] * y[GA_m] * y[GA_m] * y[GA_m] * y[GA_h] * y[GA_j] * (V - ENa);
    GA_GATE(GA_h, h_inf, 1.0 / (aa_h + bb_h));
    GA_GATE(GA_j, h_inf, 1.0 / (aa_j + bb_j));
    GA_GATE(GA_m, m_inf, aa_m * bb_m);

    /* INaL (:691-699) */
    const double aml = 0.32 * (V == -47.13 ? 10.0 : -(V + 47.13) / (exp(-0.1 * (V + 47.13)) - 1.0));
    const double bml = 0.08 * exp(-V / 11.0);
    const double INaL = c[GAC_GNaL] * y[GA_ml] * y[GA_ml] * y[GA_ml] * y[GA_hl] * (V - ENa);
    GA_GATE(GA_hl, 1.0 / (1.0 + exp((V + 91.0) / 6.1)), 600.0);
    GA_GATE(GA_ml, aml / (aml + bml), 1.0 / (aml + bml));

    /* ICaL (:701-724) */
    const double g1 = V == 0.0 ? 1.0 : vfrt / (exp(vfrt) - 1.0);
    const double g2 = V == 0.0 ? 1.0 : 2.0 * vfrt / (exp(2.0 * vfrt) - 1.0);
    const double PhiCaK = F * (0.75 * kss * exp(vfrt) - 0.75 * ko) * g1;
    const double PhiCaL = 2.0 * F * (cass * exp(2.0 * (V * F / (R * T))) - 0.341 * cao) * g2;
    const double PhiCaNa = F * (0.75 * nass * exp(vfrt) - 0.75 * nao) * g1;
    const double f = y[GA_ff] * y[GA_fs];
    const double ICaL = c[GAC_PCa] * PhiCaL * y[GA_d] * f * y[GA_fca];
    const double ICaK = 0.0003574 * c[GAC_PCa] * PhiCaK * y[GA_d] * f * y[GA_fca];
    const double ICaNa = 0.00125 * c[GAC_PCa] * PhiCaNa * y[GA_d] * f * y[GA_fca];
    const double f_inf = 1.0 / (1.0 + exp((V - -22.9) / 4.9)) + 0.35 / (1.0 + exp((45.0 - V) / 20.0));
    const double fca_inf = 0.3 / (1.0 - (ICaL > 0.0 ? 0.0 : ICaL / 0.05)) +
                           0.55 / (1.0 + cass / 0.003) + 0.15;
    GA_GATE(GA_d, 1.0 / (1.0 + exp(-(V - 3.4) / 6.2)),
            0.6 + 1.0 / (exp(-0.05 * (V + 6.0)) + exp(0.09 * (V + 14.0))));
    GA_GATE(GA_fca, fca_inf, 10.0 * CaMK_f + 0.5 + 1.0 / (1.0 + cass / 0.003));
    GA_GATE(GA_ff, f_inf, 7.0 + 1.0 / (0.0045 * exp(-(V + 20.0) / 10.0) + 0.0045 * exp((V + 20.0) / 10.0)));
    GA_GATE(GA_fs, f_inf, 70.0 + 1.0 / (3.5e-05 * exp(-(V + 5.0) / 4.0) + 3.5e-05 * exp((V + 5.0) / 6.0)));

    /* IKr, IKs, IK1 (:726-747) */
    const double IKr = c[GAC_GKr] * sqrt(ko / 5.4) * y[GA_xr] * (1.0 / (1.0 + exp((V + 22.0) / 15.0))) * (V - EK);
    GA_GATE(GA_xr, 1.0 / (1.0 + exp(-(V + 56.8) / 17.8)),
            12.98 + 1.0 / (0.3652 * exp((V - 31.66) / 3.869) + 4.123e-05 * exp(-(V - 47.78) / 20.38)));
    const double KsCa = 1.0 + 0.6 / (1.0 + pow(3.8e-05 / cai, 1.4));
    const double xs1_inf = 1.0 / (1.0 + exp(-(V - 25.1) / 37.1));
    const double IKs = c[GAC_GKs] * KsCa * y[GA_xs1] * y[GA_xs2] * (V - EKs);
    GA_GATE(GA_xs1, xs1_inf, 300.0 + 1.0 / (1e-06 * exp((V + 50.0) / 20.0) + 0.04 * exp(-(V + 50.0) / 20.0)));
    GA_GATE(GA_xs2, xs1_inf, 1.0 / (0.01 * exp((V - 50.0) / 20.0) + 0.0193 * exp(-(V + 66.54) / 31.0)));
    const double IK1 = c[GAC_GK1] * sqrt(ko / 5.4) * (1.0 / (1.0 + exp((V + 79.3 - 2.6 * ko) / 19.6))) * (V - EK);

    /* ITo gate (:749-753): d aa/dt = alpha (1 - aa) - beta aa */
    const double alpha_aa = 0.025 / (1.0 + exp((V + 58.0) / 5.0));
    const double beta_aa = 1.0 / (5.0 * (1.0 + exp((V + 19.0) / -9.0)));
    const double rito2 = 1.0 / (1.0 + exp(-(V + 10.0) / 5.0));
    r[GA_aa] = alpha_aa * (1.0 - y[GA_aa]) - beta_aa * y[GA_aa];
    g[GA_aa] = 1; xi[GA_aa] = alpha_aa / (alpha_aa + beta_aa); ta[GA_aa] = 1.0 / (alpha_aa + beta_aa);

    /* CICR (:755-769) */
    const double camk8 = 1.0 + 1.0 / (1.0 + pow(KmCaMK / CaMKa, 8.0));
    const double trel1 = 1.0 * 20.0 * camk8 / (1.0 + pow(0.5 / y[GA_cajsr], 8.0));
    const double trel2 = 50.0 * camk8 / (1.0 + pow(0.5 / y[GA_cacsr], 8.0));
    const double Jgap = (cai - cai2) / 3.0;
    const double grel = 2.0 * (1.0 - exp(-y[GA_tjsrol] / 0.5)) * exp(-y[GA_tjsrol] / 5.0);
    const double Jrel = y[GA_Jrel1] + grel * (y[GA_cajsr] - cass);
    /* A and the timer: x' = -x / tau (exact exponential decay) or x' = 1 */
    if (y[GA_tjsrol] < 5.0) {
        r[GA_A] = -y[GA_A] / 0.1; g[GA_A] = 1; xi[GA_A] = 0.0; ta[GA_A] = 0.1;
    } else {
        r[GA_A] = 1.0;
    }
    if (y[GA_cajsr] > SOICR && y[GA_A] > 45.0) {
        r[GA_tjsrol] = -y[GA_tjsrol] / 0.001; g[GA_tjsrol] = 1; xi[GA_tjsrol] = 0.0;
        ta[GA_tjsrol] = 0.001;
    } else {
        r[GA_tjsrol] = 1.0;
    }

    /* exchanger, pump, background currents (:771-875) */
    const double hca = exp(0.167 * V * F / (R * T)), hna = exp(0.5224 * V * F / (R * T));
    const double INaCa_i = gaur_ncx(nai, cai, hca, hna, 0.8 * c[GAC_Gncx]);
    const double INaCa_ss = gaur_ncx(nass, cass, hca, hna, 0.2 * c[GAC_Gncx]);
    const double Knai = 9.073 * exp(-0.155 * V * F / (3.0 * R * T));
    const double Knao = 27.78 * exp((1.0 - -0.155) * V * F / (3.0 * R * T));
    const double P = 4.2 / (1.0 + 1e-07 / 1.698e-07 + nai / 224.0 + ki / 292.0);
    const double ni = nai / Knai, no = nao / Knao, kk = ki / 0.5, kko = ko / 0.3582;
    const double dni = (1.0 + ni) * (1.0 + ni) * (1.0 + ni) + (1.0 + kk) * (1.0 + kk) - 1.0;
    const double dno = (1.0 + no) * (1.0 + no) * (1.0 + no) + (1.0 + kko) * (1.0 + kko) - 1.0;
    const double a1 = 949.5 * (ni * ni * ni) / dni;
    const double a3 = 1899.0 * (kko * kko) / dno;
    const double b2 = 39.4 * (no * no * no) / dno;
    const double b3 = 79300.0 * P * 1e-07 / (1.0 + 9.8 / 1.698e-07);
    const double b4 = 40.0 * (kk * kk) / dni;
    const double a2 = 687.2, a4 = 639.0 * 9.8 / 1.698e-07 / (1.0 + 9.8 / 1.698e-07), b1 = 182.4 * 0.05;
    const double x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
    const double x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
    const double x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
    const double x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
    const double sx = x1 + x2 + x3 + x4;
    const double JnakK = 2.0 * (x4 / sx * b1 - x3 / sx * a1);
    const double JnakNa = 3.0 * (x1 / sx * a3 - x2 / sx * b3);
    const double INaK = c[GAC_Pnak] * (1.0 * JnakNa + 1.0 * JnakK);
    const double IKb = c[GAC_GKb] * (1.0 / (1.0 + exp(-(V - 14.48) / 18.34))) * (V - EK);
    const double INab = c[GAC_PNab] * F * (nai * exp(vfrt) - nao) * g1;
    const double ICab = c[GAC_PCab] * 2.0 * F * (cai * exp(2.0 * vfrt) - 0.341 * cao) * g2;
    const double IpCa = c[GAC_GpCa] * cai / (0.0005 + cai);

    /* SR uptake and the concentrations (:877-903) */
    const double Jleak = 0.005 * y[GA_cansr] / 15.0;
    const double Jtr = (y[GA_cansr] - y[GA_cajsr]) / 100.0, Jtr2 = (y[GA_cansr] - y[GA_cacsr]) / 100.0;
    const double Jupnp = 1.0 * 0.005 * cai / (cai + 0.001), Jupnp2 = 1.0 * 0.005 * cai2 / (cai2 + 0.001);
    const double Jupp = 1.0 * 3.0 * 0.005 * cai / (cai + 0.001 - 0.0002);
    const double Jupp2 = 1.0 * 3.0 * 0.005 * cai2 / (cai2 + 0.001 - 0.0002);
    const double Jup = (1.0 - CaMK_f) * Jupnp + CaMK_f * Jupp - Jleak;
    const double Jup2 = (1.0 - CaMK_f) * Jupnp2 + CaMK_f * Jupp2 - Jleak;
    const double Bcacsr = 1.0 / (1.0 + 10.0 * 0.8 / ((0.8 + y[GA_cacsr]) * (0.8 + y[GA_cacsr])));
    const double Bcai = 1.0 / (1.0 + 0.05 * 0.00238 / ((0.00238 + cai) * (0.00238 + cai)) +
                               0.07 * 0.0005 / ((0.0005 + cai) * (0.0005 + cai)));
    const double Bcai2 = 1.0 / (1.0 + 0.05 * 0.00238 / ((0.00238 + cai2) * (0.00238 + cai2)) +
                                0.07 * 0.0005 / ((0.0005 + cai2) * (0.0005 + cai2)));
    const double Bcajsr = 1.0 / (1.0 + 10.0 * 0.8 / ((0.8 + y[GA_cajsr]) * (0.8 + y[GA_cajsr])));
    const double Bcass = 1.0 / (1.0 + 0.047 * 0.00087 / ((0.00087 + cass) * (0.00087 + cass)) +
                                1.124 * 0.0087 / ((0.0087 + cass) * (0.0087 + cass)));
    r[GA_cansr] = Jup * (vnsr1 / vnsr) + Jup2 * (vnsr2 / vnsr) - Jtr * vjsr / vnsr - Jtr2 * vcsr / vnsr;
    r[GA_ki] = -(IKr + IKs + IK1 + IKb - 2.0 * INaK) * Acap / (2.0 * F * vmyo) + JdiffK * vss / vmyo;
    r[GA_kss] = -ICaK * Acap / (2.0 * F * vss) - JdiffK;
    r[GA_nai] = -(INa + INaL + 3.0 * INaCa_i + 3.0 * INaK + INab) * Acap / (2.0 * F * vmyo) +
                JdiffNa * vss / vmyo;
    r[GA_nass] = -(ICaNa + 3.0 * INaCa_ss) * Acap / (2.0 * F * vss) - JdiffNa;
    r[GA_cacsr] = Bcacsr * (Jtr2 - y[GA_Jrel2]);
    r[GA_cai] = Bcai * (-(IpCa + ICab - 2.0 * INaCa_i) * Acap / (2.0 * 2.0 * F * vmyo1) -
                        Jup * vnsr1 / vmyo1 + Jdiff * vss / vmyo1 - Jgap);
    r[GA_cai2] = Bcai2 * (y[GA_Jrel2] * (vcsr / vmyo2) + Jgap * vmyo2 / vmyo1 - Jup2 * vnsr2 / vmyo2);
    r[GA_cajsr] = Bcajsr * (Jtr - Jrel);
    r[GA_cass] = Bcass * (-(ICaL - 2.0 * INaCa_ss) * Acap / (2.0 * 2.0 * F * vss) +
                          Jrel * vjsr / vss - Jdiff);

    /* release fluxes and ITo (:906-913) */
    const double kito2 = 1.0 - 1.0 / (1.0 + Jrel / 0.4 * Jrel / 0.4);
    const double Rel1 = (-ICaL + 2.0 * INaCa_ss) * (Acap / (2.0 * vss * F)) + Jrel * (vjsr / vss) - Jdiff;
    const double Rel2 = Jgap + y[GA_Jrel2] * (vcsr / vmyo2) - Jup2 * (vnsr2 / vmyo2);
    const double ITo = c[GAC_Gto] * y[GA_aa] * rito2 * kito2 * (V - ECl);
    const double Jrel1_inf = Rel1 > 0.0 ? 1.0 * 60.0 * Rel1 * camk8 / (1.0 + pow(0.75 / y[GA_cajsr], 8.0)) : 0.0;
    const double Jrel2_inf = Rel2 > 0.0 ? 1.0 * (250.0 * Rel2 * camk8) / (1.0 + pow(0.75 / y[GA_cacsr], 8.0)) : 0.0;
    GA_GATE(GA_Jrel1, Jrel1_inf, 0.001 > trel1 ? 0.001 : trel1);
    GA_GATE(GA_Jrel2, Jrel2_inf, 0.001 > trel2 ? 0.001 : trel2);
#undef GA_GATE

    const double Iion = INa + INaL + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + ITo + INaCa_i +
                        INaCa_ss + INaK + INab + IKb + IpCa + ICab;
    const double Ist = orc_stimulus(t, st);
    r[GA_V] = solve_vm ? -Iion - Ist : 0.0;
    *iion = Iion;
    *istim = Ist;
}

/* ============================================================== dispatch */
int orc_model_nstates(int model)
{
    switch (model) {
    case ORC_BO: return 4;
    case ORC_TNNP: return TN_NS;
    case ORC_ORD: return OR_NS;
    case ORC_AP: return 2;
    case ORC_FDA: return 4;
    case ORC_CRN: return CR_NS;
    case ORC_PYS: return PY_NS;
    case ORC_STW: return SW_NS;
    case ORC_GRD: return GR_NS;
    case ORC_TOR: return TR_NS;
    case ORC_GAUR: return GA_NS;
    default: return -1;
    }
}

int orc_model_init(int model, int tissue, double* c, double* y)
{
    memset(c, 0, sizeof(double) * ORC_MAX_CONSTS);
    switch (model) {
    case ORC_BO:
        if (tissue < 1 || tissue > 3) return -1;
        bo_init(tissue, c, y); return 0;
    case ORC_TNNP:
        if (tissue < 1 || tissue > 3) return -1;
        tnnp_init(tissue, c, y); return 0;
    case ORC_ORD:
        if (tissue != ORC_MYOCYTE) return -1;
        ord_init(tissue, c, y); return 0;
    case ORC_AP:
        if (tissue != ORC_MYOCYTE) return -1;
        ap_init(c, y); return 0;
    case ORC_FDA:
        if (tissue < ORC_DIM1 || tissue > ORC_DIM3) return -1;
        fda_init(tissue, c, y); return 0;
    case ORC_CRN:
        if (tissue != ORC_MYOCYTE) return -1;
        crn_init(c, y); return 0;
    case ORC_PYS:
        if (tissue != ORC_MYOCYTE) return -1;
        pys_init(c, y); return 0;
    case ORC_STW:
        if (tissue != ORC_MYOCYTE) return -1;
        stw_init(c, y); return 0;
    case ORC_GRD:
        if (tissue != ORC_MYOCYTE) return -1;
        grd_init(c, y); return 0;
    case ORC_TOR:
        if (tissue < 1 || tissue > 3) return -1;
        tor_init(tissue, c, y); return 0;
    case ORC_GAUR:
        if (tissue != ORC_MYOCYTE) return -1;
        gaur_init(c, y); return 0;
    default: return -1;
    }
}

int orc_model_const_index(int model, const char* name)
{
    if (model == ORC_TNNP) {
        if (!strcmp(name, "g_Na")) return TC_gNa;
        if (!strcmp(name, "g_Kr")) return TC_gKr;
        if (!strcmp(name, "g_CaL")) return TC_gCaL;
        if (!strcmp(name, "g_Ks")) return TC_gKs;
    } else if (model == ORC_ORD) {
        if (!strcmp(name, "GNa")) return OC_GNa;
        if (!strcmp(name, "GKr")) return OC_GKr;
        if (!strcmp(name, "PCa")) return OC_PCa;
        if (!strcmp(name, "GKs")) return OC_GKs;
    } else if (model == ORC_CRN) {
        if (!strcmp(name, "gNa")) return CC_gNa;
        if (!strcmp(name, "gKr")) return CC_gKr;
        if (!strcmp(name, "gCaL")) return CC_gCaL;
        if (!strcmp(name, "gKs")) return CC_gKs;
    } else if (model == ORC_PYS) {
        if (!strcmp(name, "gNa")) return PC_gNa;
        if (!strcmp(name, "gKr")) return PC_gKr;
        if (!strcmp(name, "gCaL")) return PC_gCaL;
        if (!strcmp(name, "gKs")) return PC_gKs;
    } else if (model == ORC_STW) {
        if (!strcmp(name, "g_Na")) return SC_gNa;
        if (!strcmp(name, "g_Kr")) return SC_gKr;
        if (!strcmp(name, "g_CaL")) return SC_gCaL;
        if (!strcmp(name, "g_Ks")) return SC_gKs;
    } else if (model == ORC_GAUR) {
        if (!strcmp(name, "GNa")) return GAC_GNa;
        if (!strcmp(name, "GKr")) return GAC_GKr;
        if (!strcmp(name, "PCa")) return GAC_PCa;
        if (!strcmp(name, "GKs")) return GAC_GKs;
    } else if (model == ORC_TOR) {
        if (!strcmp(name, "GNa")) return RC_GNa;
        if (!strcmp(name, "GKr")) return RC_GKr;
        if (!strcmp(name, "PCa")) return RC_PCa;
        if (!strcmp(name, "GKs")) return RC_GKs;
    } else if (model == ORC_GRD) {
        if (!strcmp(name, "g_Na")) return GC_gNa;
        if (!strcmp(name, "g_Kr")) return GC_gKr;
        if (!strcmp(name, "g_CaL")) return GC_pCa;
        if (!strcmp(name, "g_Ks")) return GC_gKs;
    } else if (model == ORC_BO) {
        if (!strcmp(name, "tauFi")) return BO_tfi;
        if (!strcmp(name, "tauSi")) return BO_tsi;
    }
    return -1;
}

/* ALGEBRAIC variables of the models that have them in the reference's order
 * (BuenoOrovio_2008.H:30-43 names, :166-225 values; AlievPanfilov_1996.H:28-32, :94-120);
 * returns their count, or -1 for a model whose algebraic list is not restated here.
 * Parity pins (tests/test_oracle_algebraic.py): no reference output holds these columns;
 * they are tied to the right-hand sides the golden traces pin (total current, stimulus,
 * and the Rush-Larsen (x_inf, tau) pairs of each Heaviside region). */
int orc_model_algebraic(int model, double t, const double* c, const double* y,
                        const orc_stim_t* st, double* a)
{
    if (model == ORC_BO) {
        const double u = y[0], v = y[1], w = y[2], s = y[3];
        const double Hv = H(u - c[BO_thv]), Hw = H(u - c[BO_thw]);
        const double Ho = H(u - c[BO_tho]), Hvm = H(u - c[BO_thvm]);
        a[5] = c[BO_tso1] + (c[BO_tso2] - c[BO_tso1]) *
               (1.0 + tanh(c[BO_kso] * (u - c[BO_uso]))) / 2.0;              /* tauSo */
        a[6] = (1.0 - Ho) * c[BO_to1] + Ho * c[BO_to2];                      /* tauO */
        a[0] = -v * Hv * (u - c[BO_thv]) * (c[BO_uu] - u) / c[BO_tfi];       /* Jfi */
        a[1] = (u - c[BO_uo]) * (1.0 - Hw) / a[6] + Hw / a[5];               /* Jso */
        a[2] = -Hw * w * s / c[BO_tsi];                                      /* Jsi */
        a[3] = a[0] + a[1] + a[2];                                           /* Jion */
        a[7] = (1.0 - Hvm) * c[BO_tv1m] + Hvm * c[BO_tv2m];                  /* tauVMinus */
        a[8] = c[BO_tw1m] + (c[BO_tw2m] - c[BO_tw1m]) *
               (1.0 + tanh(c[BO_kwm] * (u - c[BO_uwm]))) / 2.0;              /* tauWMinus */
        a[9] = (1.0 - Hw) * c[BO_ts1] + Hw * c[BO_ts2];                      /* tauS */
        a[10] = H(c[BO_thvm] - u);                                           /* vInfty */
        a[11] = (1.0 - Ho) * (1.0 - u / c[BO_twinf]) + Ho * c[BO_winfstar];  /* wInfty */
        a[4] = orc_stimulus(t, st);                                          /* Istim */
        return 12;
    }
    if (model == ORC_AP) {
        const double u = y[0], rr = y[1];
        a[0] = c[3] + c[5] * rr / (c[6] + u);                               /* AV_eps */
        a[1] = orc_stimulus(t, st);                                          /* Istim */
        a[2] = c[4] * u * (u - c[7]) * (u - 1.0) + rr * u;                   /* Iion_cm */
        return 3;
    }
    return -1;
}

void orc_model_eval(int model, int tissue, double t, const double* c, const double* y,
                    int solve_vm, const orc_stim_t* st, double* r, int* g, double* xi,
                    double* ta, double* iion, double* istim)
{
    switch (model) {
    case ORC_BO: bo_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_TNNP: tnnp_eval(tissue, t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_ORD: ord_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_AP: ap_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_FDA: fda_eval(c, y, r, g, iion); *istim = 0.0; break;
    case ORC_CRN: crn_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_PYS: pys_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_STW: stw_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_GRD: grd_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_TOR: tor_eval(tissue, t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    case ORC_GAUR: gaur_eval(t, c, y, solve_vm, st, r, g, xi, ta, iion, istim); break;
    default: break;
    }
}

double orc_model_vm_mV(int model, const double* y)
{
    /* BuenoOrovio.H:78-81, AlievPanfilov.H:78-81; TNNP/ORd store mV directly */
    if (model == ORC_BO) return y[0] * 85.7 - 84.0;
    if (model == ORC_AP) return y[0] * 100 - 80;
    return y[0];
}
