// Bueno-Orovio-Cherry-Fenton 2008 minimal ventricular model on gfx950.
// Equations: src/ionicModels/BuenoOrovio/BuenoOrovio_2008.H:166-225 (right-hand side),
// :116-163 (tissue columns); unit scaling: BuenoOrovio.C:107-152
// (u = (1000 Vm + 84)/85.7, Im = 85.7 Jion).
//
// Every gate of this model relaxes as (xinf - x)/tau with (xinf, tau) that depend on u
// only, so with u held at V^n for the step the Rush-Larsen update is the exact solution
// of the reference's gate ODEs.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct BuenoOrovio {
    static constexpr int NS = 4;   // u v w s
    static constexpr int NP = 28;  // constants, order of BuenoOrovio_2008Names.H
    static constexpr double VSCALE = 85.7, VOFF = 84.0, ISCALE = 85.7, VMUL = 1000.0;
    static constexpr double TSCALE = 1000.0;  // model ms per second
    enum { uo, uu, thv, thw, thvm, tho, tv1m, tv2m, tvp, tw1m, tw2m, kwm, uwm, twp, tfi,
           to1, to2, tso1, tso2, kso, uso, ts1, ts2, ks, us, tsi, twinf, winfstar };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue < 1 || tissue > 3) return false;
        const bool e = tissue == 1, m = tissue == 2;
        const double tab[NP] = {
            0.0, e ? 1.55 : m ? 1.61 : 1.56, 0.3, 0.13, e ? 0.006 : m ? 0.1 : 0.2,
            m ? 0.005 : 0.006, e ? 60.0 : m ? 80.0 : 75.0, e ? 1150.0 : m ? 1.4506 : 10.0,
            1.4506, e ? 60.0 : m ? 70.0 : 6.0, e ? 15.0 : m ? 8.0 : 140.0, e ? 65.0 : 200.0,
            e ? 0.03 : 0.016, e ? 200.0 : 280.0, e ? 0.11 : m ? 0.078 : 0.1,
            e ? 400.0 : m ? 410.0 : 470.0, m ? 7.0 : 6.0, e ? 30.0181 : m ? 91.0 : 40.0,
            e ? 0.9957 : m ? 0.8 : 1.2, e ? 2.0458 : m ? 2.1 : 2.0, m ? 0.6 : 0.65, 2.7342,
            e ? 16.0 : m ? 4.0 : 2.0, 2.0994, 0.9087, e ? 1.8875 : m ? 3.3849 : 2.9013,
            e ? 0.07 : m ? 0.01 : 0.0273, e ? 0.94 : m ? 0.5 : 0.78};
        for (int i = 0; i < NP; ++i) p[i] = tab[i];
        y[0] = 0.0; y[1] = 1.0; y[2] = 1.0; y[3] = 0.0;
        return true;
    }

    // Jion at (u, v, w, s): the three currents of BuenoOrovio_2008.H:171-188
    template <class T> CG_HD static T jion(T u, const T* y, const T* p)
    {
        const bool Hv = pos(u - p[thv]), Hw = pos(u - p[thw]), Ho = pos(u - p[tho]);
        const T tso = p[tso1] + (p[tso2] - p[tso1]) * (T(1) + ctanh(p[kso] * (u - p[uso]))) / T(2);
        const T to = Ho ? p[to2] : p[to1];
        const T jfi = Hv ? -y[1] * (u - p[thv]) * (p[uu] - u) / p[tfi] : T(0);
        const T jso = Hw ? T(1) / tso : (u - p[uo]) / to;
        const T jsi = Hw ? -y[2] * y[3] / p[tsi] : T(0);
        return jfi + jso + jsi;
    }

    // Rates of v, w, s at u = y[0] (BuenoOrovio_2008.H:189-220), through an update policy.
    template <class T, class U> CG_HD static void gate_rates(const T* y, U& out, const T* p)
    {
        const T u = y[0];
        const bool Hv = pos(u - p[thv]), Hw = pos(u - p[thw]), Ho = pos(u - p[tho]),
                   Hvm = pos(u - p[thvm]);
        // v: relaxes to 0 with tau_v+ above theta_v, else to H(theta_v- - u) with tau_v-
        const T vinf = Hv ? T(0) : (pos(p[thvm] - u) ? T(1) : T(0));
        const T tv = Hv ? p[tvp] : (Hvm ? p[tv2m] : p[tv1m]);
        // w: relaxes to 0 with tau_w+ above theta_w, else to w_inf with tau_w-
        const T twm = p[tw1m] + (p[tw2m] - p[tw1m]) * (T(1) + ctanh(p[kwm] * (u - p[uwm]))) / T(2);
        const T winf = Hw ? T(0) : (Ho ? p[winfstar] : T(1) - u / p[twinf]);
        const T tw = Hw ? p[twp] : twm;
        // s
        const T sinf = (T(1) + ctanh(p[ks] * (u - p[us]))) / T(2);
        const T ts = Hw ? p[ts2] : p[ts1];
        out.gate(1, vinf, tv);
        out.gate(2, winf, tw);
        out.gate(3, sinf, ts);
    }

    // Advance v, w, s by dt with u frozen; y[0] is u.
    template <class T> CG_HD static void gates(T* y, T dt, const T* p)
    {
        RlStep<T> out{y, dt};
        gate_rates(y, out, p);
    }

    // dy/dt; du/dt = -Jion - Istim when u is integrated with the gates, else 0
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, bool solve_v, T istim)
    {
        Deriv<T> out{y, d};
        gate_rates(y, out, p);
        d[0] = solve_v ? -jion(y[0], y, p) - istim : T(0);
    }

    // Single-cell step (Vm inside the ODE): RL gates, forward-Euler u with
    // du/dt = -Jion - Istim at the step start (BuenoOrovio_2008.H:221-224).
    template <class T> CG_HD static void cell_step(T* y, T dt, T istim, const T* p)
    {
        const T du = -jion(y[0], y, p) - istim;
        gates(y, dt, p);
        y[0] += dt * du;
    }
};

// Aliev-Panfilov 1996 (AlievPanfilov_1996.H:72-120; AlievPanfilov.C:107-152).
// Time is scaled by 1/12.9 and u = (1000 Vm + 80)/100; the reference returns
// Im = 100 Iion without the 1/12.9 time factor, which is kept.  Neither state has a
// (xinf, tau) form: both are integrated by forward Euler.
struct AlievPanfilov {
    static constexpr int NS = 2;
    static constexpr int NP = 8;
    static constexpr double VSCALE = 100.0, VOFF = 80.0, ISCALE = 100.0, VMUL = 1000.0;
    static constexpr double TSCALE = 1000.0 / 12.9;
    enum { Vp, Vr, Vth, eps0, kk, mu1, mu2, aa };

    static bool host_init(int tissue, double* p, double* y)
    {
        if (tissue != 4) return false;
        p[Vp] = 20.0; p[Vr] = -80.0; p[Vth] = -70.0; p[eps0] = 0.01; p[kk] = 8.0;
        p[mu1] = 0.2; p[mu2] = 0.3; p[aa] = (p[Vth] - p[Vr]) / (p[Vp] - p[Vr]);
        y[0] = 0.05; y[1] = 0.0;
        return true;
    }

    template <class T> CG_HD static T jion(T u, const T* y, const T* p)
    {
        return p[kk] * u * (u - p[aa]) * (u - T(1)) + y[1] * u;
    }

    template <class T, class U> CG_HD static void gate_rates(const T* y, U& out, const T* p)
    {
        const T u = y[0], r = y[1];
        const T e = p[eps0] + p[mu1] * r / (p[mu2] + u);
        out.rate(1, e * (-r - p[kk] * u * (u - (p[aa] + T(1)))));
    }

    template <class T> CG_HD static void gates(T* y, T dt, const T* p)
    {
        RlStep<T> out{y, dt};
        gate_rates(y, out, p);
    }

    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, bool solve_v, T istim)
    {
        Deriv<T> out{y, d};
        gate_rates(y, out, p);
        d[0] = solve_v ? -jion(y[0], y, p) - istim : T(0);
    }

    template <class T> CG_HD static void cell_step(T* y, T dt, T istim, const T* p)
    {
        const T du = -jion(y[0], y, p) - istim;
        gates(y, dt, p);
        y[0] += dt * du;
    }
};

} // namespace cg
