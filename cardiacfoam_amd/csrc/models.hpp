// Uniform device interface over the ionic models of the tissue path.
//   tissue_step: V held at the PDE value for the whole step (solveVmWithinODESolver
//                false); advances the other states by dt and returns the total ionic
//                current at the step end (model units), as the reference's solveODE
//                does with its final computeVariables call (BuenoOrovio.C:134-152,
//                TNNP.C:146-174, ORd.C:134-150, AlievPanfilov.C:134-152).
//   cell_step:   single-cell step with V integrated inside (singleCellSolver.C:186-192).
//   vt / deriv:  terms that depend on V only, and dy/dt given them (the derivatives()
//                an ODE solver calls); with V held by the PDE the V terms of a step are
//                computed once and reused by every stage.
#pragma once
#include "model_bo.hpp"
#include "model_tnnp.hpp"
#include "model_ord.hpp"
#include "model_fda.hpp"
#include "model_crn.hpp"
#include "model_pys.hpp"
#include "model_stw.hpp"
#include "model_grd.hpp"
#include "model_tor.hpp"
#include "model_gaur.hpp"

namespace cg {

enum ModelId { MODEL_BO = 0, MODEL_TNNP = 1, MODEL_ORD = 2, MODEL_AP = 3, MODEL_FDA = 4,
               MODEL_CRN = 5, MODEL_PYS = 6, MODEL_STW = 7, MODEL_GRD = 8,
               MODEL_TOR = 9, MODEL_GAUR = 10 };

constexpr int MAX_PARAMS = 48;
struct ModelParams {
    double p[MAX_PARAMS];
};

template <class M> struct Ops;

// Offset storage of the mixed path.  A state that stays near a fixed magnitude (Na_i,
// K_i, Cl_i, SR calcium) moves by far less than half an FP32 ulp per step while the cell
// rests (Na_i ~ 7 mM, ulp 4.8e-7, drifts by ~1e-8 per 0.02 ms step): stored as a plain
// FP32 value it never moves, and the resting tissue ahead of a wave drifts away from the
// FP64 path's, which shows as a conduction-velocity bias of ~1e-4 over 200 ms.  FP32
// handles therefore store such a state as s = x - ref with a per-model constant ref, the
// state's initial value: |s| starts at 0 and stays small, so its ulp resolves the
// increments.  at(q) is the ref of state q (used as its FP32 rounding on host and device
// alike), 0 for plain storage.  FP64 handles always store plain values.
template <class M> struct StateRef {
    static constexpr bool any = false;
    CG_HD static constexpr double at(int) { return 0.0; }
};
template <> struct StateRef<TNNP> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return q == TNNP::Ki ? 138.3 : q == TNNP::Nai ? 11.6 : q == TNNP::CaSR ? 0.2 : 0.0;
    }
};
template <> struct StateRef<ORd> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return (q == ORd::Nai || q == ORd::Nass) ? 7.0
             : (q == ORd::Ki || q == ORd::Kss) ? 145.0
             : (q == ORd::Cansr || q == ORd::Cajsr) ? 1.2 : 0.0;
    }
};
template <> struct StateRef<Courtemanche> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return q == Courtemanche::Nai ? 13.8169746 : q == Courtemanche::Ki ? 136.355230
             : q == Courtemanche::CaUp ? 1.5466812 : q == Courtemanche::CaRel ? 1.0765074 : 0.0;
    }
};
template <> struct StateRef<PerisYague> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return q == PerisYague::Nai ? 11.17 : q == PerisYague::Ki ? 139.0
             : q == PerisYague::Cli ? 30.0
             : (q == PerisYague::CaUp || q == PerisYague::CaRel) ? 1.488 : 0.0;
    }
};
template <> struct StateRef<Stewart> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return q == Stewart::Nai ? 8.8042029 : q == Stewart::Ki ? 136.7818942
             : q == Stewart::CaSR ? 3.1083689 : 0.0;
    }
};

template <> struct StateRef<Grandi> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return (q == Grandi::Nai || q == Grandi::Najn || q == Grandi::Nasl) ? 9.1519968
             : q == Grandi::Ki ? 120.0 : q == Grandi::Casr ? 0.50230583
             : q == Grandi::Csqn ? 1.1333754 : q == Grandi::NaBjn ? 3.6139606
             : q == Grandi::NaBsl ? 0.78860779 : 0.0;
    }
};

template <> struct StateRef<ToRORd> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return (q == ToRORd::Nai || q == ToRORd::Nass) ? 13.4
             : (q == ToRORd::Ki || q == ToRORd::Kss) ? 152.4
             : (q == ToRORd::Cli || q == ToRORd::Clss) ? 34.3
             : (q == ToRORd::Cansr || q == ToRORd::Cajsr) ? 1.8 : 0.0;
    }
};

template <> struct StateRef<Gaur> {
    static constexpr bool any = true;
    CG_HD static constexpr double at(int q)
    {
        return (q == Gaur::nai || q == Gaur::nass) ? 6.43
             : (q == Gaur::ki || q == Gaur::kss) ? 140.76
             : (q == Gaur::cansr || q == Gaur::cajsr || q == Gaur::cacsr) ? 1.3 : 0.0;
    }
};

// The Rush-Larsen tissue step on offset storage: y holds the full values (y[0] = v is
// set here), s the stored offsets; the increments of the step land in s, y is refreshed
// from them, and the total ionic current at the step end is returned.  RATES is the
// model's rates() call with the RlStep policy given.
// The single-cell step (V integrated inside) on offset storage, the same way, with the
// FP32 gate increments of rl_frac_precise.
template <class M, class T, class RATES>
CG_HD void cell_step_offsets(T* y, T* s, T dt, RATES rates)
{
    RlStep<T> out{y, dt, s, true};
    rates(out);
#pragma unroll
    for (int q = 0; q < M::NS; ++q) y[q] = T(StateRef<M>::at(q)) + s[q];
}

template <class M, class T, class RATES, class IION>
CG_HD T tissue_step_offsets(T* y, T* s, T dt, RATES rates, IION iion)
{
    RlStep<T> out{y, dt, s};
    rates(out);
#pragma unroll
    for (int q = 1; q < M::NS; ++q) y[q] = T(StateRef<M>::at(q)) + s[q];
    return iion();
}

template <> struct Ops<BuenoOrovio> {
    template <class T>
    CG_HD static T tissue_step(T u, T* y, T dt, const T* p, int, const T*)
    {
        y[0] = u;
        BuenoOrovio::gates(y, dt, p);
        return BuenoOrovio::jion(u, y, p);
    }
    template <class T> CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T*)
    {
        BuenoOrovio::cell_step(y, dt, istim, p);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T*)
    {
        return BuenoOrovio::jion(y[0], y, p);
    }
    struct None {};
    template <class T> CG_HD static None vt(T, const T*) { return None{}; }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T*, bool solve_v, T istim,
                             None)
    {
        BuenoOrovio::deriv(y, d, p, solve_v, istim);
    }
};

template <> struct Ops<AlievPanfilov> {
    template <class T>
    CG_HD static T tissue_step(T u, T* y, T dt, const T* p, int, const T*)
    {
        y[0] = u;
        AlievPanfilov::gates(y, dt, p);
        return AlievPanfilov::jion(u, y, p);
    }
    template <class T> CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T*)
    {
        AlievPanfilov::cell_step(y, dt, istim, p);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T*)
    {
        return AlievPanfilov::jion(y[0], y, p);
    }
    struct None {};
    template <class T> CG_HD static None vt(T, const T*) { return None{}; }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T*, bool solve_v, T istim,
                             None)
    {
        AlievPanfilov::deriv(y, d, p, solve_v, istim);
    }
};

template <> struct Ops<TNNP> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int tissue,
                                    const T* gs)
    {
        const auto q = TNNP::vterms(y[0], p);
        cell_step_offsets<TNNP>(
            y, s, dt, [&](RlStep<T>& out) { TNNP::rates(y, out, p, tissue, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int tissue, const T* gs)
    {
        y[0] = v;
        const auto q = TNNP::vterms(v, p);
        return tissue_step_offsets<TNNP>(
            y, s, dt,
            [&](RlStep<T>& out) { TNNP::rates(y, out, p, tissue, gs, false, T(0), q); },
            [&]() { return TNNP::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int tissue, const T* gs)
    {
        y[0] = v;
        const auto q = TNNP::vterms(v, p);
        TNNP::step(y, dt, p, tissue, gs, false, T(0), q);
        return TNNP::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int tissue, const T* gs)
    {
        const auto q = TNNP::vterms(y[0], p);
        TNNP::step(y, dt, p, tissue, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = TNNP::vterms(y[0], p);
        return TNNP::currents(y[0], y, p, q, gs).total();
    }
    template <class T> CG_HD static TNNP::VTerms<T> vt(T v, const T* p) { return TNNP::vterms(v, p); }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int tissue, const T* gs, bool solve_v,
                             T istim, const TNNP::VTerms<T>& q)
    {
        TNNP::deriv(y, d, p, tissue, gs, solve_v, istim, q);
    }
};

template <> struct Ops<ORd> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int tissue,
                                    const T* gs)
    {
        const auto q = ORd::vterms(y[0]);
        cell_step_offsets<ORd>(
            y, s, dt, [&](RlStep<T>& out) { ORd::rates(y, out, p, tissue, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int tissue, const T* gs)
    {
        y[0] = v;
        const auto q = ORd::vterms(v);
        return tissue_step_offsets<ORd>(
            y, s, dt,
            [&](RlStep<T>& out) { ORd::rates(y, out, p, tissue, gs, false, T(0), q); },
            [&]() { return ORd::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int tissue, const T* gs)
    {
        y[0] = v;
        const auto q = ORd::vterms(v);
        ORd::step(y, dt, p, tissue, gs, false, T(0), q);
        return ORd::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int tissue, const T* gs)
    {
        const auto q = ORd::vterms(y[0]);
        ORd::step(y, dt, p, tissue, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = ORd::vterms(y[0]);
        return ORd::currents(y[0], y, p, q, gs).total();
    }
    template <class T> CG_HD static ORd::VTerms<T> vt(T v, const T*) { return ORd::vterms(v); }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int tissue, const T* gs, bool solve_v,
                             T istim, const ORd::VTerms<T>& q)
    {
        ORd::deriv(y, d, p, tissue, gs, solve_v, istim, q);
    }
};

template <> struct Ops<FdaManufactured> {
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int, const T*)
    {
        y[0] = v;
        RlStep<T> out{y, dt};
        FdaManufactured::rates(y, out, p);
        return FdaManufactured::iion(y, p);
    }
    // dV/dt is 0 in this model (monodomainFDAManufactured_2014.H:68): a single cell only
    // moves u1, u2
    template <class T> CG_HD static void cell_step(T* y, T dt, T, const T* p, int, const T*)
    {
        RlStep<T> out{y, dt};
        FdaManufactured::rates(y, out, p);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T*)
    {
        return FdaManufactured::iion(y, p);
    }
    struct None {};
    template <class T> CG_HD static None vt(T, const T*) { return None{}; }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T*, bool, T, None)
    {
        Deriv<T> out{y, d};
        d[0] = T(0);
        FdaManufactured::rates(y, out, p);
    }
};

template <> struct Ops<Courtemanche> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Courtemanche::vterms(y[0], p);
        cell_step_offsets<Courtemanche>(
            y, s, dt, [&](RlStep<T>& out) { Courtemanche::rates(y, out, p, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Courtemanche::vterms(v, p);
        return tissue_step_offsets<Courtemanche>(
            y, s, dt, [&](RlStep<T>& out) { Courtemanche::rates(y, out, p, gs, false, T(0), q); },
            [&]() { return Courtemanche::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Courtemanche::vterms(v, p);
        Courtemanche::step(y, dt, p, gs, false, T(0), q);
        return Courtemanche::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Courtemanche::vterms(y[0], p);
        Courtemanche::step(y, dt, p, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = Courtemanche::vterms(y[0], p);
        return Courtemanche::currents(y[0], y, p, q, gs).total();
    }
    template <class T>
    CG_HD static Courtemanche::VTerms<T> vt(T v, const T* p)
    {
        return Courtemanche::vterms(v, p);
    }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T* gs, bool solve_v,
                             T istim, const Courtemanche::VTerms<T>& q)
    {
        Courtemanche::deriv(y, d, p, gs, solve_v, istim, q);
    }
};

template <> struct Ops<PerisYague> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = PerisYague::vterms(y[0], p);
        cell_step_offsets<PerisYague>(
            y, s, dt, [&](RlStep<T>& out) { PerisYague::rates(y, out, p, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = PerisYague::vterms(v, p);
        return tissue_step_offsets<PerisYague>(
            y, s, dt, [&](RlStep<T>& out) { PerisYague::rates(y, out, p, gs, false, T(0), q); },
            [&]() { return PerisYague::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = PerisYague::vterms(v, p);
        PerisYague::step(y, dt, p, gs, false, T(0), q);
        return PerisYague::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = PerisYague::vterms(y[0], p);
        PerisYague::step(y, dt, p, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = PerisYague::vterms(y[0], p);
        return PerisYague::currents(y[0], y, p, q, gs).total();
    }
    template <class T>
    CG_HD static PerisYague::VTerms<T> vt(T v, const T* p)
    {
        return PerisYague::vterms(v, p);
    }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T* gs, bool solve_v,
                             T istim, const PerisYague::VTerms<T>& q)
    {
        PerisYague::deriv(y, d, p, gs, solve_v, istim, q);
    }
};

template <> struct Ops<Stewart> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Stewart::vterms(y[0], p);
        cell_step_offsets<Stewart>(
            y, s, dt, [&](RlStep<T>& out) { Stewart::rates(y, out, p, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Stewart::vterms(v, p);
        return tissue_step_offsets<Stewart>(
            y, s, dt, [&](RlStep<T>& out) { Stewart::rates(y, out, p, gs, false, T(0), q); },
            [&]() { return Stewart::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Stewart::vterms(v, p);
        Stewart::step(y, dt, p, gs, false, T(0), q);
        return Stewart::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Stewart::vterms(y[0], p);
        Stewart::step(y, dt, p, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = Stewart::vterms(y[0], p);
        return Stewart::currents(y[0], y, p, q, gs).total();
    }
    template <class T>
    CG_HD static Stewart::VTerms<T> vt(T v, const T* p)
    {
        return Stewart::vterms(v, p);
    }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T* gs, bool solve_v,
                             T istim, const Stewart::VTerms<T>& q)
    {
        Stewart::deriv(y, d, p, gs, solve_v, istim, q);
    }
};

template <> struct Ops<Grandi> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Grandi::vterms(y[0], p);
        cell_step_offsets<Grandi>(
            y, s, dt, [&](RlStep<T>& out) { Grandi::rates(y, out, p, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Grandi::vterms(v, p);
        return tissue_step_offsets<Grandi>(
            y, s, dt, [&](RlStep<T>& out) { Grandi::rates(y, out, p, gs, false, T(0), q); },
            [&]() { return Grandi::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Grandi::vterms(v, p);
        Grandi::step(y, dt, p, gs, false, T(0), q);
        return Grandi::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Grandi::vterms(y[0], p);
        Grandi::step(y, dt, p, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = Grandi::vterms(y[0], p);
        return Grandi::currents(y[0], y, p, q, gs).total();
    }
    template <class T>
    CG_HD static Grandi::VTerms<T> vt(T v, const T* p)
    {
        return Grandi::vterms(v, p);
    }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T* gs, bool solve_v,
                             T istim, const Grandi::VTerms<T>& q)
    {
        Grandi::deriv(y, d, p, gs, solve_v, istim, q);
    }
};

template <> struct Ops<ToRORd> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int tissue,
                                    const T* gs)
    {
        const auto q = ToRORd::vterms(y[0], p);
        cell_step_offsets<ToRORd>(y, s, dt, [&](RlStep<T>& out) {
            ToRORd::rates(y, out, p, tissue, gs, true, istim, q);
        });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int tissue, const T* gs)
    {
        y[0] = v;
        const auto q = ToRORd::vterms(v, p);
        return tissue_step_offsets<ToRORd>(
            y, s, dt,
            [&](RlStep<T>& out) { ToRORd::rates(y, out, p, tissue, gs, false, T(0), q); },
            [&]() { return ToRORd::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int tissue, const T* gs)
    {
        y[0] = v;
        const auto q = ToRORd::vterms(v, p);
        ToRORd::step(y, dt, p, tissue, gs, false, T(0), q);
        return ToRORd::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int tissue, const T* gs)
    {
        const auto q = ToRORd::vterms(y[0], p);
        ToRORd::step(y, dt, p, tissue, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = ToRORd::vterms(y[0], p);
        return ToRORd::currents(y[0], y, p, q, gs).total();
    }
    template <class T> CG_HD static ToRORd::VTerms<T> vt(T v, const T* p)
    {
        return ToRORd::vterms(v, p);
    }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int tissue, const T* gs, bool solve_v,
                             T istim, const ToRORd::VTerms<T>& q)
    {
        ToRORd::deriv(y, d, p, tissue, gs, solve_v, istim, q);
    }
};

template <> struct Ops<Gaur> {
    template <class T>
    CG_HD static void cell_step_off(T* y, T* s, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Gaur::vterms(y[0], p);
        cell_step_offsets<Gaur>(
            y, s, dt, [&](RlStep<T>& out) { Gaur::rates(y, out, p, gs, true, istim, q); });
    }
    template <class T>
    CG_HD static T tissue_step_off(T v, T* y, T* s, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Gaur::vterms(v, p);
        return tissue_step_offsets<Gaur>(
            y, s, dt, [&](RlStep<T>& out) { Gaur::rates(y, out, p, gs, false, T(0), q); },
            [&]() { return Gaur::currents(v, y, p, q, gs).total(); });
    }
    template <class T>
    CG_HD static T tissue_step(T v, T* y, T dt, const T* p, int, const T* gs)
    {
        y[0] = v;
        const auto q = Gaur::vterms(v, p);
        Gaur::step(y, dt, p, gs, false, T(0), q);
        return Gaur::currents(v, y, p, q, gs).total();
    }
    template <class T>
    CG_HD static void cell_step(T* y, T dt, T istim, const T* p, int, const T* gs)
    {
        const auto q = Gaur::vterms(y[0], p);
        Gaur::step(y, dt, p, gs, true, istim, q);
    }
    template <class T> CG_HD static T iion(const T* y, const T* p, int, const T* gs)
    {
        const auto q = Gaur::vterms(y[0], p);
        return Gaur::currents(y[0], y, p, q, gs).total();
    }
    template <class T> CG_HD static Gaur::VTerms<T> vt(T v, const T* p)
    {
        return Gaur::vterms(v, p);
    }
    template <class T>
    CG_HD static void deriv(const T* y, T* d, const T* p, int, const T* gs, bool solve_v,
                             T istim, const Gaur::VTerms<T>& q)
    {
        Gaur::deriv(y, d, p, gs, solve_v, istim, q);
    }
};

// ALGEBRAIC variables of a model in the reference's order, as its computeVariables leaves
// them at a step's end; NA = 0 where the model's algebraic list is not evaluated on the
// device.  BuenoOrovio: BuenoOrovio_2008.H:30-43 (names), :166-225 (values);
// AlievPanfilov: AlievPanfilov_1996.H:28-32, :94-120.  istim: the single-cell stimulus at
// that time (0 in tissue, where the stimulus comes from the PDE).
template <class M> struct Algebraic {
    static constexpr int NA = 0;
    template <class T> CG_HD static void eval(const T*, const T*, T, T*) {}
};
template <> struct Algebraic<BuenoOrovio> {
    static constexpr int NA = 12;
    template <class T> CG_HD static void eval(const T* y, const T* p, T istim, T* a)
    {
        using B = BuenoOrovio;
        const T u = y[0];
        const bool Hv = pos(u - p[B::thv]), Hw = pos(u - p[B::thw]), Ho = pos(u - p[B::tho]),
                   Hvm = pos(u - p[B::thvm]);
        a[5] = p[B::tso1] + (p[B::tso2] - p[B::tso1]) * (T(1) + ctanh(p[B::kso] * (u - p[B::uso]))) / T(2);
        a[6] = Ho ? p[B::to2] : p[B::to1];
        a[0] = Hv ? -y[1] * (u - p[B::thv]) * (p[B::uu] - u) / p[B::tfi] : T(0);
        a[1] = Hw ? T(1) / a[5] : (u - p[B::uo]) / a[6];
        a[2] = Hw ? -y[2] * y[3] / p[B::tsi] : T(0);
        a[3] = a[0] + a[1] + a[2];
        a[4] = istim;
        a[7] = Hvm ? p[B::tv2m] : p[B::tv1m];
        a[8] = p[B::tw1m] + (p[B::tw2m] - p[B::tw1m]) * (T(1) + ctanh(p[B::kwm] * (u - p[B::uwm]))) / T(2);
        a[9] = Hw ? p[B::ts2] : p[B::ts1];
        a[10] = pos(p[B::thvm] - u) ? T(1) : T(0);
        a[11] = Ho ? p[B::winfstar] : T(1) - u / p[B::twinf];
    }
};
template <> struct Algebraic<AlievPanfilov> {
    static constexpr int NA = 3;
    template <class T> CG_HD static void eval(const T* y, const T* p, T istim, T* a)
    {
        using A = AlievPanfilov;
        const T u = y[0];
        a[0] = p[A::eps0] + p[A::mu1] * y[1] / (p[A::mu2] + u);
        a[1] = istim;
        a[2] = A::jion(u, y, p);
    }
};

} // namespace cg
