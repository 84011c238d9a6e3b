// monodomainFDAManufactured: the reference's manufactured-solution model
// (src/ionicModels/monodomainFDAManufactured/monodomainFDAManufactured_2014.H:40-80,
// monodomainFDAManufactured.C:100-147).  Four states (V, u1, u2, u3), model time in
// seconds, V = Vm, Im = Iion / Cm; dV/dt is 0 (the tissue integrates V).  The `tissue`
// flag carries the dimension word (1D/2D/3D -> 11/12/13), which selects Beta.  Starting
// from the fields of manufacturedFDAReference.H, the tissue step reproduces their closed
// form (chi = 3, Cm = 2, sigma = diag(0.111453302, 0.121585420, 0.030396355) on the unit
// cube).  Neither state has an (xinf, tau) form: Rush-Larsen integrates both by forward
// Euler.
#pragma once
#include "cg_common.hpp"

namespace cg {

struct FdaManufactured {
    static constexpr int NS = 4;
    static constexpr int NP = 3;
    static constexpr double VSCALE = 1.0, VOFF = 0.0, VMUL = 1.0, ISCALE = 1.0;
    static constexpr double TSCALE = 1.0;  // model time is PDE time
    enum { V, u1, u2, u3 };
    enum { Cm, Beta, Chi };

    static bool host_init(int dim, double* p, double* y)
    {
        if (dim < 11 || dim > 13) return false;
        p[Cm] = 2.0;
        p[Beta] = dim == 11 ? -1.1 : dim == 12 ? -5.9 : -8.6;
        p[Chi] = 3.0;
        y[0] = y[1] = y[2] = y[3] = 0.0;  // the manufactured fields are set by the caller
        return true;
    }

    template <class T, class U> CG_HD static void rates(const T* y, U& out, const T*)
    {
        const T a = y[u1] + y[u3] - y[V];
        const T u22 = y[u2] * y[u2];
        out.rate(u1, a * a * u22 + T(0.5) * a * u22 * (y[V] - y[u3]));
        out.rate(u2, -a * (u22 * y[u2]));
        out.rate(u3, T(0));
    }

    // Iion / Cm, the current the monodomain solver takes
    template <class T> CG_HD static T iion(const T* y, const T* p)
    {
        const T a = y[u1] + y[u3] - y[V];
        const T Iion = -T(0.5) * p[Cm] * a * (y[u2] * y[u2]) * (y[V] - y[u3]) +
                       (p[Beta] * (y[V] - y[u3])) / p[Chi];
        return Iion / p[Cm];
    }
};

} // namespace cg
