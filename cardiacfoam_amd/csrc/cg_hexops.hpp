// Stencil helpers on the uniform hex block shared by the explicit and implicit paths.
#pragma once
#include "cg_internal.hpp"

namespace cg {

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD owns a
// contiguous run of cells and the +-nx, +-nx*ny neighbour rows stay in its own L2.
// Identity unless the grid is a multiple of 8.
CG_DEV uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    return (nb & 7u) ? b : (b & 7u) * (nb >> 3) + (b >> 3);
}

// cell coordinates of linear index c (32-bit arithmetic: n < 2^31 is enforced on create)
struct IJK {
    int i, j, k;
};
CG_DEV IJK ijk_of(uint32_t c, const HexGeom& g)
{
    const uint32_t nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    const uint32_t row = c / nx;
    IJK r;
    r.i = (int)(c - row * nx);
    r.k = (int)(row / ny);
    r.j = (int)(row - (uint32_t)r.k * ny);
    return r;
}

// Neighbour indices of a cell, clamped to the cell itself on a wall.  Stencils read every
// neighbour unconditionally through these (so all loads of a row are in flight together)
// and give wall neighbours a zero weight.
struct Nbr {
    int64_t xp, xm, yp, ym, zp, zm;
};
CG_DEV Nbr nbr_of(int64_t c, const IJK& q, const HexGeom& g)
{
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    Nbr b;
    b.xp = q.i < g.nx - 1 ? c + 1 : c;
    b.xm = q.i > 0 ? c - 1 : c;
    b.yp = q.j < g.ny - 1 ? c + sy : c;
    b.ym = q.j > 0 ? c - sy : c;
    b.zp = q.k < g.nz - 1 ? c + sz : c;
    b.zm = q.k > 0 ? c - sz : c;
    return b;
}

// Orthogonal part of the FV Laplacian per unit volume, sum_f a_f (V_nb - V_c): uniform
// face coefficients (ax, ay, az) or per-face planes indexed by the lower cell.  Wall
// neighbours are clamped to the cell, so their difference vanishes (zeroGradient).
template <bool UNIFORM>
CG_DEV double hex_laplacian(const double* __restrict__ V, int64_t c, const IJK& q,
                            const HexGeom& g, const HexCoefs& a)
{
    const Nbr b = nbr_of(c, q, g);
    const double vc = V[c];
    if (UNIFORM)
        return a.ax * ((V[b.xp] - vc) + (V[b.xm] - vc)) + a.ay * ((V[b.yp] - vc) + (V[b.ym] - vc)) +
               a.az * ((V[b.zp] - vc) + (V[b.zm] - vc));
    return a.fx[c] * (V[b.xp] - vc) + a.fx[b.xm] * (V[b.xm] - vc) + a.fy[c] * (V[b.yp] - vc) +
           a.fy[b.ym] * (V[b.ym] - vc) + a.fz[c] * (V[b.zp] - vc) + a.fz[b.zm] * (V[b.zm] - vc);
}

// Explicit tensor correction of "Gauss linear corrected" with a tensor conductivity,
// div(SfGammaCorr . interp(grad V)) per unit volume, from the off-diagonal conductivity
// components od = (xy, xz, yz) per cell [3][n].  Gradients are Gauss-linear with
// zeroGradient walls: g_e(q) = (V(q+e) - V(q-e)) / (2 h_e) with a wall neighbour replaced
// by q itself.  On a wall face the cell tensor and cell gradient are used (the normal part,
// snGrad = 0, drops out of SfGammaCorr); with the neighbour index clamped to the cell the
// interior face formula 1/2(od_c + od_nb) 1/2(g_c + g_nb) gives exactly that, so every
// face takes the same branch-free path.
// mask: bit 0/1/2 set when the xy/xz/yz component is non-zero somewhere on the mesh;
// components known to be zero everywhere are skipped.
CG_DEV double hex_tensor_correction(const double* __restrict__ V, const double* __restrict__ od,
                                    int64_t n, int64_t c64, int i, int j, int k,
                                    const HexGeom& g, unsigned mask = 7u)
{
    const int c = (int)c64;
    const int idx[3] = {i, j, k};
    const int nn[3] = {g.nx, g.ny, g.nz};
    const int st[3] = {1, g.nx, g.nx * g.ny};
    const double rh2[3] = {0.5 / g.hx, 0.5 / g.hy, 0.5 / g.hz};
    // neighbour of q (same coordinates as c along e) in direction e, clamped
    auto up = [&](int q, int e) { return idx[e] < nn[e] - 1 ? q + st[e] : q; };
    auto dn = [&](int q, int e) { return idx[e] > 0 ? q - st[e] : q; };
    auto grad = [&](int q, int e) { return (V[up(q, e)] - V[dn(q, e)]) * rh2[e]; };
    // od plane of the pair (d, e), d != e: xy -> 0, xz -> 1, yz -> 2
    auto plane = [](int d, int e) { return d + e - 1; };
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int cp = up(c, d), cm = dn(c, d);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (e == d || !(mask & (1u << plane(d, e)))) continue;
            const double* odp = od + (int64_t)plane(d, e) * n;
            const double oc = odp[c], op = odp[cp], om = odp[cm];
            const double gc = grad(c, e), gp = grad(cp, e), gm = grad(cm, e);
            // (F(+d face) - F(-d face)) / h_d, F = 1/2(od_c + od_nb) 1/2(g_c + g_nb)
            const double q = (oc + op) * (gc + gp) - (oc + om) * (gc + gm);
            acc += 0.25 * q * (2.0 * rh2[d]);
        }
    }
    return acc;
}

} // namespace cg
