"""Oracle (test infrastructure, numpy): the monodomain diffusion on an arbitrary polyhedral
mesh as OpenFOAM discretises it, face by face, for checking the device's general-mesh
path.  It follows monodomainSolver.C:103-150 (fvc::laplacian / fvm::laplacian of a
volTensorField conductivity) with the Gauss linear corrected scheme of OpenFOAM's
finiteVolume library (gaussLaplacianScheme for a tensor diffusivity, correctedSnGrad,
gaussGrad with zeroGradient boundaries, surfaceInterpolation weights, nonOrthDeltaCoeffs
and nonOrthCorrectionVectors), restated from their documented formulas:

  w      = |Sf.(C_N - Cf)| / (|Sf.(Cf - C_P)| + |Sf.(C_N - Cf)|)
  sigma_f = w sigma_P + (1 - w) sigma_N          (boundary: sigma_P)
  SfGamma = Sf . sigma_f;  SfGammaSn = SfGamma . n;  SfGammaCorr = SfGamma - SfGammaSn n
  snGrad  = D (V_N - V_P) + k . interp(grad V),   D = 1/max(n.d, 0.05|d|),  k = n - D d
  flux    = SfGammaSn snGrad + SfGammaCorr . interp(grad V)
  wall    : flux = SfGammaCorr . (grad V_P - n (n . grad V_P))

The device precombines the face coefficients and evaluates each face from both of its
cells; this restatement keeps OpenFOAM's terms apart and scatters each face once.
"""
import numpy as np


def weights(m):
    Sf, Cf, C = m["Sf"], m["Cf"], m["C"]
    o, nb = m["owner"], m["neighbour"]
    sown = np.abs(np.einsum("ij,ij->i", Sf, Cf - C[o]))
    snei = np.abs(np.einsum("ij,ij->i", Sf, C[nb] - Cf))
    return snei / (sown + snei)


def grad(m, V):
    """Gauss linear gradient with zeroGradient walls, [n, 3]"""
    o, nb = m["owner"], m["neighbour"]
    w = weights(m)
    vf = w * V[o] + (1.0 - w) * V[nb]
    g = np.zeros((m["n"], 3))
    np.add.at(g, o, m["Sf"] * vf[:, None])
    np.add.at(g, nb, -m["Sf"] * vf[:, None])
    np.add.at(g, m["b_owner"], m["b_Sf"] * V[m["b_owner"]][:, None])
    return g / m["V"][:, None]


def ls_grad(m, V):
    """OpenFOAM's leastSquares gradient (leastSquaresVectors, leastSquaresGrad) with
    zeroGradient walls, [n, 3]: dd_c = sum over the cell's faces of d d^T / |d|^2 with
    d = C_N - C_P on internal faces and the wall-normal part of Cf - C_P on walls;
    grad_P += |d|^-2 dd_P^-1 d (V_N - V_P), grad_N += |d|^-2 dd_N^-1 d (V_N - V_P); a
    zeroGradient wall adds dd only (its face value is V_P)."""
    o, nb, C = m["owner"], m["neighbour"], m["C"]
    n = m["n"]
    d = C[nb] - C[o]
    w = 1.0 / np.einsum("ij,ij->i", d, d)
    dd = np.zeros((n, 3, 3))
    wdd = w[:, None, None] * d[:, :, None] * d[:, None, :]
    np.add.at(dd, o, wdd)
    np.add.at(dd, nb, wdd)
    bo = m["b_owner"]
    bn = m["b_Sf"] / np.linalg.norm(m["b_Sf"], axis=1)[:, None]
    bd = bn * np.einsum("ij,ij->i", bn, m["b_Cf"] - C[bo])[:, None]
    np.add.at(dd, bo, bd[:, :, None] * bd[:, None, :] /
              np.einsum("ij,ij->i", bd, bd)[:, None, None])
    inv = np.linalg.inv(dd)
    dV = V[nb] - V[o]
    g = np.zeros((n, 3))
    np.add.at(g, o, w[:, None] * np.einsum("fij,fj->fi", inv[o], d) * dV[:, None])
    np.add.at(g, nb, w[:, None] * np.einsum("fij,fj->fi", inv[nb], d) * dV[:, None])
    return g


def _sigma_cells(m, sigma9):
    s = np.asarray(sigma9, dtype=np.float64)
    return np.tile(s.reshape(1, 9), (m["n"], 1)) if s.size == 9 else s.reshape(m["n"], 9)


def face_fluxes(m, sigma9, V, orth=True):
    """(internal face fluxes [F], boundary face fluxes [B]) of div(sigma grad V)"""
    o, nb = m["owner"], m["neighbour"]
    S = _sigma_cells(m, sigma9).reshape(-1, 3, 3)
    w = weights(m)
    Sf = m["Sf"]
    mag = np.linalg.norm(Sf, axis=1)
    n = Sf / mag[:, None]
    d = m["C"][nb] - m["C"][o]
    D = 1.0 / np.maximum(np.einsum("ij,ij->i", n, d), 0.05 * np.linalg.norm(d, axis=1))
    k = n - D[:, None] * d
    sig_f = w[:, None, None] * S[o] + (1.0 - w)[:, None, None] * S[nb]
    SfGamma = np.einsum("fi,fij->fj", Sf, sig_f)
    SfGammaSn = np.einsum("fi,fi->f", SfGamma, n)
    SfGammaCorr = SfGamma - SfGammaSn[:, None] * n
    g = grad(m, V)
    gf = w[:, None] * g[o] + (1.0 - w)[:, None] * g[nb]
    sn = np.einsum("fi,fi->f", k, gf)
    if orth:
        sn = sn + D * (V[nb] - V[o])
    flux = SfGammaSn * sn + np.einsum("fi,fi->f", SfGammaCorr, gf)
    bo = m["b_owner"]
    bS = m["b_Sf"]
    bn = bS / np.linalg.norm(bS, axis=1)[:, None]
    bG = np.einsum("fi,fij->fj", bS, S[bo])
    bGC = bG - np.einsum("fi,fi->f", bG, bn)[:, None] * bn
    gb = g[bo] - np.einsum("fi,fi->f", g[bo], bn)[:, None] * bn
    bflux = np.einsum("fi,fi->f", bGC, gb)
    return flux, bflux


def divergence(m, flux, bflux):
    """sum of outward face fluxes per cell (volume-integrated)"""
    out = np.zeros(m["n"])
    np.add.at(out, m["owner"], flux)
    np.add.at(out, m["neighbour"], -flux)
    np.add.at(out, m["b_owner"], bflux)
    return out


def laplacian(m, sigma9, V):
    """fvc::laplacian(sigma, V) per unit volume"""
    f, b = face_fluxes(m, sigma9, V)
    return divergence(m, f, b) / m["V"]


def matrix(m, sigma9, chiCm, c, dt):
    """(diag [n], upper [F]) of chiCm fvm::ddt - fvm::laplacian, volume-integrated"""
    o, nb = m["owner"], m["neighbour"]
    S = _sigma_cells(m, sigma9).reshape(-1, 3, 3)
    w = weights(m)
    Sf = m["Sf"]
    n = Sf / np.linalg.norm(Sf, axis=1)[:, None]
    d = m["C"][nb] - m["C"][o]
    D = 1.0 / np.maximum(np.einsum("ij,ij->i", n, d), 0.05 * np.linalg.norm(d, axis=1))
    sig_f = w[:, None, None] * S[o] + (1.0 - w)[:, None, None] * S[nb]
    SfGammaSn = np.einsum("fi,fij,fj->f", Sf, sig_f, n)
    a = SfGammaSn * D
    diag = chiCm * c * m["V"] / dt
    np.add.at(diag, o, a)
    np.add.at(diag, nb, a)
    return diag, -a


def box_source(m, boxes):
    """summed intensities of the boxes (lo, hi, intensity) holding each cell centre"""
    src = np.zeros(m["n"])
    for lo, hi, inten in boxes:
        inside = np.all((m["C"] >= np.asarray(lo)) & (m["C"] <= np.asarray(hi)), axis=1)
        src[inside] += inten
    return src


def run_tissue(orc, m, sigma9, model, tissue, dt, chi, cm, n_steps, implicit=False,
               backward=False, boxes=(), V0=None, states=None, tol=1e-13, rel_tol=0.0,
               n_outer=1, residual_control=0.0, integrator=0, abs_tol=1e-6, ode_rel=1e-4):
    """The tissue step on a general mesh (myocardiumDomain.C:487-529): box stimulus at t0,
    the oracle's ionic step (orc.ionic_tissue_step), then the explicit update or the
    fvMatrix solved by the oracle's lduMatrix PCG in OpenFOAM's norm (orc.ldu_pcg_of)
    inside the PIMPLE corrector loop; activation times.  boxes: (lo, hi, t_start,
    duration, intensity).  Returns dict(V, tact, states, solves) with solves per step."""
    n = m["n"]
    chiCm = chi * cm
    c, y0 = orc.model_init(model, tissue)
    st = np.ascontiguousarray(np.tile(y0[:, None], (1, n)) if states is None else states,
                              dtype=np.float64).copy()
    V = np.full(n, -0.084) if V0 is None else np.array(V0, dtype=np.float64)
    Vold = V.copy()
    flag = np.ones(n, dtype=bool)
    tact = np.zeros(n)
    step = np.full(n, dt)
    t = 0.0
    mats = {}
    solves = []
    for s in range(n_steps):
        tn = t + dt
        t0 = tn - dt
        active = [(lo, hi, inten) for lo, hi, ts, du, inten in boxes if ts <= t0 <= ts + du]
        src = box_source(m, active)
        Im = orc.ionic_tissue_step(model, tissue, c, st, t0, dt, V, integrator=integrator,
                                   step=step, abs_tol=abs_tol, rel_tol=ode_rel)
        bdf = backward and s >= 1
        cc, c0, c00 = (1.5, 2.0, 0.5) if bdf else (1.0, 1.0, 0.0)
        if not implicit:
            Vn = (c0 * V - c00 * Vold + dt / chiCm * (laplacian(m, sigma9, V) - chiCm * Im + src)) / cc
            k = 0
        else:
            if cc not in mats:
                mats[cc] = matrix(m, sigma9, chiCm, cc, dt)
            diag, upper = mats[cc]
            b0 = m["V"] * (chiCm / dt * (c0 * V - c00 * Vold) - chiCm * Im + src)
            Vn = V.copy()
            prev, k = 0.0, 0
            for outer in range(1, 10 ** 6):
                last = n_outer <= 1 or outer >= n_outer or (outer > 1 and prev < residual_control)
                f, bf = face_fluxes(m, sigma9, Vn, orth=False)
                b = b0 + divergence(m, f, bf)
                Vn, it, r0, r1 = orc.ldu_pcg_of(n, m["owner"], m["neighbour"], diag, upper, b,
                                                Vn, tol=tol, rel_tol=rel_tol, max_iter=1000,
                                                precond=0)
                prev = r0
                k += 1
                if last:
                    break
        hit = flag & (Vn > 1e-15)
        tact[hit] = t0 + (0.0 - V[hit]) / (Vn[hit] - V[hit]) * dt
        flag[hit] = False
        Vold, V = V, Vn
        t = tn
        solves.append(k)
    return dict(V=V, tact=tact, states=st, solves=solves)
