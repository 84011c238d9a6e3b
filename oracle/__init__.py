"""ctypes binding of the CPU FP64 oracle -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
import this module, and only as the checker.  The product package
``cardiacfoam_amd`` never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

BO, TNNP, ORD, AP, FDA, CRN, PYS, STW, GRD, TOR, GAUR = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10
EPI, MCELL, ENDO, MYOCYTE = 1, 2, 3, 4
DIM1, DIM2, DIM3 = 11, 12, 13          # dimension flags of the manufactured model
MAX_STATES, MAX_CONSTS = 48, 160
NSTATES = {BO: 4, TNNP: 17, ORD: 41, AP: 2, FDA: 4, CRN: 21, PYS: 22, STW: 20, GRD: 41,
           TOR: 45, GAUR: 29}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Stim(C.Structure):
    _fields_ = [("start", C.c_double), ("period_s1", C.c_double), ("duration", C.c_double),
                ("amplitude", C.c_double), ("nstim1", C.c_int), ("period_s2", C.c_double),
                ("nstim2", C.c_int)]


class Mesh(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("hx", C.c_double),
                ("hy", C.c_double), ("hz", C.c_double), ("n_cells", C.c_long),
                ("n_faces", C.c_long), ("owner", C.POINTER(C.c_int32)),
                ("neighbour", C.POINTER(C.c_int32)), ("face_dir", C.POINTER(C.c_int8))]


class Boxes(C.Structure):
    _fields_ = [("n_boxes", C.c_int), ("box_min", C.POINTER(C.c_double)),
                ("box_max", C.POINTER(C.c_double)), ("t_start", C.POINTER(C.c_double)),
                ("duration", C.POINTER(C.c_double)), ("intensity", C.POINTER(C.c_double)),
                ("centres", C.POINTER(C.c_double))]


class TissueCfg(C.Structure):
    _fields_ = [("model", C.c_int), ("tissue", C.c_int), ("integrator", C.c_int),
                ("n_sub", C.c_int), ("abs_tol", C.c_double), ("rel_tol", C.c_double),
                ("ddt", C.c_int), ("implicit", C.c_int), ("pcg_tol", C.c_double),
                ("pcg_max_iter", C.c_int), ("chi", C.c_double), ("cm", C.c_double),
                ("dt", C.c_double), ("nthreads", C.c_int), ("norm", C.c_int),
                ("pcg_rel_tol", C.c_double), ("precond", C.c_int), ("n_outer", C.c_int),
                ("residual_control", C.c_double)]


class Tissue(C.Structure):
    _fields_ = [("mesh", Mesh), ("cfg", TissueCfg), ("nstates", C.c_int),
                ("consts", C.c_double * MAX_CONSTS), ("states", C.POINTER(C.c_double)),
                ("step", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double)),
                ("V", C.POINTER(C.c_double)), ("V0", C.POINTER(C.c_double)),
                ("V00", C.POINTER(C.c_double)), ("Iion", C.POINTER(C.c_double)),
                ("src", C.POINTER(C.c_double)), ("tact", C.POINTER(C.c_double)),
                ("flag", C.POINTER(C.c_uint8)), ("work", C.POINTER(C.c_double)),
                ("t", C.c_double), ("step_index", C.c_long), ("pcg_iters_total", C.c_long),
                ("outer_total", C.c_long), ("last_init_res", C.c_double)]


class Beat(C.Structure):
    _fields_ = [("t_up", C.c_double), ("v_base", C.c_double), ("v_peak", C.c_double),
                ("v90", C.c_double), ("t_repol90", C.c_double), ("repol_found", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_model_nstates.argtypes = [C.c_int]
        L.orc_model_init.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_model_const_index.argtypes = [C.c_int, C.c_char_p]
        L.orc_model_eval.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, C.c_int,
                                     C.POINTER(Stim), _dp, _ip, _dp, _dp,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_stimulus.argtypes = [C.c_double, C.POINTER(Stim)]
        L.orc_stimulus.restype = C.c_double
        L.orc_rl_step.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                  C.POINTER(Stim)]
        L.orc_rkf45_solve.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.POINTER(Stim),
                                      C.c_double, C.c_double, C.POINTER(C.c_double), C.c_double,
                                      C.c_double, C.c_long]
        L.orc_rkf45_solve.restype = C.c_long
        L.orc_ode_solve.argtypes = [C.c_int] + L.orc_rkf45_solve.argtypes
        L.orc_model_algebraic.argtypes = [C.c_int, C.c_double, _dp, _dp, C.POINTER(Stim), _dp]
        L.orc_model_algebraic.restype = C.c_int
        L.orc_ode_solve.restype = C.c_long
        L.orc_ionic_tissue_step.argtypes = [C.c_int, C.c_int, _dp, C.c_long, _dp, C.c_long,
                                            C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_int,
                                            _dp, C.c_double, C.c_double, C.c_int]
        L.orc_single_cell_run.argtypes = [C.c_int, C.c_int, _dp, _dp, C.POINTER(Stim), C.c_double,
                                          C.c_long, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_single_cell_trace.argtypes = [C.c_int, C.c_int, _dp, _dp, C.POINTER(Stim),
                                            C.c_double, C.c_long, _dp, _dp]
        L.orc_single_cell_trace.restype = None
        L.orc_model_vm_mV.argtypes = [C.c_int, _dp]
        L.orc_model_vm_mV.restype = C.c_double
        L.orc_mesh_build.argtypes = [C.POINTER(Mesh), C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double]
        L.orc_mesh_free.argtypes = [C.POINTER(Mesh)]
        L.orc_laplacian.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, C.c_int]
        L.orc_correction_div.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, C.c_int]
        L.orc_assemble_implicit.argtypes = [C.POINTER(Mesh), _dp, C.c_double, C.c_double,
                                            C.c_double, _dp, _dp]
        L.orc_ldu_amul.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, _dp]
        L.orc_pcg_of.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_pcg.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                              C.POINTER(C.c_double)]
        L.orc_box_stimulus.argtypes = [C.POINTER(Mesh), C.POINTER(Boxes), C.c_double, _dp]
        L.orc_activation.argtypes = [C.c_long, _dp, _dp, _u8, _dp, C.c_double, C.c_double]
        L.orc_pseudo_ecg.argtypes = [C.POINTER(Mesh), _dp, _dp, C.c_int, _dp, _dp, _dp,
                                     C.c_int]
        L.orc_explicit_dt.argtypes = [C.POINTER(Mesh), _dp, C.c_double, C.c_double]
        L.orc_explicit_dt.restype = C.c_double
        L.orc_tissue_create.argtypes = [C.POINTER(Tissue), C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.POINTER(TissueCfg), _dp]
        L.orc_tissue_free.argtypes = [C.POINTER(Tissue)]
        L.orc_tissue_step.argtypes = [C.POINTER(Tissue), C.POINTER(Boxes)]
        L.orc_detect_beats.argtypes = [_dp, _dp, C.c_long, C.POINTER(Beat), C.c_int, C.c_double,
                                       C.c_double, C.c_double]
        L.orc_bo_tangent.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                                     C.c_double, C.c_double, C.c_double, C.POINTER(Boxes), _dp,
                                     _dp, C.c_int, _dp, _dp, C.c_long, C.c_int, _lp, _dp,
                                     C.c_int, C.POINTER(C.c_double), _dp, _dp]
        _lib = L
    return _lib


# ----------------------------------------------------------------- helpers
def make_stim(start=0.0, period_s1=0.0, duration=0.0, amplitude=0.0, nstim1=None,
              period_s2=0.0, nstim2=0):
    """stimulusIO.C:142-194 defaults: nstim1 = 1 if a positive S1 period is given."""
    if nstim1 is None:
        nstim1 = 1 if period_s1 > 1e-15 else 0
    return Stim(start, period_s1, duration, amplitude, nstim1, period_s2, nstim2)


def model_init(model, tissue):
    c = np.zeros(MAX_CONSTS)
    y = np.zeros(MAX_STATES)
    if lib().orc_model_init(model, tissue, c, y) != 0:
        raise ValueError("unsupported tissue %r for model %r" % (tissue, model))
    return c, y[:NSTATES[model]].copy()


def const_index(model, name):
    return lib().orc_model_const_index(model, name.encode())


def model_eval(model, tissue, t, c, y, solve_vm=False, stim=None):
    n = NSTATES[model]
    yy = np.zeros(MAX_STATES)
    yy[:n] = y
    r = np.zeros(MAX_STATES)
    g = np.zeros(MAX_STATES, dtype=np.int32)
    xi = np.zeros(MAX_STATES)
    ta = np.zeros(MAX_STATES)
    iion = C.c_double()
    ist = C.c_double()
    lib().orc_model_eval(model, tissue, t, np.ascontiguousarray(c, dtype=np.float64), yy,
                         int(solve_vm), C.byref(stim) if stim else None, r, g, xi, ta,
                         C.byref(iion), C.byref(ist))
    return dict(rates=r[:n], gate=g[:n].astype(bool), xinf=xi[:n], tau=ta[:n],
                iion=iion.value, istim=ist.value)


EULER = 2   # integrator / method value of OpenFOAM's adaptive Euler (1: RKF45, 0: RL)


def model_algebraic(model, tissue, t, y, stim=None, consts=None):
    """ALGEBRAIC variables (reference order) of BuenoOrovio / AlievPanfilov at model time t;
    None for a model without a restated list"""
    c, _ = model_init(model, tissue)
    if consts is not None:
        c = np.array(consts, dtype=np.float64)
    yy = np.zeros(MAX_STATES)
    yy[:len(y)] = y
    a = np.zeros(64)
    st = stim if stim is not None else make_stim()
    n = lib().orc_model_algebraic(model, t, c, yy, C.byref(st), a)
    return None if n < 0 else a[:n]


def ode_solve(method, model, tissue, y, x0, x1, dx_try, abs_tol, rel_tol, solve_vm=True,
              stim=None, consts=None, max_steps=1000000000):
    """ODESolver::solve(x0, x1, y, dxTry) on one cell with `method` (1 RKF45, 2 Euler):
    (y at x1, dxTry out, accepted steps)"""
    c, _ = model_init(model, tissue)
    if consts is not None:
        c = np.array(consts, dtype=np.float64)
    yy = np.zeros(MAX_STATES)
    yy[:len(y)] = y
    st = stim if stim is not None else make_stim()
    dx = C.c_double(dx_try)
    n = lib().orc_ode_solve(method, model, tissue, c, yy, int(solve_vm), C.byref(st), x0, x1,
                            C.byref(dx), abs_tol, rel_tol, max_steps)
    return yy[:NSTATES[model]], dx.value, n


def single_cell_run(model, tissue, stim, dt_s, n_steps, integrator=1, abs_tol=1e-6,
                    rel_tol=1e-4, consts=None, y0=None):
    c, y = model_init(model, tissue)
    if consts is not None:
        c = np.array(consts, dtype=np.float64)
    if y0 is not None:
        y = np.array(y0, dtype=np.float64)
    yy = np.zeros(MAX_STATES)
    yy[:len(y)] = y
    t = np.zeros(n_steps)
    v = np.zeros(n_steps)
    rc = lib().orc_single_cell_run(model, tissue, c, yy, C.byref(stim), dt_s, n_steps,
                                   integrator, abs_tol, rel_tol, t, v)
    if rc != 0:
        raise RuntimeError("single-cell integration failed")
    return t, v, yy[:NSTATES[model]]


def single_cell_trace(model, tissue, stim, dt_model, n_steps, consts=None, y0=None):
    """Rush-Larsen single cell with the step start at k*dt_model (model units, ms for
    TNNP/ORd); returns (t, Vm mV) after each step and the final state."""
    c, y = model_init(model, tissue)
    if consts is not None:
        c = np.array(consts, dtype=np.float64)
    if y0 is not None:
        y = np.array(y0, dtype=np.float64)
    yy = np.zeros(MAX_STATES)
    yy[:len(y)] = y
    t = np.zeros(n_steps)
    v = np.zeros(n_steps)
    lib().orc_single_cell_trace(model, tissue, c, yy, C.byref(stim), dt_model, n_steps, t, v)
    return t, v, yy[:NSTATES[model]]


def ionic_tissue_step(model, tissue, consts, states, t0, dt, Vm, integrator=0, n_sub=1,
                      step=None, abs_tol=1e-6, rel_tol=1e-4, c_stride=0, nthreads=8):
    """states: [nstates, n] float64 (modified in place); returns Im (V/s)."""
    n = states.shape[1]
    Im = np.zeros(n)
    if step is None:
        step = np.full(n, dt)
    lib().orc_ionic_tissue_step(model, tissue, np.ascontiguousarray(consts, dtype=np.float64),
                                c_stride, states, n, t0, dt,
                                np.ascontiguousarray(Vm, dtype=np.float64), Im, integrator,
                                n_sub, step, abs_tol, rel_tol, nthreads)
    return Im


class OracleMesh:
    def __init__(self, nx, ny, nz, hx, hy, hz):
        self.m = Mesh()
        if lib().orc_mesh_build(C.byref(self.m), nx, ny, nz, hx, hy, hz) != 0:
            raise MemoryError
        self.n = nx * ny * nz
        nf = self.m.n_faces
        self.owner = np.ctypeslib.as_array(self.m.owner, shape=(nf,)).copy()
        self.neighbour = np.ctypeslib.as_array(self.m.neighbour, shape=(nf,)).copy()
        self.face_dir = np.ctypeslib.as_array(
            C.cast(self.m.face_dir, C.POINTER(C.c_int8)), shape=(nf,)).copy()

    def __del__(self):
        try:
            lib().orc_mesh_free(C.byref(self.m))
        except Exception:
            pass

    def laplacian(self, sigma9, V, nthreads=8):
        out = np.zeros(self.n)
        lib().orc_laplacian(C.byref(self.m), np.ascontiguousarray(sigma9, dtype=np.float64),
                            np.ascontiguousarray(V, dtype=np.float64), out, nthreads)
        return out

    def correction_div(self, sigma9, V, nthreads=8):
        out = np.zeros(self.n)
        lib().orc_correction_div(C.byref(self.m), np.ascontiguousarray(sigma9, dtype=np.float64),
                                 np.ascontiguousarray(V, dtype=np.float64), out, nthreads)
        return out

    def assemble_implicit(self, sigma9, chiCm, dt, coefft=1.0):
        diag = np.zeros(self.n)
        upper = np.zeros(self.m.n_faces)
        lib().orc_assemble_implicit(C.byref(self.m), np.ascontiguousarray(sigma9, dtype=np.float64),
                                    chiCm, dt, coefft, diag, upper)
        return diag, upper

    def amul(self, diag, upper, x):
        y = np.zeros(self.n)
        lib().orc_ldu_amul(C.byref(self.m), diag, upper, np.ascontiguousarray(x, dtype=np.float64), y)
        return y

    def pcg(self, diag, upper, b, x0, tol=1e-10, max_iter=10000):
        x = np.array(x0, dtype=np.float64)
        res = C.c_double()
        it = lib().orc_pcg(C.byref(self.m), diag, upper, np.ascontiguousarray(b, dtype=np.float64),
                           x, tol, max_iter, C.byref(res))
        return x, it, res.value

    def pcg_of(self, diag, upper, b, x0, tol=1e-11, rel_tol=0.0, max_iter=1000, precond=1):
        """OpenFOAM PCG (normalised residual; precond 0 diagonal, 1 DIC): returns
        (x, iterations, initial residual, final residual)"""
        x = np.array(x0, dtype=np.float64)
        r0, r1 = C.c_double(), C.c_double()
        it = lib().orc_pcg_of(C.byref(self.m), np.ascontiguousarray(diag, dtype=np.float64),
                              np.ascontiguousarray(upper, dtype=np.float64),
                              np.ascontiguousarray(b, dtype=np.float64), x, tol, rel_tol,
                              max_iter, precond, C.byref(r0), C.byref(r1))
        return x, it, r0.value, r1.value

    def pseudo_ecg(self, sigma9, V, positions, nthreads=8):
        """pseudo-ECG at the electrodes (orc_pseudo_ecg): (values, sums of |terms|)"""
        n = self.m.n_cells
        s = np.asarray(sigma9, dtype=np.float64)
        s = np.ascontiguousarray(np.tile(s.ravel(), n) if s.size == 9 else s.ravel())
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pos))
        ab = np.zeros(len(pos))
        lib().orc_pseudo_ecg(C.byref(self.m), s, np.ascontiguousarray(V, dtype=np.float64),
                             len(pos), pos, out, ab, nthreads)
        return out, ab

    def explicit_dt(self, sigma9, chiCm, maxCo):
        return lib().orc_explicit_dt(C.byref(self.m), np.ascontiguousarray(sigma9, dtype=np.float64),
                                     chiCm, maxCo)


def ldu_pcg_of(n_cells, lower_addr, upper_addr, diag, upper, b, x0, tol=1e-11, rel_tol=0.0,
               max_iter=1000, precond=0):
    """OpenFOAM PCG (orc_pcg_of) on an arbitrary lduAddressing: returns (x, iterations,
    initial residual, final residual)"""
    lo = np.ascontiguousarray(lower_addr, dtype=np.int32)
    up = np.ascontiguousarray(upper_addr, dtype=np.int32)
    m = Mesh()
    m.n_cells = int(n_cells)
    m.n_faces = int(lo.size)
    m.owner = lo.ctypes.data_as(C.POINTER(C.c_int32))
    m.neighbour = up.ctypes.data_as(C.POINTER(C.c_int32))
    x = np.array(x0, dtype=np.float64)
    r0, r1 = C.c_double(), C.c_double()
    it = lib().orc_pcg_of(C.byref(m), np.ascontiguousarray(diag, dtype=np.float64),
                          np.ascontiguousarray(upper, dtype=np.float64),
                          np.ascontiguousarray(b, dtype=np.float64), x, tol, rel_tol, max_iter,
                          precond, C.byref(r0), C.byref(r1))
    return x, it, r0.value, r1.value


def make_boxes(box_min, box_max, t_start, duration, intensity, centres=None):
    """centres: optional [n, 3] cell centres (m) for the inclusion test; None uses the
    block's (i + 1/2) h lattice"""
    arrs = [np.ascontiguousarray(np.ravel(a), dtype=np.float64)
            for a in (box_min, box_max, t_start, duration, intensity)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    if centres is not None:
        arrs.append(np.ascontiguousarray(np.reshape(centres, -1), dtype=np.float64))
        ptrs.append(arrs[-1].ctypes.data_as(C.POINTER(C.c_double)))
    else:
        ptrs.append(C.POINTER(C.c_double)())
    b = Boxes(len(arrs[2]), *ptrs)
    b._keep = arrs
    return b


def bo_tangent(shape, h, sig0, tissue, dt, chi, cm, boxes, kappa, theta, dkappa, dtheta,
               n_steps, probes, target, nthreads=8):
    """Loss and its directional derivatives along each row of (dkappa [K, n], dtheta
    [K, 2]) from the tangent-linear BuenoOrovio tissue run (orc_bo_tangent), eight
    directions per pass.  Returns (loss, dloss [K], Vm after the last step)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    dk = np.ascontiguousarray(np.reshape(dkappa, (-1, n)), dtype=np.float64)
    dth = np.ascontiguousarray(np.reshape(dtheta, (-1, 2)), dtype=np.float64)
    assert dk.shape[0] == dth.shape[0]
    args = (np.ascontiguousarray(kappa, dtype=np.float64),
            np.ascontiguousarray(theta, dtype=np.float64))
    pr = np.ascontiguousarray(probes, dtype=np.int64)
    tg = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
    assert tg.size == n_steps * pr.size
    s0 = np.ascontiguousarray(sig0, dtype=np.float64)
    out, loss, v = [], None, np.zeros(n)
    for a in range(0, max(dk.shape[0], 1), 8):
        kk = dk[a:a + 8]
        L, d = C.c_double(), np.zeros(8)
        rc = lib().orc_bo_tangent(nx, ny, nz, h, s0, tissue, dt, chi, cm, C.byref(boxes),
                                  args[0], args[1], kk.shape[0],
                                  np.ascontiguousarray(kk if kk.size else np.zeros((1, n))),
                                  np.ascontiguousarray(dth[a:a + 8] if kk.size else np.zeros((1, 2))),
                                  int(n_steps), pr.size, pr, tg, nthreads, C.byref(L), d, v)
        if rc != 0:
            raise ValueError("orc_bo_tangent failed (%d)" % rc)
        loss = L.value
        out.append(d[:kk.shape[0]])
    return loss, np.concatenate(out) if out else np.zeros(0), v


class OracleTissue:
    """FP64 tissue run of the reference's step order (myocardiumDomain.C:487-529)."""

    def __init__(self, nx, ny, nz, h, sigma9, model, tissue, dt, chi, cm, integrator=0,
                 n_sub=1, ddt=0, implicit=0, pcg_tol=1e-10, pcg_max_iter=20000,
                 abs_tol=1e-6, rel_tol=1e-4, nthreads=8, norm=0, pcg_rel_tol=0.0, precond=0,
                 n_outer=1, residual_control=0.0):
        """norm 0: ||r||2 <= pcg_tol ||b||2 with Jacobi PCG; norm 1: OpenFOAM's normalised
        residual (pcg_tol, pcg_rel_tol, precond 0 diagonal / 1 DIC).  n_outer and
        residual_control: PIMPLE nOuterCorrectors and the outer initial-residual tolerance."""
        hx, hy, hz = (h, h, h) if np.isscalar(h) else h
        self.T = Tissue()
        cfg = TissueCfg(model, tissue, integrator, n_sub, abs_tol, rel_tol, ddt, implicit,
                        pcg_tol, pcg_max_iter, chi, cm, dt, nthreads, norm, pcg_rel_tol,
                        precond, n_outer, residual_control)
        sig = np.ascontiguousarray(sigma9, dtype=np.float64)
        if sig.size == 9:
            sig = np.tile(sig.reshape(9), nx * ny * nz)
        rc = lib().orc_tissue_create(C.byref(self.T), nx, ny, nz, hx, hy, hz, C.byref(cfg), sig)
        if rc != 0:
            raise ValueError("orc_tissue_create failed (%d)" % rc)
        self.n = nx * ny * nz
        self.nstates = NSTATES[model]

    def __del__(self):
        try:
            lib().orc_tissue_free(C.byref(self.T))
        except Exception:
            pass

    def step(self, boxes=None):
        lib().orc_tissue_step(C.byref(self.T), C.byref(boxes) if boxes is not None else None)

    def set_const(self, k, value):
        """model constant k (the model's constant order) for every cell"""
        self.T.consts[k] = float(value)

    def _arr(self, ptr, n):
        # copies: the C buffers are freed with the tissue object
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    @property
    def V(self):
        return self._arr(self.T.V, self.n)

    def set_V(self, V, old=True):
        """Vm (and, with old, its old levels) from an array of n values"""
        for ptr in ((self.T.V, self.T.V0, self.T.V00) if old else (self.T.V,)):
            np.ctypeslib.as_array(ptr, shape=(self.n,))[:] = V

    def set_states(self, S):
        """all states, [nstates, n]"""
        np.ctypeslib.as_array(self.T.states, shape=(self.nstates * self.n,))[:] = \
            np.asarray(S, dtype=np.float64).ravel()

    @property
    def solver_counts(self):
        """(PCG iterations, linear solves) over all steps, initial residual of the last solve"""
        return self.T.pcg_iters_total, self.T.outer_total, self.T.last_init_res

    @property
    def Iion(self):
        return self._arr(self.T.Iion, self.n)

    @property
    def tact(self):
        return self._arr(self.T.tact, self.n)

    @property
    def flag(self):
        return np.ctypeslib.as_array(self.T.flag, shape=(self.n,)).copy()

    @property
    def states(self):
        return self._arr(self.T.states, self.n * self.nstates).reshape(self.nstates, self.n)

    @property
    def t(self):
        return self.T.t


def classify_beats(beats):
    """classify_early_strokes (postProcessing_restCurves.py:203-263) on detect_beats'
    output: a beat is early when the next upstroke comes before its 90 % repolarisation
    or it never repolarised; early beats lose their APD and DI; a beat's DI is valid when
    the beat before it was not early and it repolarised itself.  Adds the keys early,
    valid_apd, valid_di in place; returns the indices of the early beats."""
    for b in beats:
        b["early"] = False
        b["valid_apd"] = bool(b["repol_found"])
        b["valid_di"] = False
    early = []
    for i in range(len(beats) - 1):
        cur, nxt = beats[i], beats[i + 1]
        if not cur["repol_found"] or nxt["t_up"] < cur["t_repol90"]:
            cur["early"] = True
            cur["valid_apd"] = False
            cur["valid_di"] = False
            early.append(i)
    for i in range(1, len(beats)):
        if not beats[i - 1]["early"] and beats[i]["repol_found"]:
            beats[i]["valid_di"] = True
    return early


def apd_di(beats):
    """compute_apd_di (postProcessing_restCurves.py:270-297) on classified beats: the last
    two beats as S1, S2; (APD90 of S2, DI = S2 upstroke - S1 repolarisation), or
    (None, None) when there are fewer than two beats, S1 is early, either APD or S2's DI
    is not valid, or APD or DI is not positive."""
    if len(beats) < 2:
        return None, None
    s1, s2 = beats[-2], beats[-1]
    if s1["early"] or not (s1["valid_apd"] and s2["valid_apd"] and s2["valid_di"]):
        return None, None
    di = s2["t_up"] - s1["t_repol90"]
    apd = s2["t_repol90"] - s2["t_up"]
    if di <= 0 or apd <= 0:
        return None, None
    return apd, di


def detect_beats(t, vm, max_beats=256, dvdt_threshold=5000.0, baseline_s=0.002, peak_s=2.0):
    beats = (Beat * max_beats)()
    n = lib().orc_detect_beats(np.ascontiguousarray(t, dtype=np.float64),
                               np.ascontiguousarray(vm, dtype=np.float64), len(t), beats,
                               max_beats, dvdt_threshold, baseline_s, peak_s)
    return [dict(t_up=b.t_up, v_base=b.v_base, v_peak=b.v_peak, v90=b.v90,
                 t_repol90=b.t_repol90, repol_found=bool(b.repol_found)) for b in beats[:n]]
