"""Host mirror of MyocardiumDomain + MonodomainSolver on a uniform hex block.

A ``Monodomain`` owns the device mesh, the ionic model and the tissue state and runs
the reference's step (myocardiumDomain.C:487-529): box stimulus at t0, solveODE,
solveDiffusionExplicit/Implicit, updateActivationTime.  Construction reads the same
electroProperties keys as MyocardiumDomain / MonodomainSolver
(myocardiumDomain.C:180-277, monodomainSolver.C:48-100, stimulusIO.C:305-390).
"""
import ctypes as C
import functools
import os
import weakref

import numpy as np

from . import _capi
from ._capi import check, lib
from .foam_dict import symm_to_full
from .ionic import IonicModel, FatalIOError


def _dp(a):
    return a.ctypes.data_as(C.c_void_p)


class _Mesh:
    """The device mesh handle and its conductivity (shared by HexMesh and PolyMesh)."""

    def __del__(self):
        if getattr(self, "m", None):
            try:
                lib().cg_mesh_destroy(self.m)
            except Exception:
                pass
            self.m = None

    def set_conductivity(self, sigma):
        """sigma: 9 values (uniform tensor), 6 (symmTensor) or [n, 9]/[n, 3, 3] per cell."""
        s = np.asarray(sigma, dtype=np.float64)
        if s.size == 6:
            s = np.asarray(symm_to_full(s.ravel()))
        if s.size == 9:
            s = np.ascontiguousarray(s.ravel())
            check(lib().cg_mesh_set_conductivity(self.m, _dp(s), 1))
        else:
            s = np.ascontiguousarray(s.reshape(self.n, 9))
            check(lib().cg_mesh_set_conductivity(self.m, _dp(s), 0))
        self._sigma = s

    def read_conductivity(self, path):
        """Per-cell conductivity from a volTensorField file (the `conductivity` field that
        monodomainSolver.C:48-100 reads when it is present); True if the file exists."""
        from . import foam_io
        if not os.path.exists(path):
            return False
        f = foam_io.read_field(path)
        if f.cls != "volTensorField":
            raise FatalIOError("%s: conductivity must be a volTensorField" % path)
        v = f.expand(self.n)
        self.set_conductivity(v if not f.uniform else v[0])
        return True

    def explicit_dt(self, chi, cm, max_co):
        dt = C.c_double()
        check(lib().cg_mesh_explicit_dt(self.m, chi, cm, max_co, C.byref(dt)))
        return dt.value


class HexMesh(_Mesh):
    """Uniform nx*ny*nz hex block, blockMesh cell order c = i + nx*(j + ny*k)."""

    def __init__(self, nx, ny, nz, h, centres=None):
        """centres: optional [n, 3] cell centres as the fvMesh holds them (mesh().C()); the
        box stimulus then tests them instead of the (i + 1/2) h lattice."""
        hx, hy, hz = (h, h, h) if np.isscalar(h) else h
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.h = (float(hx), float(hy), float(hz))
        self.n = self.nx * self.ny * self.nz
        self._centres = None
        m = C.c_void_p()
        check(lib().cg_mesh_create_hex(self.nx, self.ny, self.nz, *self.h, C.byref(m)))
        self.m = m
        if centres is not None:
            self.set_cell_centres(centres)

    @classmethod
    def from_block_mesh_dict(cls, path):
        """The mesh blockMesh writes for system/blockMeshDict (path to the file or to the
        case): one straight-edged, axis-aligned hex block with simpleGrading (1 1 1), the
        form of every tissue tutorial.  Cell centres are those OpenFOAM computes
        (`meshes.block_cell_centres`), boundary faces in the dictionary's patch order."""
        import os.path as op
        from . import foam_dict, meshes
        if op.isdir(path):
            path = op.join(path, "system", "blockMeshDict")
        d = foam_dict.read(path)
        scale = float(d.get("scale", d.get("convertToMeters", 1.0)))
        v = np.asarray(d["vertices"], dtype=np.float64)
        blocks = d["blocks"]
        if len(blocks) != 5 or blocks[0] != "hex" or blocks[3] != "simpleGrading":
            raise FatalIOError("%s: the hex path takes one hex block with simpleGrading" % path)
        labels, n, grading = blocks[1], [int(a) for a in blocks[2]], blocks[4]
        if any(float(g) != 1.0 for g in np.ravel(grading)) or d.get("edges"):
            raise FatalIOError("%s: the hex path takes uniform grading and straight edges"
                               % path)
        bv = v[np.asarray(labels)]
        lo, hi = bv[0], bv[6]
        want = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
                         [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                         [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]])
        if not np.array_equal(bv, want) or np.any(hi <= lo):
            raise FatalIOError("%s: the block is not an axis-aligned box in hex vertex order"
                               % path)
        # boundary patches: each block face listed, matched to its side of the hex model
        side_of = {frozenset(labels[q] for q in f): meshes.SIDES[s]
                   for s, f in enumerate(meshes.HEX_FACES)}
        patches = []
        bd = d.get("boundary", [])
        for q in range(0, len(bd), 2):
            faces = bd[q + 1].get("faces", [])
            patches.append(tuple(side_of[frozenset(int(a) for a in f)] for f in faces))
        listed = [s for p in patches for s in p]
        if sorted(listed) != sorted(meshes.SIDES):
            # sides not listed go to blockMesh's defaultFaces patch, after the others
            patches.append(tuple(s for s in meshes.SIDES if s not in listed))
        if np.any(lo != 0.0):
            raise FatalIOError("%s: the hex path takes a block with its corner at the origin"
                               % path)
        h = tuple((hi[q] - lo[q]) * scale / n[q] for q in range(3))
        centres = meshes.block_cell_centres(lo, hi, n, scale, patches=tuple(patches))
        return cls(n[0], n[1], n[2], h, centres=centres)

    def set_cell_centres(self, centres):
        """The fvMesh's cell centres [n, 3] (m) for the box stimulus test
        (cg_mesh_set_cell_centres)."""
        c = np.ascontiguousarray(centres, dtype=np.float64).reshape(self.n, 3)
        check(lib().cg_mesh_set_cell_centres(self.m, _dp(c)))
        self._centres = c

    def cell_centres(self):
        if self._centres is not None:
            return self._centres
        i = np.arange(self.nx)
        j = np.arange(self.ny)
        k = np.arange(self.nz)
        K, J, I = np.meshgrid(k, j, i, indexing="ij")
        return np.stack([(I.ravel() + 0.5) * self.h[0], (J.ravel() + 0.5) * self.h[1],
                         (K.ravel() + 0.5) * self.h[2]], axis=1)


class PolyMesh(_Mesh):
    """Any polyhedral mesh as the fvMesh provides it (cg_mesh_create_general): internal
    faces in lduAddressing order (owner < neighbour) with area vectors Sf [F, 3] pointing
    from owner to neighbour and centres Cf [F, 3], cell centres C [n, 3] and volumes V [n],
    boundary faces by owner cell and outward area vector (zeroGradient for Vm)."""

    def __init__(self, owner, neighbour, Sf, Cf, C_, V, boundary_owner, boundary_Sf):
        own = np.ascontiguousarray(owner, dtype=np.int32)
        nbr = np.ascontiguousarray(neighbour, dtype=np.int32)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Sf, Cf, C_, V)]
        bown = np.ascontiguousarray(boundary_owner, dtype=np.int32)
        bsf = np.ascontiguousarray(boundary_Sf, dtype=np.float64)
        self.n = int(arrs[3].size)
        self.C = arrs[2].reshape(self.n, 3).copy()
        self.V = arrs[3].copy()
        m = C.c_void_p()
        check(lib().cg_mesh_create_general(self.n, own.size, own.ctypes.data_as(C.c_void_p),
                                           nbr.ctypes.data_as(C.c_void_p), _dp(arrs[0]),
                                           _dp(arrs[1]), _dp(arrs[2]), _dp(arrs[3]), bown.size,
                                           bown.ctypes.data_as(C.c_void_p), _dp(bsf),
                                           C.byref(m)))
        self.m = m

    def cell_centres(self):
        return self.C


def fv_solution_controls(fv_solution, field="Vm"):
    """Solver controls of `field` from a parsed fvSolution, looked up as fvMatrix::solve and
    pimpleControl do: the first solvers entry whose key (possibly a regular expression such
    as "Vm|VmFinal") matches the field (tolerance, relTol, maxIter with OpenFOAM's default
    1000), and PIMPLE nOuterCorrectors (default 1) and residualControl tolerance (default
    0).  The residual norm is OpenFOAM's."""
    import re
    entry = None
    for key, d in fv_solution.get("solvers", {}).items():
        if re.fullmatch(key.strip('"'), field):
            entry = d
            break
    if entry is None:
        raise FatalIOError("keyword %s is undefined in solvers" % field)
    if "tolerance" not in entry:
        raise FatalIOError("keyword tolerance is undefined in solvers/%s" % field)
    pimple = fv_solution.get("PIMPLE", {})
    rc = 0.0
    for key, d in pimple.get("residualControl", {}).items():
        if re.fullmatch(key.strip('"'), field):
            rc = float(d.get("tolerance", 0.0))
            break
    return {"norm": "OpenFOAM", "tolerance": float(entry["tolerance"]),
            "relTol": float(entry.get("relTol", 0.0)), "maxIter": int(entry.get("maxIter", 1000)),
            "nOuterCorrectors": int(pimple.get("nOuterCorrectors", 1)), "residualControl": rc}


def _scheme_entry(table, name, kind):
    """fvSchemes lookup as fvSchemes::ddtScheme / gradScheme / laplacianScheme do: the
    term's own entry (a key or a quoted regular expression), else `default` unless that
    is `none`."""
    import re
    table = table or {}
    for key, v in table.items():
        if key == name or (key.startswith('"') and re.fullmatch(key.strip('"'), name)):
            return v
    d = table.get("default")
    if d is None or d == "none":
        raise FatalIOError("keyword %s is undefined in dictionary %sSchemes" % (name, kind))
    return d


def _words(v):
    return [str(a) for a in (v if isinstance(v, list) else [v])]


def fv_schemes_selection(fv_schemes, general_mesh=False):
    """The schemes system/fvSchemes selects for the monodomain equation
    chi Cm fvm::ddt(Vm) == fvc/fvm::laplacian(conductivity, Vm) - chi Cm Iion + source
    (monodomainSolver.C:103-150): {"ddt": "Euler" | "backward", "grad": ..., "laplacian":
    ...}.  Raises FatalIOError for what the device path does not implement, rather than
    substituting another scheme:
    - ddt(Vm): Euler or backward (BDF2 with an Euler first step);
    - grad(Vm) (the tensor correction's gradient): Gauss linear, or leastSquares on the
      hex block, where with zeroGradient walls it equals Gauss linear (both give the
      central difference, the wall cells the one-sided half difference);
    - laplacian(conductivity,Vm): Gauss linear corrected (uncorrected and orthogonal are
      the same operator on the orthogonal hex block);
    - interpolate(conductivity): linear."""
    ddt = " ".join(_words(_scheme_entry(fv_schemes.get("ddtSchemes"), "ddt(Vm)", "ddt")))
    if ddt not in ("Euler", "backward"):
        raise FatalIOError("ddt scheme '%s' for Vm is not implemented by the device path "
                           "(Euler, backward)" % ddt)
    grad = " ".join(_words(_scheme_entry(fv_schemes.get("gradSchemes"), "grad(Vm)", "grad")))
    ok_grad = ["Gauss linear"] + ([] if general_mesh else ["leastSquares"])
    if grad not in ok_grad:
        raise FatalIOError("grad scheme '%s' for Vm is not implemented by the device path on "
                           "this mesh (%s)" % (grad, ", ".join(ok_grad)))
    lap = " ".join(_words(_scheme_entry(fv_schemes.get("laplacianSchemes"),
                                        "laplacian(conductivity,Vm)", "laplacian")))
    ok_lap = ["Gauss linear corrected"] + ([] if general_mesh else
                                           ["Gauss linear uncorrected",
                                            "Gauss linear orthogonal"])
    if lap not in ok_lap:
        raise FatalIOError("laplacian scheme '%s' is not implemented by the device path on "
                           "this mesh (%s)" % (lap, ", ".join(ok_lap)))
    interp = " ".join(_words(_scheme_entry(fv_schemes.get("interpolationSchemes"),
                                           "interpolate(conductivity)", "interpolation")))
    if interp != "linear":
        raise FatalIOError("interpolation scheme '%s' is not implemented (linear)" % interp)
    return {"ddt": ddt, "grad": grad, "laplacian": lap, "interpolation": interp}


def _vm_fed_by(domain_ref):
    d = domain_ref()
    if d is None or d.time[1] == 0:
        return None
    return d.field(_capi.FIELD_VM_OLD)


class Monodomain:
    """Device monodomain tissue: MyocardiumDomain state + MonodomainSolver operator."""

    def __init__(self, mesh, coeffs, dt=None, max_co=None, precision=None, tol=1e-8,
                 max_iter=2000, ddt=None, vm0=-0.084, case_dir=None, start_time="0",
                 fv_schemes=None):
        """case_dir: an OpenFOAM case.  Its system/fvSchemes selects the schemes
        (fv_schemes_selection; a parsed dictionary may be passed as fv_schemes instead),
        and its <start_time>/ directory may hold the fields the reference reads if
        present: conductivity (monodomainSolver.C:54-92, taking precedence over the
        dictionary's value), Vm and activationTime.  Without either, ddt ("Euler" or
        "backward") is the caller's choice, Euler when not given."""
        self.mesh = mesh
        if fv_schemes is None and case_dir:
            from . import foam_dict
            path = os.path.join(case_dir, "system", "fvSchemes")
            if not os.path.exists(path):
                raise FatalIOError("cannot find file %s" % path)
            fv_schemes = foam_dict.read(path)
        self.schemes = (fv_schemes_selection(fv_schemes, isinstance(mesh, PolyMesh))
                        if fv_schemes is not None else None)
        if ddt is None:
            ddt = self.schemes["ddt"] if self.schemes else "Euler"
        if ddt not in ("Euler", "backward"):
            raise FatalIOError("ddt scheme '%s' is not implemented (Euler, backward)" % ddt)
        self.coeffs = coeffs
        cpath = os.path.join(case_dir, start_time, "conductivity") if case_dir else None
        if not (cpath and mesh.read_conductivity(cpath)):
            if "conductivity" not in coeffs:
                raise FatalIOError("keyword conductivity is undefined")
            cond = coeffs["conductivity"]
            mesh.set_conductivity(cond if not isinstance(cond, np.ndarray) else cond)
        self.chi = float(coeffs["chi"])
        self.cm = float(coeffs["cm"])
        alg = coeffs.get("solutionAlgorithm", "implicit")
        if alg not in ("explicit", "implicit"):
            raise FatalIOError("Unrecognised solution algorithm %s" % alg)
        self.algorithm = alg
        if alg == "explicit" and dt is None:
            # electroModel::setDeltaT (electroModel.C:238-257): dt from maxCo, once
            dt = mesh.explicit_dt(self.chi, self.cm, 0.1 if max_co is None else max_co)
        if dt is None:
            raise ValueError("dt is required for the implicit algorithm")
        self.dt = float(dt)
        self.ionic = IonicModel.New(coeffs, mesh.n, self.dt, False, precision=precision)
        # the tissue feeds its cells V^n each step: the ionic io hooks read it back (a weak
        # reference, so the handles are freed with the domain)
        self.ionic._vm_fed = functools.partial(_vm_fed_by, weakref.ref(self))
        if coeffs.get("singleCellStimulus"):
            raise FatalIOError("Detected active ionic-model stimulus protocol while using a "
                               "reaction-diffusion myocardium domain.")
        self.ddt = ddt
        cfg = _capi.TissueConfig(
            _capi.EXPLICIT if alg == "explicit" else _capi.IMPLICIT,
            _capi.DDT_BACKWARD if ddt == "backward" else _capi.DDT_EULER,
            self.chi, self.cm, self.dt, vm0, tol, max_iter)
        t = C.c_void_p()
        check(lib().cg_tissue_create(mesh.m, self.ionic.h, C.byref(cfg), C.byref(t)))
        self.t = t
        stim = coeffs.get("externalStimulus")
        if stim:
            self.set_external_stimulus(stim)
        if case_dir:
            self.read_fields(case_dir, start_time)

    def __del__(self):
        if getattr(self, "t", None):
            try:
                lib().cg_tissue_destroy(self.t)
            except Exception:
                pass
            self.t = None

    def set_external_stimulus(self, d):
        """externalStimulus sub-dictionary (stimulusIO.C:305-390)."""
        if "stimulusLocationMinList" in d:
            lo = np.asarray(d["stimulusLocationMinList"], dtype=np.float64).reshape(-1, 3)
            hi = np.asarray(d["stimulusLocationMaxList"], dtype=np.float64).reshape(-1, 3)
        else:
            lo = np.asarray([d["stimulusLocationMin"]], dtype=np.float64)
            hi = np.asarray([d["stimulusLocationMax"]], dtype=np.float64)
        nb = len(lo)
        t0 = np.asarray(d.get("stimulusStartTimeList", [d.get("stimulusStartTime", 0.0)] * nb),
                        dtype=np.float64)
        dur = np.asarray(d.get("stimulusDurationList", [d.get("stimulusDuration", 0.0)] * nb),
                         dtype=np.float64)
        inten = np.asarray(d.get("stimulusIntensityList", [d.get("stimulusIntensity", 0.0)] * nb),
                           dtype=np.float64)
        for name, a in (("stimulusStartTimeList", t0), ("stimulusDurationList", dur),
                        ("stimulusIntensityList", inten)):
            if len(a) != nb:
                raise FatalIOError("%s must have the same size as the stimulus box list." % name)
        self.set_boxes(lo, hi, t0, dur, inten)

    def set_boxes(self, lo, hi, t0, dur, inten):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (lo, hi, t0, dur, inten)]
        check(lib().cg_tissue_set_stimulus_boxes(self.t, len(arrs[2]), *[_dp(a) for a in arrs]))
        self._boxes = arrs

    def step(self, n=1):
        check(lib().cg_tissue_step(self.t, int(n)))

    def sync(self):
        check(lib().cg_tissue_sync(self.t))

    def field(self, which):
        out = np.empty(self.mesh.n)
        check(lib().cg_tissue_get_field(self.t, which, _dp(out)))
        return out

    def set_field(self, which, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().cg_tissue_set_field(self.t, which, _dp(a)))

    @property
    def Vm(self):
        return self.field(_capi.FIELD_VM)

    @property
    def Iion(self):
        return self.field(_capi.FIELD_IION)

    @property
    def activationTime(self):
        return self.field(_capi.FIELD_ACTIVATION)

    @property
    def time(self):
        t = C.c_double()
        s = C.c_int64()
        check(lib().cg_tissue_time(self.t, C.byref(t), C.byref(s)))
        return t.value, s.value

    def set_profiling(self, on=True):
        check(lib().cg_tissue_set_profiling(self.t, int(on)))

    def kernel_times(self):
        """{kernel: (total ms, launches)} of the HIP-event timer."""
        names = C.create_string_buffer(32 * 16)
        ms = (C.c_double * 16)()
        cnt = (C.c_int64 * 16)()
        n = C.c_int()
        check(lib().cg_tissue_kernel_times(self.t, 16, names, ms, cnt, C.byref(n)))
        out = {}
        for k in range(n.value):
            nm = names.raw[32 * k:32 * (k + 1)].split(b"\0")[0].decode()
            out[nm] = (ms[k], cnt[k])
        return out

    def apply_operator(self, x, y, ddt_coeff=1.0, precision="fp64"):
        """y = A x on the device with the implicit operator (fvm::ddt + fvm::laplacian,
        orthogonal part, per unit volume).  x, y: contiguous CUDA tensors of float64
        (precision "fp64") or float32 ("mixed").  Asynchronous on the tissue's stream."""
        prec = {"fp64": 0, "mixed": 1}[precision]
        want = 8 if prec == 0 else 4
        for a in (x, y):
            if not (a.is_cuda and a.is_contiguous() and a.element_size() == want
                    and a.numel() == self.mesh.n):
                raise ValueError("apply_operator needs contiguous device tensors of %d cells "
                                 "and %d-byte elements" % (self.mesh.n, want))
        check(lib().cg_tissue_apply_operator(self.t, float(ddt_coeff), prec,
                                             C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())))

    def solver_stats(self):
        it = C.c_int64()
        r = C.c_double()
        check(lib().cg_tissue_solver_stats(self.t, C.byref(it), C.byref(r)))
        return it.value, r.value

    def set_solver_controls(self, norm="OpenFOAM", tolerance=1e-11, relTol=0.0, maxIter=1000,
                            nOuterCorrectors=1, residualControl=0.0):
        """fvSolution controls of the Vm solve: norm "OpenFOAM" (lduMatrix normalised
        residual) or "L2" (relative); PIMPLE nOuterCorrectors and the residualControl
        tolerance of the outer loop."""
        k = {"L2": _capi.NORM_L2, "OpenFOAM": _capi.NORM_OPENFOAM}[norm]
        check(lib().cg_tissue_set_solver_controls(self.t, k, float(tolerance), float(relTol),
                                                  int(maxIter), int(nOuterCorrectors),
                                                  float(residualControl)))

    def set_fv_solution(self, fv_solution, field="Vm"):
        """Applies fv_solution_controls(fv_solution, field)."""
        self.set_solver_controls(**fv_solution_controls(fv_solution, field))

    def diffusion_step(self, Iion, source=None):
        """solveDiffusionExplicit / solveDiffusionImplicit (monodomainSolver.C:103-150) with
        the caller's ionic current Iion [n] (V/s) and optional source [n] (A/m^3; default:
        the stimulus boxes): one step without the tissue's ionic model
        (cg_tissue_diffusion_step)."""
        i = np.ascontiguousarray(Iion, dtype=np.float64)
        src = None if source is None else np.ascontiguousarray(source, dtype=np.float64)
        if i.size != self.mesh.n or (src is not None and src.size != self.mesh.n):
            raise ValueError("fields of %d cells expected" % self.mesh.n)
        check(lib().cg_tissue_diffusion_step(self.t, _dp(i), None if src is None else _dp(src),
                                             0))

    def pseudo_ecg(self, positions):
        """PseudoECGSolver::compute (pseudoECGSolver.C:36-83) of the current Vm at the
        electrode positions [E][3] (m): the dipole sum per electrode, [E]."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pos))
        check(lib().cg_tissue_pseudo_ecg(self.t, len(pos), _dp(pos), _dp(out)))
        return out

    # ------------------------------------------------------------ on-disk fields
    def external_stimulus_current(self):
        """sourceField_ of the last step (myocardiumDomain.C:307-342), A/m^3: the summed
        intensities of the boxes active at that step's start time t0 (inclusive ends)
        whose closed extent holds the cell centre; zero before the first step."""
        t, steps = self.time
        out = np.zeros(self.mesh.n)
        boxes = getattr(self, "_boxes", None)
        if steps == 0 or boxes is None:
            return out
        t0 = (t + 0.0) - self.dt
        lo, hi, ts, dur, inten = boxes
        c = self.mesh.cell_centres()
        for b in range(len(ts)):
            if t0 < ts[b] or t0 > ts[b] + dur[b]:
                continue
            inside = np.all((c >= lo[3 * b:3 * b + 3]) & (c <= hi[3 * b:3 * b + 3]), axis=1)
            out[inside] += inten[b]
        return out

    def exported_fields(self):
        """The ionic fields of outputVariables.ionic.export as MyocardiumDomain::exportStates
        fills them at an output time, after the step (electroModel.C:321-330,
        ionicModelIO.C:434-500): [(name, values)] for the names that resolve
        (exportedFieldNamesRef), each the raw model-unit value: Vm-like names and states
        from the cells' states, whose V is the one the last step fed them (the model's V of
        Vm at the step's start, BuenoOrovio.C:119-122); algebraic variables and RATES_ as
        computeVariables leaves them at the step's end (BuenoOrovio.C:134-147), with no
        stimulus and dV/dt = 0 in tissue mode."""
        from . import foam_io
        out = (self.coeffs or {}).get("outputVariables", {})
        ionic = out.get("ionic") if isinstance(out, dict) else None
        export = ionic.get("export") if isinstance(ionic, dict) else None
        if not export:
            return []
        plan = foam_io.trace_plan(self.ionic.model_key, self.ionic.stateVariableNames(), export,
                                  self.ionic.algebraic_names())
        S = self.ionic.io_states()
        srcs = {src for _, src, _ in plan}
        R = self.ionic.eval("rates", S) if "rate" in srcs else None
        A = self.ionic.eval("algebraic", S) if "alg" in srcs else None
        got = []
        for name, src, i in plan:
            got.append((name, S[i] if src in ("vm", "state") else R[i] if src == "rate"
                        else A[i]))
        return got

    def write_fields(self, case_dir, precision=6, time_name=None):
        """MyocardiumDomain::write (myocardiumDomain.C:617-638): Vm,
        externalStimulusCurrent, ionicCurrent and activationTime, then the exported ionic
        fields (exported_fields, dimensionless) as OpenFOAM ASCII volScalarFields in
        <case_dir>/<timeName>/.  Returns that directory.  The write order is the
        reference's, so an export list naming Vm (the NiedererEtAl2012 case's (Vm Jsi))
        leaves the file Vm holding the raw V state in model units, as the reference's
        second write of that name does (electroModel.C:216-224 writes the registered
        fields first, MyocardiumDomain::write the exported ones after Vm_)."""
        from . import foam_io
        tn = time_name or foam_io.time_name(self.time[0])
        d = os.path.join(case_dir, tn)
        for name, v in (("Vm", self.Vm), ("externalStimulusCurrent",
                                          self.external_stimulus_current()),
                        ("ionicCurrent", self.Iion), ("activationTime", self.activationTime)):
            foam_io.write_field(os.path.join(d, name), v, location=tn, precision=precision)
        for name, v in self.exported_fields():
            foam_io.write_field(os.path.join(d, name), v, name=name, location=tn,
                                dimensions=(0,) * 7, precision=precision)
        return d

    def read_fields(self, case_dir, time_name="0"):
        """The READ_IF_PRESENT fields of MyocardiumDomain (myocardiumDomain.C:196-251) from
        <case_dir>/<timeName>/: Vm (also the old level, so a first backward step starts
        from it) and activationTime.  Returns the names read."""
        from . import foam_io
        got = []
        for name, field in (("Vm", _capi.FIELD_VM), ("activationTime", _capi.FIELD_ACTIVATION)):
            path = os.path.join(case_dir, time_name, name)
            if not os.path.exists(path):
                continue
            v = foam_io.read_field(path).expand(self.mesh.n)
            self.set_field(field, v)
            if field == _capi.FIELD_VM:
                self.set_field(_capi.FIELD_VM_OLD, v)
            got.append(name)
        return got

    def set_activation_threshold(self, v):
        """Activation threshold in V for the steps still to run (0, the default, is the
        reference's, myocardiumDomain.C:362-371)."""
        check(lib().cg_tissue_set_activation_threshold(self.t, float(v)))

    def solver_report(self):
        """last step: (linear solves, initial residual of the last corrector, final residual)"""
        n = C.c_int()
        r0, r1 = C.c_double(), C.c_double()
        check(lib().cg_tissue_solver_report(self.t, C.byref(n), C.byref(r0), C.byref(r1)))
        return n.value, r0.value, r1.value
