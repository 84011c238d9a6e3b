"""ctypes binding of libcardiacgpu.so (the C-ABI declared in include/cardiacgpu.h).

The library is loaded from this package directory; there is no fallback.  A missing
library or a failing call raises immediately.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcardiacgpu.so")

CG_OK = 0
ERRORS = {1: "ARG", 2: "MODEL", 3: "TISSUE", 4: "HIP", 5: "NOMEM", 6: "NONFINITE",
          7: "NOT_CONVERGED", 8: "STIMULUS", 9: "STATE", 10: "ODE"}
FP64, MIXED = 0, 1
DEVICE_PTRS = 1
ODE_RUSH_LARSEN, ODE_RKF45, ODE_EULER = 0, 1, 2
NORM_L2, NORM_OPENFOAM = 0, 1
EXPLICIT, IMPLICIT = 0, 1
DDT_EULER, DDT_BACKWARD = 0, 1
FIELD_VM, FIELD_IION, FIELD_ACTIVATION, FIELD_VM_OLD = 0, 1, 2, 3
RATES, ALGEBRAIC = 0, 1



class CgError(RuntimeError):
    pass


class Stimulus(C.Structure):
    _fields_ = [("stim_start", C.c_double), ("stim_period_s1", C.c_double),
                ("stim_duration", C.c_double), ("stim_amplitude", C.c_double),
                ("nstim1", C.c_int), ("stim_period_s2", C.c_double), ("nstim2", C.c_int)]


class TissueConfig(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("ddt", C.c_int), ("chi", C.c_double),
                ("cm", C.c_double), ("dt", C.c_double), ("vm0", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int)]


_lib = None


def lib():
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CgError("libcardiacgpu.so not built (run python -m cardiacfoam_amd.build): %s"
                      % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, dp, ip64 = C.c_void_p, C.POINTER(C.c_double), C.c_int64
    L.cg_last_error.restype = C.c_char_p
    L.cg_version.restype = C.c_char_p
    L.cg_model_id.argtypes = [C.c_char_p]
    L.cg_tissue_flag.argtypes = [C.c_char_p]
    L.cg_model_nstates.argtypes = [C.c_int]
    L.cg_model_state_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_size_t]
    L.cg_model_units.argtypes = [C.c_int, dp, dp, dp, dp]
    L.cg_tissue_set_activation_threshold.argtypes = [vp, C.c_double]
    L.cg_ionic_create.argtypes = [C.c_int, C.c_int, ip64, C.c_int, C.c_int, C.POINTER(vp)]
    L.cg_ionic_destroy.argtypes = [vp]
    L.cg_ionic_nstates.argtypes = [vp]
    L.cg_ionic_set_constant.argtypes = [vp, C.c_int, C.c_double]
    L.cg_ionic_set_conductance_scale.argtypes = [vp, vp]
    L.cg_ionic_set_stimulus.argtypes = [vp, C.POINTER(Stimulus)]
    L.cg_tissue_set_solver_controls.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int,
                                                C.c_int, C.c_double]
    L.cg_tissue_solver_report.argtypes = [vp, C.POINTER(C.c_int), dp, dp]
    ip32 = C.POINTER(C.c_int32)
    L.cg_ldu_create.argtypes = [ip64, ip64, ip32, ip32, C.POINTER(vp)]
    L.cg_ldu_destroy.argtypes = [vp]
    L.cg_ldu_width.argtypes = [vp]
    L.cg_ldu_set_coeffs.argtypes = [vp, vp, vp, C.c_int]
    L.cg_ldu_amul.argtypes = [vp, vp, vp, C.c_int]
    L.cg_ldu_amul_timed.argtypes = [vp, vp, vp, C.POINTER(C.c_float)]
    L.cg_ldu_solve.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                               C.POINTER(C.c_int), dp, dp]
    L.cg_ionic_set_integrator.argtypes = [vp, C.c_int, C.c_double, C.c_double, ip64,
                                          C.c_double]
    L.cg_ionic_step.argtypes = [vp, C.c_double, C.c_double, vp, vp, C.c_int]
    L.cg_ionic_run_cells.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.cg_tissue_pseudo_ecg.argtypes = [vp, C.c_int, vp, vp]
    L.cg_tissue_pseudo_ecg_noalloc.argtypes = [vp, C.c_int, vp, vp]
    L.cg_tissue_diffusion_step.argtypes = [vp, vp, vp, C.c_int]
    L.cg_mesh_create_general.argtypes = [C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp,
                                         C.c_int64, vp, vp, vp]
    L.cg_ionic_run_cells_trace.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int64, vp]
    L.cg_format_scalars.argtypes = [vp, C.c_int64, C.c_int, C.c_int, C.c_int64, vp, C.c_int64,
                                    C.POINTER(C.c_int64)]
    L.cg_parse_scalars.argtypes = [vp, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.cg_ionic_get_states.argtypes = [vp, C.c_int, C.c_int, vp]
    L.cg_ionic_set_states.argtypes = [vp, C.c_int, C.c_int, vp]
    L.cg_mesh_create_hex.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.POINTER(vp)]
    L.cg_mesh_match_ldu.argtypes = [C.c_int, C.c_int, C.c_int, ip64, vp, vp]
    L.cg_mesh_destroy.argtypes = [vp]
    L.cg_mesh_ncells.argtypes = [vp]
    L.cg_mesh_ncells.restype = C.c_int64
    L.cg_mesh_set_conductivity.argtypes = [vp, vp, C.c_int]
    L.cg_mesh_set_cell_centres.argtypes = [vp, vp]
    L.cg_mesh_explicit_dt.argtypes = [vp, C.c_double, C.c_double, C.c_double, dp]
    L.cg_mesh_laplacian.argtypes = [vp, vp, vp]
    L.cg_tissue_create.argtypes = [vp, vp, C.POINTER(TissueConfig), C.POINTER(vp)]
    L.cg_tissue_destroy.argtypes = [vp]
    L.cg_tissue_set_stimulus_boxes.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    L.cg_tissue_step.argtypes = [vp, C.c_int]
    L.cg_tissue_sync.argtypes = [vp]
    L.cg_tissue_get_field.argtypes = [vp, C.c_int, vp]
    L.cg_tissue_set_field.argtypes = [vp, C.c_int, vp]
    L.cg_tissue_time.argtypes = [vp, dp, C.POINTER(C.c_int64)]
    L.cg_tissue_solver_stats.argtypes = [vp, C.POINTER(C.c_int64), dp]
    L.cg_tissue_set_profiling.argtypes = [vp, C.c_int]
    L.cg_tissue_kernel_times.argtypes = [vp, C.c_int, C.c_char_p, dp, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int)]
    L.cg_tissue_apply_operator.argtypes = [vp, C.c_double, C.c_int, vp, vp]
    L.cg_ionic_set_protocols.argtypes = [vp, vp]
    L.cg_ionic_set_run_lengths.argtypes = [vp, vp, ip64]
    L.cg_ionic_enable_beats.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
    L.cg_batch_run.argtypes = [C.POINTER(vp), C.c_int, C.c_double, C.c_int]
    L.cg_batch_autotune.argtypes = [C.POINTER(vp), C.c_int, C.c_double, C.c_int, C.c_char_p,
                                    C.c_size_t]
    L.cg_ionic_get_beats.argtypes = [vp, vp, vp]
    L.cg_ionic_eval.argtypes = [vp, C.c_int, ip64, vp, vp, vp]
    L.cg_model_nalgebraic.argtypes = [C.c_int]
    L.cg_model_algebraic_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_size_t]
    L.cg_ionic_get_restitution.argtypes = [vp, vp, vp]
    L.cg_adjoint_create.argtypes = [vp, C.c_int, vp, C.c_int, C.POINTER(vp)]
    L.cg_adjoint_destroy.argtypes = [vp]
    L.cg_adjoint_forward.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_int64)]
    L.cg_adjoint_backward.argtypes = [vp, C.c_int64, C.c_int, vp, vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != CG_OK:
        msg = lib().cg_last_error().decode(errors="replace")
        raise CgError("cardiacgpu error %s (%d): %s" % (ERRORS.get(rc, "?"), rc, msg))
    return rc


def header_symbols(header=None):
    """Function names declared in include/cardiacgpu.h."""
    import re
    if header is None:
        header = os.path.join(os.path.dirname(HERE), "include", "cardiacgpu.h")
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", text)))


def missing_symbols():
    """Names of header functions the loaded library does not export."""
    L = lib()
    out = []
    for s in header_symbols():
        try:
            getattr(L, s)
        except AttributeError:
            out.append(s)
    return out
