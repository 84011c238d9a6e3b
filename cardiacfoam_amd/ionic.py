"""Host mirror of cardiacFoam's ionicModel runtime selection (the GPU is the engine).

``IonicModel.New(dict, nIntegrationPoints, initialDeltaT, solveVmWithinODESolver)``
follows ionicModel::New (src/ionicModels/ionicModel/ionicModel.C:76-94): the model is
looked up by the ``ionicModel`` key, the tissue by ``tissue``
(ionicSelector.C:32-66), and the same errors are raised for unknown names.
``solveODE(stepStartTime, deltaT, Vm, Im)`` has the meaning of ionicModel.H:250-256
(seconds, volts, V/s) and runs on the device through the C-ABI.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib

MODEL_KEYS = ["BuenoOrovio", "TNNP", "ORd", "AlievPanfilov", "monodomainFDAManufactured",
              "Courtemanche", "PerisYague_2022", "Stewart", "Grandi", "ToRORd_dynCl",
              "Gaur"]
SUPPORTED_TISSUES = {
    "BuenoOrovio": ["endocardialCells", "mCells", "epicardialCells"],
    "TNNP": ["endocardialCells", "mCells", "epicardialCells"],
    "ORd": ["myocyte"],
    "AlievPanfilov": ["myocyte"],
    "Courtemanche": ["myocyte"],  # Courtemanche.C:92-95
    "PerisYague_2022": ["myocyte"],  # PerisYague.C:92-95
    "Stewart": ["myocyte"],  # Stewart.C:93-96
    "Grandi": ["myocyte"],  # Grandi.C:93-96
    "ToRORd_dynCl": ["endocardialCells", "mCells", "epicardialCells"],  # ToRORd_dynCl.C:93-96
    "Gaur": ["myocyte"],  # Gaur.C:93-96
    # the manufactured model's supportedDimensions() (monodomainFDAManufactured.C:85-88)
    "monodomainFDAManufactured": ["1D", "2D", "3D"],
}
# models selected by `dimension` instead of `tissue`
DIMENSION_MODELS = ("monodomainFDAManufactured",)


class FatalIOError(KeyError):
    """Raised where the reference calls FatalIOErrorInLookup / FatalErrorInFunction."""


def _ptr(a):
    """Host numpy array -> (pointer, flags); torch device tensor -> device pointer."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError("fields must be C-contiguous float64 arrays")
        return a.ctypes.data_as(C.c_void_p), 0
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(a, torch.Tensor):
        if a.dtype != torch.float64 or not a.is_contiguous():
            raise TypeError("tensors must be contiguous float64")
        if a.is_cuda:
            return C.c_void_p(a.data_ptr()), _capi.DEVICE_PTRS
        return C.c_void_p(a.data_ptr()), 0
    raise TypeError("unsupported field type %r" % type(a))


class IonicModel:
    """Device-resident ionic model for nIntegrationPoints cells (SoA states)."""

    def __init__(self, model_key, tissue, n, solve_vm=False, precision="fp64"):
        L = lib()
        self.model_key = model_key
        self.model = L.cg_model_id(model_key.encode())
        if self.model < 0:
            raise FatalIOError("Unknown ionicModel type %s\nValid ionicModel types: %s"
                               % (model_key, MODEL_KEYS))
        if tissue not in SUPPORTED_TISSUES[model_key]:
            raise FatalIOError("Unsupported tissue '%s' allowed: %s"
                               % (tissue, SUPPORTED_TISSUES[model_key]))
        self.tissue_name = tissue
        self.tissue = L.cg_tissue_flag(tissue.encode())
        self.n = int(n)
        self.solve_vm = bool(solve_vm)
        # tissue handles: callable giving the Vm (V) the last step was fed, or None
        self._vm_fed = None
        self.precision = precision
        prec = {"fp64": _capi.FP64, "mixed": _capi.MIXED}[precision]
        h = C.c_void_p()
        check(L.cg_ionic_create(self.model, self.tissue, self.n, int(solve_vm), prec, C.byref(h)))
        self.h = h
        self.nstates = L.cg_ionic_nstates(h)

    @staticmethod
    def New(dict_, nIntegrationPoints, initialDeltaT, solveVmWithinODESolver=False,
            precision=None):
        """ionicModel::New.  ``precision`` may also come from the dictionary key
        ``precision`` (fp64 | mixed), an addition of this backend."""
        if "ionicModel" not in dict_:
            raise FatalIOError("keyword ionicModel is undefined")
        key = dict_["ionicModel"]
        if key not in MODEL_KEYS:
            raise FatalIOError("Unknown ionicModel type %s\nValid ionicModel types: %s"
                               % (key, MODEL_KEYS))
        word_key = "dimension" if key in DIMENSION_MODELS else "tissue"
        if word_key not in dict_:
            raise FatalIOError("Normal model requires 'tissue' entry" if word_key == "tissue"
                               else "Manufactured model requires 'dimension' entry")
        prec = precision or dict_.get("precision", "fp64")
        m = IonicModel(key, str(dict_[word_key]).strip('"'), nIntegrationPoints,
                       solveVmWithinODESolver, prec)
        m.initialDeltaT = initialDeltaT
        solver = dict_.get("solver", "RushLarsen")
        # ODESolver::New keys (ionicModel.H:233-242); OpenFOAM's defaults when absent.
        # The kept per-cell step starts at initialDeltaT (ionicModel.C:38).
        m.set_integrator(solver, float(dict_.get("absTol", 1e-15)),
                         float(dict_.get("relTol", 1e-4)), int(dict_.get("maxSteps", 10000)),
                         float(initialDeltaT))
        stim = dict_.get("singleCellStimulus")
        if stim:
            m.set_stimulus_protocol(stim)
        return m

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().cg_ionic_destroy(h)
            except Exception:
                pass
            self.h = None

    # ------------------------------------------------------------------ API
    def solveODE(self, stepStartTime, deltaT, Vm, Im):
        pv, fv = _ptr(Vm)
        pi, fi = _ptr(Im)
        if fv != fi:
            raise TypeError("Vm and Im must both live on the host or both on the device")
        if len(Vm) != self.n or len(Im) != self.n:
            raise ValueError("field size %d/%d != nIntegrationPoints %d" % (len(Vm), len(Im), self.n))
        check(lib().cg_ionic_step(self.h, float(stepStartTime), float(deltaT), pv, pi, fv))
        # the V state of the step (BuenoOrovio.C:119-122) for the io hooks: a copy, since
        # the caller's Vm moves on with the diffusion step
        if fv:
            v = Vm.detach().clone()
            self._vm_fed = lambda v=v: v.cpu().numpy()
        else:
            v = np.array(Vm, dtype=np.float64, copy=True)
            self._vm_fed = lambda v=v: v

    ODE_SOLVERS = {"RushLarsen": _capi.ODE_RUSH_LARSEN, "RKF45": _capi.ODE_RKF45,
                   "Euler": _capi.ODE_EULER}

    def set_integrator(self, solver="RKF45", absTol=1e-15, relTol=1e-4, maxSteps=10000,
                       initial_step=0.0):
        """Per-cell ODE integrator: "RKF45" or "Euler" (OpenFOAM's adaptive ODESolvers of
        the reference's `solver RKF45;` / `solver Euler;`), or "RushLarsen".  initial_step (model time units) seeds the kept
        per-cell trial steps; 0 seeds them with the first step's deltaT in seconds."""
        if solver not in self.ODE_SOLVERS:
            raise FatalIOError("Unknown ODESolver type %s\nValid ODESolver types on the device: %s"
                               % (solver, sorted(self.ODE_SOLVERS)))
        check(lib().cg_ionic_set_integrator(self.h, self.ODE_SOLVERS[solver], float(absTol),
                                            float(relTol), int(maxSteps), float(initial_step)))
        self.solver = solver

    def run_cells(self, t0, dt, n_steps):
        check(lib().cg_ionic_run_cells(self.h, float(t0), float(dt), int(n_steps)))

    def run_cells_trace(self, t0, dt, n_steps, cell=0):
        """run_cells that returns cell `cell` after every step: (states, rates), each
        [n_steps, nstates] in model units (RATES at the step's end time)."""
        tr = np.empty((int(n_steps), 2, self.nstates))
        check(lib().cg_ionic_run_cells_trace(self.h, float(t0), float(dt), int(n_steps),
                                             int(cell), tr.ctypes.data_as(C.c_void_p)))
        return tr[:, 0, :], tr[:, 1, :]

    def set_stimulus_protocol(self, d):
        """singleCellStimulus sub-dictionary (stimulusIO.C:142-194)."""
        keys = ("stim_start", "stim_period_S1", "stim_duration", "stim_amplitude")
        if any(k in d for k in keys + ("nstim1", "stim_period_S2", "nstim2")) and \
                not all(k in d for k in keys):
            raise FatalIOError("Incomplete single-cell stimulus protocol. Required keys are: "
                               "(stim_start, stim_period_S1, stim_duration, stim_amplitude).")
        s1 = float(d.get("stim_period_S1", 0.0))
        s = _capi.Stimulus(float(d.get("stim_start", 0.0)), s1, float(d.get("stim_duration", 0.0)),
                           float(d.get("stim_amplitude", 0.0)),
                           int(d.get("nstim1", 1 if s1 > 1e-15 else 0)),
                           float(d.get("stim_period_S2", 0.0)), int(d.get("nstim2", 0)))
        check(lib().cg_ionic_set_stimulus(self.h, C.byref(s)))

    def set_constant(self, k, value):
        check(lib().cg_ionic_set_constant(self.h, int(k), float(value)))

    def set_conductance_scale(self, scale):
        if scale is None:
            check(lib().cg_ionic_set_conductance_scale(self.h, None))
            return
        a = np.ascontiguousarray(scale, dtype=np.float64).reshape(4, self.n)
        check(lib().cg_ionic_set_conductance_scale(self.h, a.ctypes.data_as(C.c_void_p)))

    def states(self, first=0, count=None):
        count = self.nstates - first if count is None else count
        out = np.empty((count, self.n))
        check(lib().cg_ionic_get_states(self.h, first, count, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_states(self, values, first=0):
        a = np.ascontiguousarray(values, dtype=np.float64)
        a = a.reshape(-1, self.n)
        check(lib().cg_ionic_set_states(self.h, first, a.shape[0], a.ctypes.data_as(C.c_void_p)))

    # --------------------------------------------------- batched single cells
    def set_protocols(self, protocols):
        """Per-cell S1/S2 protocols: a structured sequence of n dicts with the
        singleCellStimulus keys (model time units), or None for the handle's own protocol."""
        if protocols is None:
            check(lib().cg_ionic_set_protocols(self.h, None))
            return
        if len(protocols) != self.n:
            raise ValueError("need one protocol per cell")
        arr = (_capi.Stimulus * self.n)()
        for i, d in enumerate(protocols):
            s1 = float(d.get("stim_period_S1", 0.0))
            arr[i] = _capi.Stimulus(float(d.get("stim_start", 0.0)), s1,
                                    float(d.get("stim_duration", 0.0)),
                                    float(d.get("stim_amplitude", 0.0)),
                                    int(d.get("nstim1", 1 if s1 > 1e-15 else 0)),
                                    float(d.get("stim_period_S2", 0.0)), int(d.get("nstim2", 0)))
        check(lib().cg_ionic_set_protocols(self.h, C.cast(arr, C.c_void_p)))

    # numpy layout of cg_stimulus (4 doubles, int, double, int; natural alignment)
    PROTOCOL_DTYPE = np.dtype([("stim_start", "f8"), ("stim_period_S1", "f8"),
                               ("stim_duration", "f8"), ("stim_amplitude", "f8"),
                               ("nstim1", "i4"), ("stim_period_S2", "f8"), ("nstim2", "i4")],
                              align=True)

    def set_protocol_arrays(self, stim_start, stim_period_S1, stim_duration, stim_amplitude,
                            nstim1, stim_period_S2=0.0, nstim2=0):
        """Per-cell protocols from arrays (or scalars broadcast to every cell)."""
        a = np.zeros(self.n, dtype=self.PROTOCOL_DTYPE)
        a["stim_start"], a["stim_period_S1"] = stim_start, stim_period_S1
        a["stim_duration"], a["stim_amplitude"] = stim_duration, stim_amplitude
        a["nstim1"], a["stim_period_S2"], a["nstim2"] = nstim1, stim_period_S2, nstim2
        assert a.itemsize == C.sizeof(_capi.Stimulus)
        check(lib().cg_ionic_set_protocols(self.h, a.ctypes.data_as(C.c_void_p)))

    def set_run_lengths(self, steps=None, n_steps_all=0):
        if steps is None:
            check(lib().cg_ionic_set_run_lengths(self.h, None, int(n_steps_all)))
            return
        a = np.ascontiguousarray(steps, dtype=np.int64)
        if a.size != self.n:
            raise ValueError("need one run length per cell")
        check(lib().cg_ionic_set_run_lengths(self.h, a.ctypes.data_as(C.c_void_p), 0))

    def enable_beats(self, max_beats=32, dvdt_threshold=5.0, baseline_s=2e-3, peak_window_s=2.0):
        """On-device detect_beats (postProcessing_restCurves.py:115-200); threshold in V/s."""
        check(lib().cg_ionic_enable_beats(self.h, int(max_beats), float(dvdt_threshold),
                                          float(baseline_s), float(peak_window_s)))
        self.max_beats = int(max_beats)

    def restitution(self):
        """classify_early_strokes + compute_apd_di (postProcessing_restCurves.py:203-297) of
        every cell's beats from the last batch run, on the device: (apd[n], di[n] in s, NaN
        where the pair is rejected, flags[n, max_beats]: 1 early, 2 valid APD, 4 valid DI, 8
        more beats than max_beats)."""
        mb = getattr(self, "max_beats", 0)
        out = np.zeros((self.n, 2))
        flags = np.zeros((self.n, max(mb, 1)), dtype=np.uint8)
        check(lib().cg_ionic_get_restitution(self.h, out.ctypes.data_as(C.c_void_p),
                                             flags.ctypes.data_as(C.c_void_p)))
        return out[:, 0], out[:, 1], flags[:, :mb]

    def beats(self):
        """(counts[n], records[n, max_beats, 5]) of the last batch run: t_up (s), Vm base,
        peak, V90 (V), t_repol90 (s, NaN if not reached); unused records NaN."""
        mb = getattr(self, "max_beats", 0)
        cnt = np.zeros(self.n, dtype=np.int32)
        rec = np.zeros((self.n, mb, 5))
        check(lib().cg_ionic_get_beats(self.h, cnt.ctypes.data_as(C.c_void_p),
                                       rec.ctypes.data_as(C.c_void_p)))
        return cnt, rec

    def stateVariableNames(self):
        buf = C.create_string_buffer(64)
        names = []
        for i in range(self.nstates):
            check(lib().cg_model_state_name(self.model, i, buf, 64))
            names.append(buf.value.decode())
        return names

    def algebraic_names(self):
        """the model's ALGEBRAIC list in the reference's order where the device evaluates it
        (BuenoOrovio, AlievPanfilov), else []"""
        buf = C.create_string_buffer(64)
        out = []
        for i in range(max(lib().cg_model_nalgebraic(self.model), 0)):
            check(lib().cg_model_algebraic_name(self.model, i, buf, 64))
            out.append(buf.value.decode())
        return out

    def io_states(self):
        """STATES_ [nstates, n] as the reference holds them after a step (its ioStatesPtr,
        ionicModel.H:112-118): the device states, with a tissue cell's V the one its last
        step was fed ((Vm vmul + voff)/vscale, BuenoOrovio.C:119-122)"""
        S = self.states()
        vm = self._vm_fed() if (self._vm_fed is not None and not self.solve_vm) else None
        if vm is not None:
            S[0] = self.model_v(vm)
        return S

    def stateVariableValues(self, integrationPtI=0):
        """ionicModel.H:391-406: the states of one integration point"""
        if not 0 <= integrationPtI < self.n:
            raise FatalIOError("Requested integration point %d but valid range is [0, %d]."
                               % (integrationPtI, self.n - 1))
        return self.io_states()[:, integrationPtI].copy()

    def algebraicVariableNames(self):
        return self.algebraic_names()

    def vmRest(self):
        """ionicModel.H:305-316: cell 0's V in volts through the model's Vm transform (its
        initial state until the first step)"""
        _, voff, vscale, _ = self.units()
        return float(self.io_states()[0, 0] * vscale - voff) * 1e-3

    def _cai_index(self):
        # ionicVariableCompatibility.C:154-181 findCaiStateIndex
        from .foam_io import _canonical
        for i, nm in enumerate(self.stateVariableNames()):
            c = _canonical(nm)
            if c in ("cai", "calciumcai", "intracellularcai", "intracellularcalcium") or \
                    "cai" in c:
                return i
        return -1

    def hasSignal(self, s):
        """ionicModel.C:120-148 for the coupling signals "VM" (every device model: V is
        state 0 under the model's Vm transform) and "CAI" (a state named like Ca_i)"""
        if s == "VM":
            return True
        if s == "CAI":
            return self._cai_index() >= 0
        return False

    def signals(self, s):
        """ionicModel.C:150-204 signal(i, s) of every integration point: VM in mV, CAI in
        mM; a signal the model does not provide is a FatalError"""
        if not self.hasSignal(s):
            raise FatalIOError("Requested coupling signal %s from ionicModel, but this ionic "
                               "model does not provide it." % s)
        S = self.io_states()
        if s == "VM":
            _, voff, vscale, _ = self.units()
            return S[0] * vscale - voff
        return S[self._cai_index()]

    def signal(self, i, s):
        if not 0 <= i < self.n:
            raise FatalIOError("Requested coupling signal index i=%d but valid "
                               "integration-point range is [0, %d]." % (i, self.n - 1))
        return float(self.signals(s)[i])

    def units(self):
        """(vmul, voff, vscale, t_per_s) of the model (cg_model_units): the V state fed
        from Vm is (Vm vmul + voff)/vscale, the trace's Vm in mV is V vscale - voff, and
        model time is t_s t_per_s"""
        return model_units(self.model)

    def model_time(self, t_s):
        return np.asarray(t_s, dtype=np.float64) * self.units()[3]

    def model_v(self, Vm):
        """the V state solveODE feeds a tissue cell from Vm in V (BuenoOrovio.C:119-122)"""
        vmul, voff, vscale, _ = self.units()
        return (np.asarray(Vm, dtype=np.float64) * vmul + voff) / vscale

    def eval(self, kind, states, t_model=None):
        """RATES ("rates") or ALGEBRAIC ("algebraic") of the model at rows of states
        [nstates, n] (model units, V as state 0) and model times t_model [n]
        (cg_ionic_eval, on the device): [nvars, n]."""
        k = {"rates": _capi.RATES, "algebraic": _capi.ALGEBRAIC}[kind]
        st = np.ascontiguousarray(states, dtype=np.float64)
        if st.ndim != 2 or st.shape[0] != self.nstates:
            raise ValueError("states must be [nstates, n]")
        n = st.shape[1]
        t = None if t_model is None else np.ascontiguousarray(t_model, dtype=np.float64)
        nv = self.nstates if k == _capi.RATES else max(lib().cg_model_nalgebraic(self.model), 0)
        out = np.zeros((max(nv, 1), n))
        check(lib().cg_ionic_eval(self.h, k, n, st.ctypes.data_as(C.c_void_p),
                                  None if t is None else t.ctypes.data_as(C.c_void_p),
                                  out.ctypes.data_as(C.c_void_p)))
        return out[:nv]


def model_units(model):
    """cg_model_units of a model id or key: (vmul, voff, vscale, t_per_s)"""
    L = lib()
    if isinstance(model, str):
        model = L.cg_model_id(model.encode())
    out = [C.c_double() for _ in range(4)]
    check(L.cg_model_units(int(model), *[C.byref(x) for x in out]))
    return tuple(x.value for x in out)


def _handles(models):
    arr = (C.c_void_p * len(models))()
    for i, m in enumerate(models):
        arr[i] = m.h
    return arr


def batch_run(models, dt, chunk=20000):
    """Runs every cell of every model for its run length in one hipGraph launch."""
    arr = _handles(models)
    check(lib().cg_batch_run(arr, len(models), float(dt), int(chunk)))


def batch_autotune(models, dt, trial_steps=400):
    """Picks the fastest kernel shape per model on a short trial; returns the report."""
    arr = _handles(models)
    buf = C.create_string_buffer(4096)
    check(lib().cg_batch_autotune(arr, len(models), float(dt), int(trial_steps), buf, 4096))
    return buf.value.decode()
