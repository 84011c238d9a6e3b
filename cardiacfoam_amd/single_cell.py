"""Host mirror of SingleCellSolver (src/electroModels/myocardiumModels/singleCellSolver/
singleCellSolver.C): one integration point with Vm inside the ODE, advanced on the device,
and the trace file postProcessing/<ionicModel>_<tissueName>_<protocolSuffix>.txt that the
reference writes every step from writeAfterTime on (singleCellSolver.C:98-120, 179-229).
"""
import numpy as np

from .foam_io import SingleCellTrace
from .ionic import IonicModel


class SingleCellSolver:
    def __init__(self, coeffs, delta_t, case_dir, precision=None):
        """coeffs: the singleCellSolverCoeffs dictionary; delta_t: controlDict deltaT in
        s, also the initial ODE step (singleCellSolver.C:48-57 passes runTime.deltaT)."""
        self.coeffs = coeffs
        self.dt = float(delta_t)
        self.ionic = IonicModel.New(coeffs, 1, self.dt, True, precision=precision)
        self.write_after = float(coeffs.get("writeAfterTime", 0.0))
        self.trace = SingleCellTrace(case_dir, self.ionic.model_key, self.ionic.tissue_name,
                                     coeffs, self.ionic.stateVariableNames(),
                                     self.ionic.algebraic_names())
        self.t = 0.0
        self.steps = 0

    def run(self, end_time, chunk=8192):
        """Time loop to endTime (Time::run: while t < endTime - deltaT/2; t += deltaT),
        one evolve() per step; rows of steps ending at or after writeAfterTime are
        appended to the trace (ionicModelIO::shouldWriteStep)."""
        while self.t < end_time - 0.5 * self.dt:
            k = 0
            ts = []
            t = self.t
            while k < chunk and t < end_time - 0.5 * self.dt:
                t = t + self.dt
                ts.append(t)
                k += 1
            states, rates = self.ionic.run_cells_trace(self.t, self.dt, k, 0)
            ts = np.asarray(ts)
            keep = ts >= self.write_after
            if keep.any():
                alg = None
                if self.trace.needs_alg:
                    # computeVariables at each step's end (its stimulus included), on the
                    # device from the recorded states
                    alg = self.ionic.eval("algebraic", states[keep].T,
                                          self.ionic.model_time(ts[keep])).T
                self.trace.append(ts[keep], states[keep], rates[keep], alg)
            self.t = t
            self.steps += k
        return self.trace.path
