"""Golden vectors and case set-ups taken from the reference's regression tests (data
only: inputs, expected outputs and tolerances).

SINGLE_CELL: tutorials/regressionTests/singleCell/constant/electroProperties,
             system/controlDict (deltaT 5e-5, endTime 1) and singleCell.reference:2-6.
NIEDERER:    tutorials/regressionTests/NiedererEtAl2012/constant/electroProperties,
             system/blockMeshDict (100 15 35 over 20x3x7 mm), system/controlDict
             (endTime 0.04, maxCo 0.1), system/fvSchemes (backward), system/Niedererpoints,
             NiedererEtAl2012.reference:9-14 (expected, tolerance) and
             log.Alltest-regression:6-11 (values of the passing run).
"""

SINGLE_CELL = dict(
    # singleCellSolverCoeffs of the case (ionicModel, tissue, integrator, stimulus, the
    # outputVariables export list) and the trace file singleCell.reference:2-6 names
    coeffs={"ionicModel": "BuenoOrovio", "tissue": "endocardialCells",
            "solutionAlgorithm": "explicit", "solver": "RKF45", "initialODEStep": 1e-5,
            "maxSteps": 1000000000, "absTol": 1e-6, "relTol": 1e-4,
            "singleCellStimulus": {"stim_start": 20, "stim_duration": 1,
                                   "stim_amplitude": 0.4, "stim_period_S1": 1000,
                                   "nstim1": 3, "stim_period_S2": 0, "nstim2": 0},
            "writeAfterTime": 0, "outputVariables": {"ionic": {"export": ["Vm", "s"],
                                                              "debug": ["Vm"]}}},
    trace_file="BuenoOrovio_endocardialCells_S1_1000.txt",
    stimulus=dict(start=20.0, period_s1=1000.0, duration=1.0, amplitude=0.4, nstim1=3,
                  period_s2=0.0, nstim2=0),
    dt=5e-5, n_steps=20000, abs_tol=1e-6, rel_tol=1e-4,
    # (time s, Vm mV, tolerance mV)
    reference=[(0.0200000, -84.0005041, 5e-3), (0.0215000, 42.3123891, 5e-3),
               (0.0220000, 44.8863342, 5e-3), (0.0300000, 32.7821149, 5e-3),
               (0.5000000, -83.6520453, 5e-3)],
)

# tutorials/manufacturedSolutions/singleCellBidomain/postProcessing/
# AlievPanfilov_myocyte_S1_1000.txt, the reference's trace of an AlievPanfilov myocyte
# (rows kept in aliev_panfilov_trace.txt).  The case's electroProperties now select
# another model; the trace's own name and rows fix the rest: the run's integrator
# controls are the case's (RKF45, absTol 1e-6, relTol 1e-4, deltaT 5e-5 s, endTime 1),
# the initial state relaxes from -75 mV (u = 0.05) to rest, and the upstroke starts at
# 0.258 s = 20 model time units of 12.9 ms with dV/dt = 3014 mV/s, i.e. an S1 stimulus
# of amplitude 0.4 from model time 20 (period 1000 from the file name, duration 1 as
# the singleCell tutorials use).
ALIEV_PANFILOV = dict(
    coeffs={"ionicModel": "AlievPanfilov", "tissue": "myocyte", "solutionAlgorithm": "explicit",
            "solver": "RKF45", "initialODEStep": 1e-5, "maxSteps": 1000000000,
            "absTol": 1e-6, "relTol": 1e-4,
            "singleCellStimulus": {"stim_start": 20, "stim_duration": 1, "stim_amplitude": 0.4,
                                   "stim_period_S1": 1000, "nstim1": 1, "stim_period_S2": 0,
                                   "nstim2": 0},
            "writeAfterTime": 0, "outputVariables": {"ionic": {"export": ["Vm"]}}},
    stimulus=dict(start=20.0, period_s1=1000.0, duration=1.0, amplitude=0.4, nstim1=1,
                  period_s2=0.0, nstim2=0),
    trace_file="AlievPanfilov_myocyte_S1_1000.txt",
    fixture="aliev_panfilov_trace.txt",
    dt=5e-5, n_steps=20000, abs_tol=1e-6, rel_tol=1e-4, upstroke=0.258,
)

NIEDERER = dict(
    shape=(100, 15, 35), h=0.2e-3,
    sigma9=[0.1334177215, 0, 0, 0, 0.01760617761, 0, 0, 0, 0.01760617761],
    chi=140000.0, cm=0.01, max_co=0.1, end_time=0.04,
    box=dict(min=[0.0, 0.0, 5.5e-3], max=[1.5e-3, 1.5e-3, 7e-3], t_start=0.0,
             duration=2e-3, intensity=50000.0),
    # probe point -> cell (i, j, k) containing it; probe 8 sits on the face between
    # i = 49 and 50 and goes to the lower cell, as polyMesh::findCell's first match.
    # The stimulus box edges (x = y = 1.5 mm, z = 5.5 mm) lie on cell-centre planes: the
    # stimulated set is the cells whose OpenFOAM-computed centre the inclusive box holds
    # (block_mesh_dict, boundary patch order below).
    probes=[
        dict(name="p0", point=(0.0, 0.0, 0.007), cell=(0, 0, 34), expected=0.00120503, tol=1e-5),
        dict(name="p2", point=(0.019999, 0.0, 0.007), cell=(99, 0, 34), expected=0.03271, tol=1e-4),
        dict(name="p4", point=(0.0, 0.003, 0.007), cell=(0, 14, 34), expected=0.0105757, tol=1e-4),
        dict(name="p8", point=(0.01, 0.0015, 0.0035), cell=(49, 7, 17), expected=0.0218433, tol=1e-4),
        dict(name="p3", point=(0.019999, 0.0, 0.0), cell=(99, 0, 0), expected=0.0, tol=1e-6),
        dict(name="p7", point=(0.019999, 0.003, 0.0), cell=(99, 14, 0), expected=0.0, tol=1e-6),
    ],
    # log.Alltest-regression:6-9: the values of the reference's own passing run, printed
    # to the digits below (a restatement that reproduces them is within half a unit of
    # the last digit)
    passing_log={"p0": 0.0012048, "p2": 0.0327112, "p4": 0.0105765, "p8": 0.0218436},
    passing_log_halfunit={"p0": 5e-8, "p2": 5e-8, "p4": 5e-8, "p8": 5e-8},
    # system/blockMeshDict: vertices (0 0 0)..(20 3 7), scale 0.001, one patch `walls`
    # listing the block faces (3 7 6 2) (0 4 7 3) (2 6 5 1) (1 5 4 0) (0 3 2 1) (4 5 6 7)
    block_mesh_dict=dict(lo=(0.0, 0.0, 0.0), hi=(20.0, 3.0, 7.0), scale=0.001,
                         patches=(("ymax", "xmin", "xmax", "ymin", "zmin", "zmax"),)),
    # NiedererEtAl2012.reference:9-14 as (file, time, awk column, expected, tolerance);
    # probes of system/Niedererpoints in their listed order, column 2 = probe 0
    probe_file="Niedererpoints/0/activationTime",
    probe_locations=[(0.0, 0.0, 0.007), (0.0, 0.0, 0.0), (0.019999, 0.0, 0.007),
                     (0.019999, 0.0, 0.0), (0.0, 0.003, 0.007), (0.0, 0.003, 0.0),
                     (0.019999, 0.003, 0.007), (0.019999, 0.003, 0.0), (0.01, 0.0015, 0.0035)],
    probe_rows=[("Niedererpoints/0/activationTime", 0.04, 2, 0.00120503, 1e-5),
                ("Niedererpoints/0/activationTime", 0.04, 4, 0.03271, 1e-4),
                ("Niedererpoints/0/activationTime", 0.04, 6, 0.0105757, 1e-4),
                ("Niedererpoints/0/activationTime", 0.04, 10, 0.0218433, 1e-4),
                ("Niedererpoints/0/activationTime", 0.04, 5, 0.0, 1e-6),
                ("Niedererpoints/0/activationTime", 0.04, 9, 0.0, 1e-6)],
    write_interval=5e-3,
)
