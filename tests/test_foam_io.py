"""On-disk formats on the CPU (host code of the library: number formatting and parsing
need no GPU): OpenFOAM ASCII field files, probe files, the single-cell trace, and the
regression scripts' comparison restated in tests/regression_check.py.  The number forms
are pinned by C's printf (%g, %f), which is what std::ostream produces."""
import math
import os

import numpy as np
import pytest

from tests import regression_check as rc
from tests.golden import cases


@pytest.fixture(scope="module")
def fio():
    from cardiacfoam_amd import foam_io
    return foam_io


def _special():
    return np.array([0.0, -0.0, 1.0, -0.084, 1e-5, 1.2e-5, 9.999995e-5, 123456.0, 123456.7,
                     1234567.0, 0.1 + 0.2, 42.3123891, -84.0005041, 5e-324, 1.7976931348623157e308,
                     2.5e-3, 1e-15, 33.0, 1e6])


def test_general_format_is_printf_g(fio):
    rng = np.random.default_rng(0)
    v = np.concatenate([_special(), rng.normal(size=3000) * 10.0 ** rng.integers(-12, 12, 3000)])
    for p in (6, 10, 17):
        got = fio.format_numbers(v, p)
        want = "".join("%.*g\n" % (p, x) for x in v).encode()
        assert got == want


def test_fixed_format_and_rows(fio):
    v = np.array([[0.02, -84.0005041, 0.9], [0.5, 42.31238912345, 1e-9], [1e7, -0.0, 3.25]])
    got = fio.format_numbers(v, 7, fixed=True, row=3)
    want = "".join(" ".join("%.7f" % x for x in r) + "\n" for r in v).encode()
    assert got == want


def test_large_list_uses_threads_and_keeps_order(fio):
    v = np.random.default_rng(1).normal(size=300_001)
    got = fio.format_numbers(v, 6)
    assert got.count(b"\n") == v.size
    back, _ = fio.parse_numbers(got, v.size)
    np.testing.assert_array_equal(back, np.array([float("%.6g" % x) for x in v]))


def test_parse_round_trip_17_digits_is_exact(fio):
    v = np.random.default_rng(2).normal(size=20000) * 1e-3
    back, end = fio.parse_numbers(fio.format_numbers(v, 17), v.size)
    assert np.array_equal(back, v)


def test_parse_lists_signs_and_errors(fio):
    from cardiacfoam_amd import FatalIOError  # noqa: F401
    text = b"(1 2 3)\n(+4 -5e-3 6E+02)\n)\n;"
    vals, end = fio.parse_numbers(text, 6)
    np.testing.assert_array_equal(vals, [1, 2, 3, 4, -5e-3, 600])
    assert text[end:end + 1] == b")"
    with pytest.raises(RuntimeError):
        fio.parse_numbers(b"1 2 x3", 3)
    with pytest.raises(RuntimeError):
        fio.parse_numbers(b"1 2", 3)


def test_scalar_field_round_trip_and_layout(fio, tmp_path):
    v = np.array([-0.084, 0.01, 1.2e-5, 42.0, -7.5])
    p = fio.write_field(str(tmp_path / "0.005" / "Vm"), v, location="0.005")
    text = open(p).read()
    assert "class       volScalarField;" in text and 'location    "0.005";' in text
    assert "dimensions      [1 2 -3 0 0 -1 0];" in text
    assert "internalField   nonuniform List<scalar> \n5\n(\n-0.084\n0.01\n1.2e-05\n42\n-7.5\n)\n;" in text
    assert "walls\n    {\n        type            zeroGradient;" in text
    f = fio.read_field(p)
    assert f.cls == "volScalarField" and f.name == "Vm" and not f.uniform
    assert f.dimensions == [1, 2, -3, 0, 0, -1, 0]
    np.testing.assert_array_equal(f.values, v)
    assert f.boundary["walls"]["type"] == "zeroGradient"


def test_uniform_and_tensor_fields(fio, tmp_path):
    p = fio.write_field(str(tmp_path / "0" / "activationTime"), np.zeros(7))
    assert "internalField   uniform 0;" in open(p).read()
    f = fio.read_field(p)
    assert f.uniform and np.array_equal(f.expand(7), np.zeros(7))
    sig = np.random.default_rng(3).normal(size=(6, 9))
    p = fio.write_field(str(tmp_path / "0" / "conductivity"), sig, precision=17)
    text = open(p).read()
    assert "class       volTensorField;" in text and "List<tensor> \n6\n(\n(" in text
    assert "dimensions      [-1 -3 3 0 0 2 0];" in text
    f = fio.read_field(p)
    assert f.values.shape == (6, 9) and np.array_equal(f.values, sig)
    p = fio.write_field(str(tmp_path / "0" / "f0"), np.tile([1.0, 0.0, 0.0], (4, 1)))
    f = fio.read_field(p)
    assert f.cls == "volVectorField" and f.uniform
    np.testing.assert_array_equal(f.expand(4), np.tile([1.0, 0.0, 0.0], (4, 1)))


def test_reads_an_openfoam_written_initial_field(fio, tmp_path):
    """The layout OpenFOAM's blockMesh tutorials ship for 0/Vm: banner comment, FoamFile
    header, uniform internal value, a zeroGradient patch."""
    text = """/*--------------------------------*- C++ -*----------------------------------*\\
| =========                 |                                                 |
\\*---------------------------------------------------------------------------*/
FoamFile
{
    version     2.0;
    format      ascii;
    class       volScalarField;
    location    "0";
    object      Vm;
}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

dimensions      [1 2 -3 0 0 -1 0];

internalField   uniform -0.084;

boundaryField
{
    walls
    {
        type            zeroGradient;
    }
}

// ************************************************************************* //
"""
    p = tmp_path / "Vm"
    p.write_text(text)
    f = fio.read_field(str(p))
    assert f.uniform and f.values == -0.084 and f.boundary["walls"]["type"] == "zeroGradient"
    q = tmp_path / "Vm2"
    q.write_text(text.replace("uniform -0.084;", "nonuniform List<scalar> 3(1 2.5 -3e-2);"))
    np.testing.assert_array_equal(fio.read_field(str(q)).values, [1, 2.5, -3e-2])
    r = tmp_path / "Vm3"
    r.write_text(text.replace("uniform -0.084;", "nonuniform List<scalar> 4(1 2.5 -3e-2);"))
    with pytest.raises(RuntimeError):
        fio.read_field(str(r))
    b = tmp_path / "Vm4"
    b.write_text(text.replace("format      ascii;", "format      binary;"))
    with pytest.raises(KeyError):
        fio.read_field(str(b))


def test_probe_cells_and_file(fio, tmp_path):
    c = cases.NIEDERER
    nx, ny, nz = c["shape"]
    pr = fio.Probes(c["shape"], c["h"], c["probe_locations"], ["activationTime"],
                    str(tmp_path / "postProcessing"), name="Niedererpoints")
    want = {p["point"]: p["cell"] for p in c["probes"]}
    for loc, cell in zip(c["probe_locations"], pr.cells):
        if tuple(loc) in want:
            i, j, k = want[tuple(loc)]
            assert cell == i + nx * (j + ny * k)
    field = np.arange(nx * ny * nz, dtype=np.float64) * 1e-6
    pr.write(0.005, {"activationTime": field})
    pr.write(0.04, {"activationTime": field * 2})
    lines = open(pr.paths["activationTime"]).read().splitlines()
    assert lines[0].startswith("# Probe 0 (0 0 0.007)") and lines[8].startswith("# Probe 8")
    assert lines[9].split()[:2] == ["#", "Probe"] and lines[10].split() == ["#", "Time"]
    row = lines[-1].split()
    assert float(row[0]) == 0.04 and len(row) == 10
    np.testing.assert_allclose([float(x) for x in row[1:]], field[pr.cells] * 2, rtol=1e-5)
    with pytest.raises(KeyError):
        fio.Probes(c["shape"], c["h"], [(0.03, 0, 0)], ["Vm"], str(tmp_path))


def test_protocol_suffix_and_trace_plan(fio):
    sc = cases.SINGLE_CELL["coeffs"]
    assert fio.protocol_suffix(sc) == "S1_1000"
    assert fio.protocol_suffix({}) == "noStim"
    assert fio.protocol_suffix({"singleCellStimulus": {"stim_period_S1": 1000, "nstim1": 0}}) \
        == "noStim"
    assert fio.protocol_suffix({"singleCellStimulus": {"stim_period_S1": 1000.5, "nstim1": 10,
                                                       "stim_period_S2": 350, "nstim2": 2}}) \
        == "S1_1000.5_S2_350"
    names = ["u", "v", "w", "s"]
    assert fio.trace_plan("BuenoOrovio", names, ["Vm", "s"]) == [("Vm", "vm", 0),
                                                                ("s", "state", 3)]
    assert fio.trace_plan("BuenoOrovio", names, ["VM", "RATES_w", "rates_S"]) == \
        [("VM", "vm", 0), ("RATES_w", "rate", 2), ("rates_S", "rate", 3)]
    full = fio.trace_plan("BuenoOrovio", names, None)
    assert [c[0] for c in full] == ["Vm", "v", "w", "s", "RATES_u", "RATES_v", "RATES_w",
                                    "RATES_s"]
    # without the model's algebraic list an unresolved name may be one: refused
    with pytest.raises(KeyError):
        fio.trace_plan("TNNP", ["V", "K_i"], ["V", "i_Na"])
    # with it, an unknown name is dropped with a warning (ionicModelIO.C:388-419)
    alg = ["Jfi", "Jso", "Jsi", "Jion", "Istim", "tauSo", "tauO", "tauVMinus", "tauWMinus",
           "tauS", "vInfty", "wInfty"]
    with pytest.warns(UserWarning, match="nothing"):
        plan = fio.trace_plan("BuenoOrovio", names, ["Vm", "Jsi", "nothing", "J_ion"], alg)
    assert plan == [("Vm", "vm", 0), ("Jsi", "alg", 2), ("J_ion", "alg", 3)]
    full = fio.trace_plan("BuenoOrovio", names, None, alg)
    assert [c[0] for c in full][4:16] == alg


def test_trace_file_layout_and_regression_check(fio, tmp_path):
    sc = cases.SINGLE_CELL
    tr = fio.SingleCellTrace(str(tmp_path), "BuenoOrovio", "endocardialCells", sc["coeffs"],
                             ["u", "v", "w", "s"])
    assert os.path.basename(tr.path) == sc["trace_file"]
    t = np.array([0.02, 0.0215, 0.022, 0.03, 0.5])
    # states whose Vm column (85.7 u - 84) is each golden value
    vm = np.array([r[1] for r in sc["reference"]])
    st = np.zeros((5, 4))
    st[:, 0] = (vm + 84.0) / 85.7
    st[:, 3] = 0.25
    tr.append(t, st, np.zeros((5, 4)))
    lines = open(tr.path).read().splitlines()
    assert lines[0] == "time Vm s"
    assert lines[1].split()[0] == "0.0200000" and lines[2].split()[2] == "0.2500000"
    rows = [(sc["trace_file"], tt, e, tol) for tt, e, tol in sc["reference"]]
    passed, failed = rc.check_single_cell(str(tmp_path / "postProcessing"), rows)
    assert len(passed) == 5 and not failed
    # a value off by more than the tolerance and a missing time both fail
    bad = [(sc["trace_file"], 0.0215, 42.3123891 + 6e-3, 5e-3),
           (sc["trace_file"], 0.0216, 0.0, 1.0)]
    passed, failed = rc.check_single_cell(str(tmp_path / "postProcessing"), bad)
    assert not passed and len(failed) == 2
    names, data = fio.read_trace(tr.path)
    assert names == ["time", "Vm", "s"] and data.shape == (5, 3)
    np.testing.assert_allclose(data[:, 1], vm, atol=5e-8)


def test_probe_regression_check(fio, tmp_path):
    c = cases.NIEDERER
    pr = fio.Probes(c["shape"], c["h"], c["probe_locations"], ["activationTime"],
                    str(tmp_path), name="Niedererpoints")
    field = np.zeros(int(np.prod(c["shape"])))
    expected = {2: 0.00120503, 4: 0.03271, 6: 0.0105757, 10: 0.0218433}
    for col, val in expected.items():
        field[pr.cells[col - 2]] = val
    pr.write(0.0399, {"activationTime": field})
    passed, failed = rc.check_probes(str(tmp_path), c["probe_rows"])
    assert len(passed) == 6 and not failed
    field[pr.cells[2]] += 2e-4
    pr.write(0.04, {"activationTime": field})
    passed, failed = rc.check_probes(str(tmp_path), c["probe_rows"])
    assert len(failed) == 1 and failed[0][2] == 4


def test_time_name(fio):
    assert [fio.time_name(t) for t in (0.0, 0.005, 0.04, 1e-5, 0.1 + 0.2, 12.5)] == \
        ["0", "0.005", "0.04", "1e-05", "0.3", "12.5"]
    assert math.isclose(float(fio.time_name(0.0400000001)), 0.04)


def test_pseudo_ecg_file_and_electrodes(fio, tmp_path):
    from cardiacfoam_amd import parse_dictionary
    d = parse_dictionary("""
        ECG { ecgSolver pseudoECG;
              electrodePositions { E1 (-0.5 0.5 0.5); E2 (1.5 0.5 0.5); E3 (1.2 0.23 0.61); } }
    """)
    names, pos = fio.electrodes(d["ECG"])
    assert names == ["E1", "E2", "E3"] and pos.shape == (3, 3) and pos[2, 1] == 0.23
    f = fio.PseudoEcgFile(str(tmp_path), names)
    f.write(0.01, [1.5e-3, -2.25e-5, 0.0])
    lines = open(f.path).read().splitlines()
    assert lines[0] == "# time  E1  E2  E3"
    assert lines[1] == "1.00000000e-02  1.50000000e-03  -2.25000000e-05  0.00000000e+00"
