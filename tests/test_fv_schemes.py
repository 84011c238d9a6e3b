"""system/fvSchemes on the host mirror (no GPU): every tissue tutorial's schemes parse into
the device path's selection, unsupported schemes are refused instead of replaced, and the
least-squares gradient the tutorials select equals Gauss linear on the hex block (the
oracle's restatement of leastSquaresVectors / leastSquaresGrad)."""
import json
import os

import numpy as np
import pytest

import cardiacfoam_amd as cf
from cardiacfoam_amd.tissue import fv_schemes_selection
from tests.golden import make_fv_schemes

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "fv_schemes.json")


def _fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def test_fixture_matches_reference_files():
    """tests/golden/fv_schemes.json is the parse of the reference's case files (checked
    where the reference tree is present)."""
    ref = "/root/reference"
    if not os.path.isdir(os.path.join(ref, "tutorials")):
        pytest.skip("reference tree not present")
    assert make_fv_schemes.read(ref) == _fixture()


def test_every_tutorial_selects_a_device_scheme():
    got = {case: fv_schemes_selection(d) for case, d in _fixture().items()}
    for case, sel in got.items():
        assert sel["ddt"] == "backward", case           # every tissue tutorial
        assert sel["laplacian"] == "Gauss linear corrected", case
        assert sel["grad"] in ("leastSquares", "Gauss linear"), case
    assert got["regressionTests/NiedererEtAl2012"]["grad"] == "leastSquares"
    # ddt(Vm) backward with default none
    assert got["manufacturedSolutions/monodomainPseudoECG"]["ddt"] == "backward"


def test_refusals():
    base = _fixture()["regressionTests/NiedererEtAl2012"]
    for ddt in ("CrankNicolson 0.9", "steadyState", "bounded Euler"):
        d = dict(base, ddtSchemes={"default": ddt.split() if " " in ddt else ddt})
        with pytest.raises(cf.FatalIOError):
            fv_schemes_selection(d)
    with pytest.raises(cf.FatalIOError):           # default none and no ddt(Vm)
        fv_schemes_selection(dict(base, ddtSchemes={"default": "none"}))
    with pytest.raises(cf.FatalIOError):           # leastSquares on a general mesh
        fv_schemes_selection(base, general_mesh=True)
    with pytest.raises(cf.FatalIOError):
        fv_schemes_selection(dict(base, laplacianSchemes={"default": ["Gauss", "harmonic",
                                                                      "corrected"]}))
    assert fv_schemes_selection(dict(base, gradSchemes={"default": ["Gauss", "linear"]}),
                                general_mesh=True)["grad"] == "Gauss linear"
    d = dict(base, ddtSchemes={"default": "backward", "ddt(Vm)": "Euler"})
    assert fv_schemes_selection(d)["ddt"] == "Euler"


def test_monodomain_reads_case_schemes(tmp_path):
    """Monodomain(case_dir=...) takes ddt from system/fvSchemes; a case without the file
    is refused (no silent default); the non-reference Coeffs key is gone."""
    sysdir = tmp_path / "system"
    sysdir.mkdir()
    (sysdir / "fvSchemes").write_text(
        "ddtSchemes { default CrankNicolson 0.9; }\ngradSchemes { default leastSquares; }\n"
        "laplacianSchemes { default Gauss linear corrected; }\n"
        "interpolationSchemes { default linear; }\n")
    coeffs = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells", "chi": 1.4e5,
              "cm": 0.01, "solutionAlgorithm": "explicit", "conductivity": [0.1] * 9}
    with pytest.raises(cf.FatalIOError, match="CrankNicolson"):
        cf.Monodomain(cf.HexMesh(4, 3, 2, 1e-4), coeffs, dt=1e-6, case_dir=str(tmp_path))
    with pytest.raises(cf.FatalIOError, match="fvSchemes"):
        cf.Monodomain(cf.HexMesh(4, 3, 2, 1e-4), coeffs, dt=1e-6,
                      case_dir=str(tmp_path / "nocase"))


def test_least_squares_equals_gauss_on_the_hex_block():
    """OpenFOAM's leastSquares gradient with zeroGradient walls equals Gauss linear on a
    uniform block (so the hex path's tensor correction is the tutorials' scheme), and
    differs on a skewed mesh, where only leastSquares is exact for linear fields in
    interior cells."""
    from cardiacfoam_amd import meshes
    from oracle import general as og
    L = (1.4e-3, 1e-3, 0.8e-3)
    rng = np.random.default_rng(0)
    m = meshes.block(7, 5, 4, lengths=L)
    V = rng.normal(size=m["n"])
    g, ls = og.grad(m, V), og.ls_grad(m, V)
    assert np.abs(g - ls).max() <= 1e-13 * np.abs(g).max()
    w = meshes.block(7, 5, 4, lengths=L, warp=meshes.smooth_warp(0.05, L))
    X = w["C"]
    Vl = 3 * X[:, 0] - 2 * X[:, 1] + 0.5 * X[:, 2]
    interior = np.ones(w["n"], bool)
    interior[w["b_owner"]] = False
    assert np.abs(og.ls_grad(w, Vl)[interior] - [3, -2, 0.5]).max() < 1e-12
    assert np.abs(og.grad(w, Vl)[interior] - [3, -2, 0.5]).max() > 1e-2
