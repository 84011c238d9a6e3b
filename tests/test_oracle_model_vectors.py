"""The oracle's model right-hand sides against the reference's own generated expressions
(tests/golden/model_vectors.json, written by tests/golden/make_model_vectors.py from the
models' headers: initConsts and the function each model's derivatives() calls, at seeded
states along an action potential).

Pins every model the oracle restates -- BuenoOrovio, TNNP (3 tissues), ORd, AlievPanfilov,
Courtemanche, PerisYague_2022, Stewart, Grandi, ToRORd_dynCl (3 tissues) and Gaur -- to
the reference's text: the initial states exactly, and RATES in tissue mode (V frozen) and in single-cell mode (dV/dt = -Iion) to
1e-13 of each rate's range over the samples (the same libm; the two differ only in the
order of a few operations).  ORd carries the documented substitution of the published
values for the 27 constants its initConsts leaves unset, in both.  A constant off by one
part in 1e6 is seen by the vectors."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "model_vectors.json")
IDS = {"BuenoOrovio": 0, "TNNP": 1, "ORd": 2, "AlievPanfilov": 3, "Courtemanche": 5,
       "PerisYague_2022": 6, "Stewart": 7, "Grandi": 8, "ToRORd_dynCl": 9, "Gaur": 10}
TOL = 1e-13


def _fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def _cases():
    d = _fixture()
    return [(key, int(t)) for key, blk in d.items() for t in blk if t != "state_names"]


def _rates(orc, m, tis, c, S, solve_vm):
    return np.array([orc.model_eval(m, tis, 0.0, c, s, solve_vm=solve_vm)["rates"] for s in S])


def test_fixture_matches_reference_files():
    """tests/golden/model_vectors.json is what the generator makes from the reference's
    headers (checked where the reference tree is present)."""
    ref = "/root/reference"
    if not os.path.isdir(os.path.join(ref, "src", "ionicModels")):
        pytest.skip("reference tree not present")
    from tests.golden import make_model_vectors
    assert make_model_vectors.make(ref) == _fixture()


def test_state_names_follow_the_reference():
    """the STATES enums of the models' Names.H, in order, are the library's state names"""
    import ctypes as C
    from cardiacfoam_amd import _capi
    L = _capi.lib()
    buf = C.create_string_buffer(64)
    for key, blk in _fixture().items():
        m = L.cg_model_id(key.encode())
        got = []
        for i in range(L.cg_model_nstates(m)):
            _capi.check(L.cg_model_state_name(m, i, buf, 64))
            got.append(buf.value.decode())
        assert got == blk["state_names"], key


@pytest.mark.parametrize("key,tissue", _cases())
def test_oracle_against_reference_expressions(orc, key, tissue):
    v = _fixture()[key][str(tissue)]
    m = IDS[key]
    c, y0 = orc.model_init(m, tissue)
    np.testing.assert_array_equal(y0, v["init_states"])
    S = np.array(v["states"])
    for mode, solve_vm in (("rates_tissue", False), ("rates_cell", True)):
        want = np.array(v[mode])
        got = _rates(orc, m, tissue, c, S, solve_vm)
        scale = np.max(np.abs(want), axis=0) + 1e-300
        err = np.max(np.abs(got - want) / scale)
        assert err <= TOL, (key, tissue, mode, err)
    # dV/dt: 0 with V frozen, -Iion inside the ODE
    assert np.all(np.array(v["rates_tissue"])[:, 0] == 0.0)


@pytest.mark.parametrize("key", ["TNNP", "ORd", "Courtemanche", "PerisYague_2022", "Stewart",
                                 "Grandi", "ToRORd_dynCl", "Gaur"])
def test_vectors_see_a_constant_off_by_one_part_in_a_million(orc, key):
    """Planted in the oracle's constants on the CPU: each constant that some rate depends
    on at these states, scaled by 1 + 1e-6, moves a rate beyond the pin's tolerance."""
    d = _fixture()[key]
    tissue = int(next(t for t in d if t != "state_names"))
    m = IDS[key]
    c, _ = orc.model_init(m, tissue)
    S = np.array(d[str(tissue)]["states"])
    base = _rates(orc, m, tissue, c, S, True)
    scale = np.max(np.abs(base), axis=0) + 1e-300
    seen, used = 0, 0
    for k in np.nonzero(c)[0]:
        cc = c.copy()
        cc[k] *= 1.0 + 1e-6
        err = np.max(np.abs(_rates(orc, m, tissue, cc, S, True) - base) / scale)
        if err > 0.0:
            used += 1
            seen += err > 10 * TOL
    assert used > 0 and seen == used, (key, seen, used)
