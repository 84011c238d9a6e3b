"""Writes tests/golden/fv_schemes.json: the ddt/grad/laplacian/interpolation entries of
every tutorial fvSchemes that drives the electro (monodomain) equation, parsed from the
reference's case files (data only).  Run in a container that holds the reference:
    python tests/golden/make_fv_schemes.py /root/reference
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cardiacfoam_amd import foam_dict  # noqa: E402

# tutorials whose electroProperties select a myocardium solver; the electro-mechanical
# cases keep the electro region's schemes under system/electro/
CASES = [
    "HeartSimTemplate",
    "Niederer/NiedererEtAl2011MonodomainPurkinje",
    "Niederer/NiedererEtAl2011verification",
    "Niederer/electroMechanicalNiedererEtAl2011/system/electro",
    "electroMechanicalNiedererEtAl2011/system/electro",
    "manufacturedSolutions/monodomainPseudoECG",
    "manufacturedSolutions/singleCellMonodomain",
    "regressionTests/NiedererEtAl2012",
    "regressionTests/singleCell",
    "singleCellprotocols/restitutionCurves_s1s2Protocol",
    "singleCellprotocols/singleCell",
]
KEYS = ("ddtSchemes", "gradSchemes", "laplacianSchemes", "interpolationSchemes")


def read(ref):
    out = {}
    for case in CASES:
        d = os.path.join(ref, "tutorials", case)
        path = os.path.join(d, "fvSchemes") if d.endswith("electro") else \
            os.path.join(d, "system", "fvSchemes")
        parsed = foam_dict.read(path)
        out[case] = {k: parsed.get(k) for k in KEYS}
    return out


if __name__ == "__main__":
    data = read(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    with open(os.path.join(HERE, "fv_schemes.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
