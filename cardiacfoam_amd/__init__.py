"""cardiacfoam_amd -- MI355X (gfx950) implementation of cardiacFoam's monodomain tissue
step: per-cell ionic update (BuenoOrovio, TNNP, ORd, AlievPanfilov, Courtemanche,
PerisYague_2022, Stewart, Grandi, ToRORd_dynCl, Gaur) and transmembrane
diffusion (explicit face-flux update, or implicit assembly + Jacobi-PCG), behind the
C-ABI in include/cardiacgpu.h.

The compute runs only in libcardiacgpu.so; this package is the host-side mirror of the
reference's ionicModel / MyocardiumDomain interface.
"""
from . import _capi
from .foam_dict import parse as parse_dictionary, read as read_dictionary
from .ionic import IonicModel, FatalIOError, batch_run, batch_autotune
from .tissue import HexMesh, PolyMesh, Monodomain
from .adjoint import TissueAdjoint, probe_traces
from .ldu import LduMatrix

__all__ = ["LduMatrix", "IonicModel", "FatalIOError", "batch_run", "batch_autotune", "HexMesh", "PolyMesh", "Monodomain", "TissueAdjoint", "probe_traces", "parse_dictionary",
           "read_dictionary", "library_path", "version"]


def library_path():
    return _capi.LIB_PATH


def version():
    return _capi.lib().cg_version().decode()
