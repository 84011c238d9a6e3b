"""Polyhedral test meshes: the package's block generator (cardiacfoam_amd/meshes.py, with
OpenFOAM's primitiveMesh geometry) and its smooth warp."""
from cardiacfoam_amd.meshes import block, mesh_args, smooth_warp  # noqa: F401
