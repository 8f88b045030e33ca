"""MI355X (gfx950) DG-Euler explicit time step: HIP kernels behind a C ABI (include/dgm.h).

The drop-in surface mirrors the Rust crate Decem5150/DGM_in_rust: :class:`SpatialDisc`,
:class:`TemperalDisc`, :class:`Solver` (see :mod:`dgm_in_rust_amd.solver`). Meshes come from
:mod:`dgm_in_rust_amd.meshgen` (Gmsh 4.1 reader, NACA 0012 O-grid, periodic square).
"""
from .meshgen import MeshInput, naca0012_ogrid, periodic_square, read_gmsh, write_gmsh  # noqa: F401
from .solver import Context, Solver, SpatialDisc, TemperalDisc  # noqa: F401
from ._lib import DgmError  # noqa: F401
from .batch import Batch  # noqa: F401
