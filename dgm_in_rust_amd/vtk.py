"""Result output: the legacy-ASCII VTK file of write_to_vtk (src/io/write_results.rs:8-92).

The file holds the mesh vertices (z = 0), the triangles in element order, and three cell fields taken
from mode 0 of each element (the cell mean of the Dubiner basis): density, velocity (ρu/ρ, ρv/ρ, 0)
and pressure (γ − 1)(E − ½ρ(u² + v²)) (:26-73). The layout is the one the reference's writer
produces: numbers in shortest round-trip decimal without an exponent ("20", "0.05", "NaN"), each
section on one line, the blank lines between sections (`outputfiles/mesh.vtk`, byte for byte in
`tests/test_vtk.py`). The reference writes `outputfiles/solutions/step=N.vtk` at the end of
`time_march` (src/temporal_disc.rs:56); `write_solution` does the same on request.
"""
import os
from typing import Optional

import numpy as np


def format_f64(x: float) -> str:
    """A float as the reference prints it: the shortest digits that read back to the same double,
    positional, without a trailing ".0"; NaN as "NaN", infinities as "inf" / "-inf"."""
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def cell_fields(U: np.ndarray, hcr: float):
    """Density, velocity [ne][3] and pressure from mode 0 of U [ne][4][Np], in the reference's
    operation order (src/io/write_results.rs:28-52)."""
    rho = np.ascontiguousarray(U[:, 0, 0], dtype=np.float64)
    ru = np.ascontiguousarray(U[:, 1, 0], dtype=np.float64)
    rv = np.ascontiguousarray(U[:, 2, 0], dtype=np.float64)
    en = np.ascontiguousarray(U[:, 3, 0], dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        u = ru / rho
        v = rv / rho
        p = (hcr - 1.0) * (en - 0.5 * rho * (u * u + v * v))
    vel = np.zeros((len(rho), 3))
    vel[:, 0] = u
    vel[:, 1] = v
    return rho, vel, p


def vtk_text(vxy: np.ndarray, element_vertices: np.ndarray, U: np.ndarray, hcr: float) -> str:
    """The whole file as text."""
    vxy = np.asarray(vxy, dtype=np.float64)
    ev = np.asarray(element_vertices, dtype=np.int64)
    npt, nc = vxy.shape[0], ev.shape[0]
    if U.shape[0] != nc:
        raise ValueError(f"state has {U.shape[0]} elements, mesh has {nc}")
    pts = np.zeros((npt, 3))
    pts[:, :2] = vxy
    rho, vel, p = cell_fields(U, hcr)
    fmt = lambda a: " ".join(format_f64(x) for x in np.ravel(a))
    cells = " ".join(f"3 {a} {b} {c}" for a, b, c in ev.tolist())
    lines = ["# vtk DataFile Version 2.0", "", "ASCII", "", "DATASET UNSTRUCTURED_GRID",
             f"POINTS {npt} double", fmt(pts), "",
             f"CELLS {nc} {4 * nc}", cells, "",
             f"CELL_TYPES {nc}", *(["5"] * nc), "", "",
             f"POINT_DATA {npt}", "",
             f"CELL_DATA {nc}", "",
             "SCALARS Density double 1", "LOOKUP_TABLE default", fmt(rho), "",
             "VECTORS Velocity double", fmt(vel), "",
             "SCALARS Pressure double 1", "LOOKUP_TABLE default", fmt(p), "", ""]
    return "\n".join(lines)


def write_vtk(path: str, ctx, U: np.ndarray) -> None:
    """Writes the state U [ne][4][Np] of the context's mesh to path."""
    text = vtk_text(ctx.mesh.vxy, ctx.connectivity()["element_vertices"], U, ctx.params.hcr)
    with open(path, "w", newline="\n") as f:
        f.write(text)


def write_solution(ctx, U: np.ndarray, step: int, out_dir: Optional[str] = None) -> str:
    """write_to_vtk's file name and folder: <out_dir>/step=<step>.vtk, the folder created if missing
    (outputfiles/solutions by default, src/io/write_results.rs:82-91). Returns the path."""
    out_dir = out_dir or os.path.join("outputfiles", "solutions")
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"step={step}.vtk")
    write_vtk(path, ctx, U)
    return path
