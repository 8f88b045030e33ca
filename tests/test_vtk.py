"""Result output (CPU): the VTK writer against the reference's own output file.

Pins: outputfiles/mesh.vtk, which the reference's write_to_vtk (src/io/write_results.rs:8-92) wrote
for inputfiles/airfoil.msh with an all-NaN state; the product's writer gives the same file byte for
byte (size and SHA-256 in tests/golden/airfoil_mesh_vtk.json). The number format is checked against
what fixes it (shortest round-trip decimal, positional, Rust's NaN/inf spellings) and the cell
fields against the freestream state they come from.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from dgm_in_rust_amd import Context, vtk

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_airfoil_file_matches_reference_bytes(airfoil_mesh):
    ref = json.load(open(os.path.join(GOLD, "airfoil_mesh_vtk.json")))
    ctx = Context(airfoil_mesh, order=2, host_only=True)
    U = np.full(ctx.shape, np.nan)
    raw = vtk.vtk_text(airfoil_mesh.vxy, ctx.connectivity()["element_vertices"], U, 1.4).encode()
    assert len(raw) == ref["bytes"]
    assert hashlib.sha256(raw).hexdigest() == ref["sha256"]


@pytest.mark.parametrize("x,s", [(0.0, "0"), (-0.0, "-0"), (1.0, "1"), (-20.0, "-20"), (0.1, "0.1"),
                                 (0.141414, "0.141414"), (1.0e5, "100000"), (2.5e-7, "0.00000025"),
                                 (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf")])
def test_number_format(x, s):
    assert vtk.format_f64(x) == s


def test_number_format_round_trips():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000),
                         [1.0e300, -1.0e-300, 5e-324, np.finfo(np.float64).max]])
    for x in xs:
        s = vtk.format_f64(x)
        assert "e" not in s and "E" not in s
        assert float(s) == x
        # shortest: the same digits as Python's shortest repr
        digits = lambda t: t.lstrip("-").replace(".", "").strip("0")
        assert digits(s) == digits(repr(float(x)).split("e")[0])


def test_freestream_fields_and_file_name(airfoil_mesh, tmp_path):
    mach, aoa, rho, p = 0.5, 1.25, 1.0, 1.0e5
    ctx = Context(airfoil_mesh, order=2, host_only=True, mach=mach, aoa_deg=aoa)
    U = ctx.initial_solution(rho, aoa, mach, p)
    path = vtk.write_solution(ctx, U, 500, out_dir=str(tmp_path / "solutions"))
    assert os.path.basename(path) == "step=500.vtk"
    lines = open(path).read().split("\n")
    ne = ctx.ne

    def field(name, n):
        i = next(k for k, l in enumerate(lines) if l.startswith(name))
        k = i + 2 if lines[i + 1].startswith("LOOKUP_TABLE") else i + 1
        v = np.array([float(t) for t in lines[k].split()])
        assert len(v) == n
        return v

    d = field("SCALARS Density", ne)
    vel = field("VECTORS Velocity", 3 * ne).reshape(ne, 3)
    pr = field("SCALARS Pressure", ne)
    c = np.sqrt(1.4 * p / rho)
    a = np.deg2rad(aoa)
    assert np.allclose(d, rho, rtol=1e-13)
    assert np.allclose(vel[:, 0], mach * c * np.cos(a), rtol=1e-12)
    assert np.allclose(vel[:, 1], mach * c * np.sin(a), rtol=1e-10)
    assert np.all(vel[:, 2] == 0.0)
    assert np.allclose(pr, p, rtol=1e-11)
    # the fields come from mode 0 only: higher modes change nothing
    V = U.copy()
    V[:, :, 1:] += 7.0
    assert vtk.vtk_text(airfoil_mesh.vxy, ctx.connectivity()["element_vertices"], V, 1.4) == open(path).read()
