"""Writes the fixtures under tests/golden/ from the reference's data files.

Run in the build container only (it reads the reference tree): copies the shipped input mesh
inputfiles/airfoil.msh (Gmsh 4.1, input data) and extracts from outputfiles/mesh.vtk the points
and triangle connectivity that the reference wrote after GmshParser + process_mesh
(src/io/write_results.rs:9-24), i.e. the element -> vertex order, as airfoil_vtk.npz.
The VTK cell data is all NaN. The size and SHA-256 of the whole file go to airfoil_mesh_vtk.json,
so that the product's VTK writer can be checked byte for byte against the reference's writer.
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    shutil.copyfile(os.path.join(REF, "inputfiles", "airfoil.msh"), os.path.join(HERE, "airfoil.msh"))
    raw = open(os.path.join(REF, "outputfiles", "mesh.vtk"), "rb").read()
    with open(os.path.join(HERE, "airfoil_mesh_vtk.json"), "w") as f:
        json.dump({"file": "outputfiles/mesh.vtk", "bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest()}, f,
                  indent=1)
        f.write("\n")
    lines = raw.decode().split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("POINTS"))
    n = int(lines[i].split()[1])
    vals = []
    k = i + 1
    while len(vals) < 3 * n:
        vals += [float(x) for x in lines[k].split()]
        k += 1
    pts = np.array(vals).reshape(n, 3)
    j = next(k for k, l in enumerate(lines) if l.startswith("CELLS"))
    nc, ntok = int(lines[j].split()[1]), int(lines[j].split()[2])
    tok = []
    k = j + 1
    while len(tok) < ntok:
        tok += [int(x) for x in lines[k].split()]
        k += 1
    cells, o = [], 0
    while o < ntok:
        assert tok[o] == 3
        cells.append(tok[o + 1:o + 4])
        o += 4
    assert len(cells) == nc
    np.savez_compressed(os.path.join(HERE, "airfoil_vtk.npz"), points=pts, cells=np.array(cells, np.int64))
    print("wrote", n, "points and", nc, "cells")


if __name__ == "__main__":
    sys.exit(main())
