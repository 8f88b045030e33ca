"""Mesh producers and the Gmsh 4.1 reader/writer.

Every producer returns a :class:`MeshInput`, the structure that ``GmshParser::parse``
(src/io/mesh_parser.rs:28-158) hands to ``process_mesh`` (src/initialization.rs:112-261):
vertices, boundary line blocks tagged Wall / Farfield, and triangle blocks with
counter-clockwise node order. ``process_mesh`` numbering therefore applies unchanged.

* :func:`naca0012_ogrid` — O-grid around a NACA 0012 section (4-digit thickness formula,
  cosine-clustered surface points, geometric radial growth, seed-chosen quad diagonals,
  seeded interior jitter), with a Wall block on the surface and a Farfield block outside.
* :func:`periodic_square` — doubly periodic jittered square (no boundary blocks). Vertices
  are identified across the period; ``elem_xy`` carries each triangle's unwrapped
  coordinates (the reference has no periodic boundary; this is an extension).
* :func:`read_gmsh` / :func:`write_gmsh` — Gmsh 4.1 ASCII in the subset the reference parser reads.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

WALL = 0
FARFIELD = 1


@dataclasses.dataclass
class MeshInput:
    vxy: np.ndarray                  # [nv, 2] float64
    eb_type: np.ndarray              # [neb] int32 (0 wall, 1 farfield)
    eb_count: np.ndarray             # [neb] int64
    eb_nodes: np.ndarray             # [sum eb_count, 2] int64, 0-based
    el_count: np.ndarray             # [nelb] int64
    el_nodes: np.ndarray             # [ne, 3] int64, 0-based, CCW
    elem_xy: Optional[np.ndarray] = None   # [ne, 3, 2] float64 (periodic meshes)
    name: str = "mesh"

    @property
    def n_elements(self) -> int:
        return int(self.el_nodes.shape[0])

    @property
    def n_vertices(self) -> int:
        return int(self.vxy.shape[0])

    def contiguous(self) -> "MeshInput":
        f = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
        return MeshInput(f(self.vxy, np.float64), f(self.eb_type, np.int32), f(self.eb_count, np.int64),
                         f(self.eb_nodes, np.int64).reshape(-1, 2), f(self.el_count, np.int64),
                         f(self.el_nodes, np.int64).reshape(-1, 3), f(self.elem_xy, np.float64), self.name)


def naca_thickness(x: np.ndarray, t: float = 0.12) -> np.ndarray:
    """NACA 4-digit half thickness y_t(x) with the standard (open trailing edge) coefficients."""
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1015 * x ** 4)


def naca0012_ogrid(n_surface: int = 160, n_layers: int = 50, growth: Optional[float] = None, r_far: float = 20.0,
                   jitter: float = 0.05, seed: int = 0, name: Optional[str] = None) -> MeshInput:
    """O-grid of 2 * n_surface * n_layers triangles around a unit-chord NACA 0012.

    Surface points are cosine-clustered in x (angle theta_i = 2 pi (i + 1/2) / n_surface,
    x = (1 + cos theta) / 2, upper surface for sin theta > 0); the far-field circle of radius
    ``r_far`` about the mid-chord uses the same angles. Layer j sits at the blend fraction
    s_j = (g^j - 1) / (g^n - 1) between surface and circle, so the first layer is
    r_far * (g - 1) / (g^n - 1) thick (0.0349 chords for n = 50, g = 1.08). When ``growth`` is
    None, g = 1.08^(50 / n_layers), i.e. the 50-layer grid refined uniformly in the radial index.
    Interior nodes move by up to ``jitter`` times the smaller local spacing in each coordinate;
    each quad takes a seed-chosen diagonal.
    """
    if n_surface < 8 or n_layers < 1:
        raise ValueError("n_surface >= 8 and n_layers >= 1 required")
    g = 1.08 ** (50.0 / n_layers) if growth is None else float(growth)
    rng = np.random.default_rng(seed)
    th = 2.0 * np.pi * (np.arange(n_surface) + 0.5) / n_surface
    xs = 0.5 * (1.0 + np.cos(th))
    ys = np.where(np.sin(th) > 0, 1.0, -1.0) * naca_thickness(xs)
    xf = 0.5 + r_far * np.cos(th)
    yf = r_far * np.sin(th)
    j = np.arange(n_layers + 1)
    s = (g ** j - 1.0) / (g ** n_layers - 1.0) if abs(g - 1.0) > 1e-14 else j / n_layers
    X = (1.0 - s)[:, None] * xs[None, :] + s[:, None] * xf[None, :]   # [n_layers+1, n_surface]
    Y = (1.0 - s)[:, None] * ys[None, :] + s[:, None] * yf[None, :]
    if jitter > 0.0 and n_layers >= 2:
        # local spacing: the smaller of the radial and circumferential neighbour distances
        dr = np.hypot(np.diff(X, axis=0), np.diff(Y, axis=0))            # [n_layers, ns]
        dc = np.hypot(X - np.roll(X, -1, axis=1), Y - np.roll(Y, -1, axis=1))
        h = np.minimum(np.minimum(dr[:-1], dr[1:]), np.minimum(dc[1:-1], np.roll(dc[1:-1], 1, axis=1)))
        jx = rng.uniform(-jitter, jitter, size=h.shape) * h
        jy = rng.uniform(-jitter, jitter, size=h.shape) * h
        X[1:-1] += jx
        Y[1:-1] += jy
    vxy = np.stack([X.ravel(), Y.ravel()], axis=1)
    vid = lambda jj, ii: jj * n_surface + (ii % n_surface)
    diag = rng.integers(0, 2, size=(n_layers, n_surface))
    tris: List[List[int]] = []
    for jj in range(n_layers):
        for ii in range(n_surface):
            # quad a (inner, theta_i) -> d (outer, theta_i) -> c (outer, theta_i+1) -> b (inner, theta_i+1) is CCW
            a, b = vid(jj, ii), vid(jj, ii + 1)
            c, d = vid(jj + 1, ii + 1), vid(jj + 1, ii)
            if diag[jj, ii] == 0:
                tris += [[a, d, c], [a, c, b]]
            else:
                tris += [[a, d, b], [d, c, b]]
    el = np.asarray(tris, dtype=np.int64)
    _check_ccw(vxy, el)
    # Wall: inner ring; each line runs like the triangle edge that owns it (b -> a, decreasing angle).
    wall = np.stack([vid(0, np.arange(n_surface) + 1), vid(0, np.arange(n_surface))], axis=1)
    # Farfield: outer ring, owned edge runs d -> c (increasing angle).
    far = np.stack([vid(n_layers, np.arange(n_surface)), vid(n_layers, np.arange(n_surface) + 1)], axis=1)
    return MeshInput(vxy=vxy, eb_type=np.array([WALL, FARFIELD], np.int32),
                     eb_count=np.array([n_surface, n_surface], np.int64),
                     eb_nodes=np.concatenate([wall, far]).astype(np.int64),
                     el_count=np.array([el.shape[0]], np.int64), el_nodes=el,
                     name=name or f"naca0012_ogrid_{n_surface}x{n_layers}_s{seed}")


def periodic_square(n: int, length: float = 10.0, jitter: float = 0.10, seed: int = 0,
                    name: Optional[str] = None) -> MeshInput:
    """Doubly periodic [0, L]^2 from n x n quads, 2 n^2 triangles, random diagonals.

    Node (i, j) is jittered by up to ``jitter`` times the spacing in each coordinate; images
    across the period share the vertex id and the jitter.
    """
    if n < 3:
        raise ValueError("n >= 3 required for a periodic square")
    rng = np.random.default_rng(seed)
    h = length / n
    base = np.arange(n) * h
    jx = rng.uniform(-jitter, jitter, size=(n, n)) * h
    jy = rng.uniform(-jitter, jitter, size=(n, n)) * h
    X = base[None, :] + jx            # [j, i]
    Y = base[:, None] + jy
    vxy = np.stack([X.ravel(), Y.ravel()], axis=1)

    def node(jj, ii):
        return (jj % n) * n + (ii % n), X[jj % n, ii % n] + length * (ii // n), Y[jj % n, ii % n] + length * (jj // n)

    diag = rng.integers(0, 2, size=(n, n))
    tris, txy = [], []
    for jj in range(n):
        for ii in range(n):
            a, b, c, d = node(jj, ii), node(jj, ii + 1), node(jj + 1, ii + 1), node(jj + 1, ii)
            quads = ([a, b, c], [a, c, d]) if diag[jj, ii] == 0 else ([a, b, d], [b, c, d])
            for t in quads:
                tris.append([p[0] for p in t])
                txy.append([[p[1], p[2]] for p in t])
    el = np.asarray(tris, dtype=np.int64)
    exy = np.asarray(txy, dtype=np.float64)
    _check_ccw_xy(exy)
    return MeshInput(vxy=vxy, eb_type=np.zeros(0, np.int32), eb_count=np.zeros(0, np.int64),
                     eb_nodes=np.zeros((0, 2), np.int64), el_count=np.array([el.shape[0]], np.int64),
                     el_nodes=el, elem_xy=exy, name=name or f"periodic_square_{n}_s{seed}")


def _check_ccw(vxy: np.ndarray, el: np.ndarray) -> None:
    _check_ccw_xy(vxy[el])


def _check_ccw_xy(exy: np.ndarray) -> None:
    x, y = exy[..., 0], exy[..., 1]
    J = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
    if not np.all(J > 0):
        raise ValueError(f"mesh has {int(np.sum(J <= 0))} non-CCW triangles")


def box_farfield(n: int = 4, length: float = 1.0, seed: int = 0, jitter: float = 0.1) -> MeshInput:
    """Small square [0, L]^2 with every boundary edge in one Farfield block (tests)."""
    rng = np.random.default_rng(seed)
    h = length / n
    X, Y = np.meshgrid(np.arange(n + 1) * h, np.arange(n + 1) * h)
    inner = (slice(1, n), slice(1, n))
    X[inner] += rng.uniform(-jitter, jitter, size=(n - 1, n - 1)) * h
    Y[inner] += rng.uniform(-jitter, jitter, size=(n - 1, n - 1)) * h
    vxy = np.stack([X.ravel(), Y.ravel()], axis=1)
    vid = lambda j, i: j * (n + 1) + i
    diag = rng.integers(0, 2, size=(n, n))
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = vid(j, i), vid(j, i + 1), vid(j + 1, i + 1), vid(j + 1, i)
            tris += [[a, b, c], [a, c, d]] if diag[j, i] == 0 else [[a, b, d], [b, c, d]]
    el = np.asarray(tris, np.int64)
    _check_ccw(vxy, el)
    bnd = []
    for i in range(n):
        bnd.append([vid(0, i), vid(0, i + 1)])          # bottom, left to right
    for j in range(n):
        bnd.append([vid(j, n), vid(j + 1, n)])          # right, upwards
    for i in range(n, 0, -1):
        bnd.append([vid(n, i), vid(n, i - 1)])          # top, right to left
    for j in range(n, 0, -1):
        bnd.append([vid(j, 0), vid(j - 1, 0)])          # left, downwards
    return MeshInput(vxy=vxy, eb_type=np.array([FARFIELD], np.int32), eb_count=np.array([len(bnd)], np.int64),
                     eb_nodes=np.asarray(bnd, np.int64), el_count=np.array([el.shape[0]], np.int64), el_nodes=el,
                     name=f"box_farfield_{n}_s{seed}")


def read_gmsh(path: str) -> MeshInput:
    """Gmsh 4.1 ASCII reader with the state machine of src/io/mesh_parser.rs:28-158.

    Physical names come from $PhysicalNames, entity -> physical tag from the 9th field of
    $Entities lines with at least 11 fields, coordinates from $Nodes lines with exactly three
    fields, and element blocks of dimension 1 (boundary lines) and 2 (triangles).
    """
    phys, ent = {}, {}
    verts: List[List[float]] = []
    eblocks, etypes, elblocks = [], [], []
    state, remaining = "idle", 0
    with open(path, "r") as f:
        for raw in f:
            line = raw.rstrip("\n").rstrip("\r")
            if line.startswith("$End"):
                state = "idle"
                continue
            for tag, st in (("$PhysicalNames", "phys0"), ("$Entities", "ent0"), ("$Nodes", "nodes0"),
                            ("$Elements", "elem0")):
                if line.startswith(tag):
                    state = st
                    break
            else:
                v = line.split()
                if state == "phys0":
                    state = "phys"
                elif state == "phys":
                    if len(v) == 3:
                        phys[int(v[1])] = v[2].strip('"')
                elif state == "ent0":
                    state = "ent"
                elif state == "ent":
                    if len(v) >= 11:
                        ent[int(v[0])] = int(v[8])
                elif state == "nodes0":
                    state = "nodes"
                elif state == "nodes":
                    if len(v) == 3:
                        verts.append([float(v[0]), float(v[1])])
                elif state == "elem0":
                    state = "block"
                elif state == "block":
                    dim, tag, num = int(v[0]), int(v[1]), int(v[3])
                    name = phys[ent[tag]]
                    if dim == 1:
                        eblocks.append([])
                        t = name.lower()
                        etypes.append(WALL if t == "wall" else (FARFIELD if t == "farfield" else -1))
                        state = "edges"
                    elif dim == 2:
                        elblocks.append([])
                        state = "tris"
                    remaining = num
                elif state == "edges":
                    eblocks[-1].append([int(v[1]) - 1, int(v[2]) - 1])
                    remaining -= 1
                    if remaining == 0:
                        state = "block"
                elif state == "tris":
                    elblocks[-1].append([int(v[1]) - 1, int(v[2]) - 1, int(v[3]) - 1])
                    remaining -= 1
                    if remaining == 0:
                        state = "block"
    if any(t < 0 for t in etypes):
        raise ValueError("boundary block with a physical name other than Wall / Farfield")
    eb_nodes = np.asarray([p for b in eblocks for p in b], np.int64).reshape(-1, 2)
    el_nodes = np.asarray([t for b in elblocks for t in b], np.int64).reshape(-1, 3)
    return MeshInput(vxy=np.asarray(verts, np.float64), eb_type=np.asarray(etypes, np.int32),
                     eb_count=np.asarray([len(b) for b in eblocks], np.int64), eb_nodes=eb_nodes,
                     el_count=np.asarray([len(b) for b in elblocks], np.int64), el_nodes=el_nodes,
                     name=path.rsplit("/", 1)[-1])


def write_gmsh(mesh: MeshInput, path: str) -> None:
    """Write a non-periodic MeshInput as Gmsh 4.1 ASCII readable by :func:`read_gmsh` and the reference parser."""
    if mesh.elem_xy is not None:
        raise ValueError("periodic meshes have no Gmsh form the reference parser reads")
    nb = len(mesh.eb_count)
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        f.write(f"$PhysicalNames\n3\n1 1 \"Wall\"\n1 2 \"Farfield\"\n2 3 \"Fluid\"\n$EndPhysicalNames\n")
        f.write(f"$Entities\n0 {nb} 1 0\n")
        for b in range(nb):
            tag = 1 if mesh.eb_type[b] == WALL else 2
            f.write(f"{b + 1} 0 0 0 0 0 0 1 {tag} 2 0 0\n")
        f.write(f"{nb + 1} 0 0 0 0 0 0 1 3 {nb} " + " ".join(str(i + 1) for i in range(nb)) + "\n")
        f.write("$EndEntities\n")
        nv = mesh.n_vertices
        f.write(f"$Nodes\n1 {nv} 1 {nv}\n2 {nb + 1} 0 {nv}\n")
        for i in range(nv):
            f.write(f"{i + 1}\n")
        for x, y in mesh.vxy:
            f.write(f"{float(x)!r} {float(y)!r} 0\n")
        f.write("$EndNodes\n")
        ne = mesh.n_elements
        nl = int(mesh.eb_count.sum())
        f.write(f"$Elements\n{nb + 1} {ne + nl} 1 {ne + nl}\n")
        tag, off = 1, 0
        for b in range(nb):
            cnt = int(mesh.eb_count[b])
            f.write(f"1 {b + 1} 1 {cnt}\n")
            for a, c in mesh.eb_nodes[off:off + cnt]:
                f.write(f"{tag} {a + 1} {c + 1}\n")
                tag += 1
            off += cnt
        f.write(f"2 {nb + 1} 2 {ne}\n")
        for a, b, c in mesh.el_nodes:
            f.write(f"{tag} {a + 1} {b + 1} {c + 1}\n")
            tag += 1
        f.write("$EndElements\n")
