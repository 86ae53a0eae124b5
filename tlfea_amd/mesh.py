"""Mesh construction and I/O for TET10 / TET4 meshes.

Conventions shared with the element code (reference `src/elements/FEAT10Data.cuh:41-51`):

* connectivity is int32, column-major ``conn[e + a * n_elem]`` (the Eigen ``MatrixXi``
  default storage the reference copies with ``.data()`` at `FEAT10Data.cuh:444-446`);
* TET10 node order is 4 corners then the edge midsides of the edges
  (0,1), (1,2), (0,2), (0,3), (1,3), (2,3) (`FEAT10Data.cu:125-135`);
* positions are three separate float64 arrays x, y, z.

The structured cantilever beams of the benchmark configurations have no generator in the
reference; ``kuhn_beam`` builds them on the 2x lattice (every lattice point is either a hex
corner or a midside of a Kuhn edge, so an ``nx x ny x nz`` cell beam has
``(2nx+1)(2ny+1)(2nz+1)`` nodes).  Three of the six Kuhn tets of a cube are negatively
oriented; since the element code uses the signed ``det J`` as the volume weight
(`FEAT10Data.cu:169-173`, `FEAT10DataFunc.cuh:431-433`) those tets get two corners (and
the matching midsides) swapped so every element has ``det J > 0``.
"""
from __future__ import annotations

import itertools
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

# edges of a TET10 in the reference's midside order
T10_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))
# TetGen order -> standard order (reference `src/utils/cpu_utils.cc:606`)
TETGEN_TO_STANDARD = (0, 1, 2, 3, 6, 7, 9, 5, 8, 4)


@dataclass
class TetMesh:
    """A tetrahedral mesh with SoA node positions and column-major connectivity."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    elems: np.ndarray  # (n_elem, nn) int32, row per element
    fixed: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    f_ext: np.ndarray | None = None  # (3 * n_nodes,) AoS [3*node + d]
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_elem(self) -> int:
        return int(self.elems.shape[0])

    @property
    def nn(self) -> int:
        return int(self.elems.shape[1])

    @property
    def n_dof(self) -> int:
        return 3 * self.n_nodes

    def conn_colmajor(self) -> np.ndarray:
        """Connectivity as the flat column-major int32 array ``conn[e + a*n_elem]``."""
        return np.ascontiguousarray(self.elems.T, dtype=np.int32).reshape(-1)

    def xyz(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], axis=1)


# --------------------------------------------------------------------------------------
# Structured Kuhn beam
# --------------------------------------------------------------------------------------
def _kuhn_tets():
    """The six Kuhn tets of the unit cube as corner offsets, each positively oriented."""
    unit = np.eye(3, dtype=np.int64)
    tets = []
    for perm in itertools.permutations(range(3)):
        v0 = np.zeros(3, np.int64)
        v1 = v0 + unit[perm[0]]
        v2 = v1 + unit[perm[1]]
        v3 = v2 + unit[perm[2]]
        corners = [v0, v1, v2, v3]
        det = np.linalg.det(np.stack([v1 - v0, v2 - v0, v3 - v0]).astype(float))
        if det < 0:  # swap corners 1 and 2 to make the signed volume positive
            corners = [v0, v2, v1, v3]
        tets.append(np.stack(corners))
    return tets


def kuhn_beam(nx: int, ny: int, nz: int, lx: float = 1.0, ly: float = 0.1, lz: float = 0.1,
              tet10: bool = True) -> TetMesh:
    """Structured beam ``[0,lx] x [0,ly] x [0,lz]`` of ``nx*ny*nz`` hex cells, 6 tets each.

    TET10 nodes live on the 2x lattice with id ``(I*NY + J)*NZ + K`` where
    ``NY = 2ny+1`` and ``NZ = 2nz+1`` (so node ids grow along x slab by slab).
    TET4 meshes use the 1x lattice with the same id formula and ``NY = ny+1``.
    Elements are ordered cell by cell (x outermost) and six tets per cell.
    """
    s = 2 if tet10 else 1
    NX, NY, NZ = s * nx + 1, s * ny + 1, s * nz + 1
    I, J, K = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    x = (I.reshape(-1) * (lx / (s * nx))).astype(np.float64)
    y = (J.reshape(-1) * (ly / (s * ny))).astype(np.float64)
    z = (K.reshape(-1) * (lz / (s * nz))).astype(np.float64)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ci.reshape(-1), cj.reshape(-1), ck.reshape(-1)], axis=1)  # (nc,3)

    def node_id(lat):  # lat: (..., 3) lattice coordinates
        return (lat[..., 0] * NY + lat[..., 1]) * NZ + lat[..., 2]

    per_cell = []
    for corners in _kuhn_tets():
        lat_c = s * (cells[:, None, :] + corners[None, :, :])  # (nc,4,3)
        ids = [node_id(lat_c[:, a, :]) for a in range(4)]
        if tet10:
            for (a, b) in T10_EDGES:
                ids.append(node_id((lat_c[:, a, :] + lat_c[:, b, :]) // 2))
        per_cell.append(np.stack(ids, axis=1))
    elems = np.stack(per_cell, axis=1).reshape(-1, 10 if tet10 else 4).astype(np.int32)
    m = TetMesh(x, y, z, elems)
    m.meta.update(dict(kind="kuhn_beam", nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz))
    return m


def clamp_and_tip_load(m: TetMesh, total_load, tol_frac: float = 5e-4) -> TetMesh:
    """Fix nodes at x = x_min and spread ``total_load`` (3-vector or -z scalar) evenly over
    the x = x_max nodes, as the reference beam drivers do
    (`examples/beam_linear_static.cc:151-176`)."""
    L = m.x.max() - m.x.min()
    tol = tol_frac * L
    fixed = np.nonzero(np.abs(m.x - m.x.min()) < tol)[0].astype(np.int32)
    tip = np.nonzero(np.abs(m.x - m.x.max()) < tol)[0]
    load = np.asarray(total_load, dtype=np.float64)
    if load.ndim == 0:
        load = np.array([0.0, 0.0, float(load)])
    f = np.zeros(3 * m.n_nodes)
    for d in range(3):
        f[3 * tip + d] = load[d] / len(tip)
    m.fixed = fixed
    m.f_ext = f
    m.meta["tip_nodes"] = tip.astype(np.int32)
    return m


def beam_config(name: str) -> TetMesh:
    """The two fixed benchmark beams: ``"C1"`` (20x2x2 cells, 480 TET10, 1000 N) and
    ``"C2"`` (200x20x20 cells, 480,000 TET10, 1e5 N), both 1.0 x 0.1 x 0.1 m, clamped at
    x = 0 and loaded in -z at x = L."""
    if name == "C1":
        return clamp_and_tip_load(kuhn_beam(20, 2, 2), -1000.0)
    if name == "C2":
        return clamp_and_tip_load(kuhn_beam(200, 20, 20), -1.0e5)
    raise ValueError(f"unknown beam config {name!r}")


def random_beam_batch(n: int, seed: int):
    """The batched-instance sampler: ``n`` beams with cells nx~U{10..80}, ny,nz~U{2..8},
    length U(0.5,2) m, square cross-section side U(0.05,0.2) m, E log-uniform [1e6,1e10] Pa,
    nu~U(0.2,0.45), rho~U(1000,8000) and a tip load of uniform random direction with
    log-uniform magnitude in [1,1e4] N.  Returns a list of (mesh, material dict)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(10, 81))
        ny = int(rng.integers(2, 9))
        nz = int(rng.integers(2, 9))
        L = float(rng.uniform(0.5, 2.0))
        a = float(rng.uniform(0.05, 0.2))
        E = float(10 ** rng.uniform(6, 10))
        nu = float(rng.uniform(0.2, 0.45))
        rho = float(rng.uniform(1000, 8000))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        mag = float(10 ** rng.uniform(0, 4))
        m = clamp_and_tip_load(kuhn_beam(nx, ny, nz, L, a, a), d * mag)
        out.append((m, dict(E=E, nu=nu, rho=rho)))
    return out


# --------------------------------------------------------------------------------------
# Readers / writers
# --------------------------------------------------------------------------------------
def read_tetgen(node_file: str, ele_file: str) -> TetMesh:
    """TetGen ``.node``/``.ele`` reader with the reference's adaptive 0/1 base and the
    TetGen -> standard TET10 remap (`src/utils/cpu_utils.cc:596-735`)."""
    with open(node_file) as fh:
        lines = [ln.split() for ln in fh if ln.strip() and not ln.lstrip().startswith("#")]
    n_nodes, dim = int(lines[0][0]), int(lines[0][1])
    if dim != 3:
        raise ValueError(f"only 3D nodes are supported, found {dim}D")
    rows = [(int(r[0]), float(r[1]), float(r[2]), float(r[3])) for r in lines[1:1 + n_nodes]]
    min_id = min(r[0] for r in rows)
    off = 0 if min_id == 0 else 1
    xyz = np.zeros((n_nodes, 3))
    for nid, a, b, c in rows:
        idx = nid - off
        if 0 <= idx < n_nodes:
            xyz[idx] = (a, b, c)
    with open(ele_file) as fh:
        lines = [ln.split() for ln in fh if ln.strip() and not ln.lstrip().startswith("#")]
    n_el, npe = int(lines[0][0]), int(lines[0][1])
    if npe != 10:
        raise ValueError(f"only T10 elements (10 nodes) are supported, found {npe}")
    recs = [(int(r[0]), [int(v) for v in r[1:11]]) for r in lines[1:1 + n_el]]
    min_eid = min(r[0] for r in recs)
    min_nid = min(min(r[1]) for r in recs)
    eoff = 0 if min_eid == 0 else 1
    noff = 0 if min_nid == 0 else 1
    elems = np.zeros((n_el, 10), np.int32)
    for eid, ids in recs:
        tg = [v - noff for v in ids]
        std = [tg[TETGEN_TO_STANDARD[i]] for i in range(10)]
        idx = eid - eoff
        if 0 <= idx < n_el:
            elems[idx] = std
    return TetMesh(xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), elems)


def read_vtu_ascii(path: str) -> TetMesh:
    """ASCII VTU reader for VTK_QUADRATIC_TETRA (24) or VTK_TETRA (10) meshes, standing in
    for the `mophi::LoadVtu` call of `examples/beam_linear_static.cc:69-99`.  VTK's quadratic
    tet midside order is the reference's edge order, so no remap is applied."""
    root = ET.parse(path).getroot()
    piece = root.find(".//Piece")
    npts = int(piece.get("NumberOfPoints"))
    ncells = int(piece.get("NumberOfCells"))

    def arr(name_pred, dtype):
        for da in root.iter("DataArray"):
            if name_pred(da):
                txt = "".join(t for t in da.itertext())
                txt = re.sub(r"<[^>]*>", " ", txt)
                return np.array(txt.split(), dtype=dtype)
        raise ValueError("missing DataArray")

    pts_da = piece.find("Points").find("DataArray")
    pts = np.array(pts_da.text.split(), dtype=np.float64)[: 3 * npts].reshape(npts, 3)
    cells = piece.find("Cells")
    byname = {da.get("Name"): da for da in cells.iter("DataArray")}
    conn = np.array(byname["connectivity"].text.split(), dtype=np.int64)
    offs = np.array(byname["offsets"].text.split(), dtype=np.int64)
    types = np.array(byname["types"].text.split(), dtype=np.int64)
    assert len(offs) == ncells and len(types) == ncells
    if np.all(types == 24):
        nn = 10
    elif np.all(types == 10):
        nn = 4
    else:
        raise ValueError("mixed or unsupported VTK cell types")
    elems = conn.reshape(ncells, nn).astype(np.int32)
    return TetMesh(pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), elems)


def write_vtk_legacy(path: str, m: TetMesh, x=None, y=None, z=None) -> None:
    """Legacy ASCII VTK writer with the layout of the reference's `WriteOutputVTK`
    (`src/elements/FEAT10Data.cu:759-794`): POINTS as float, CELLS with ``nn+1`` ints per
    row and CELL_TYPES 24 (TET10) or 10 (TET4)."""
    x = m.x if x is None else x
    y = m.y if y is None else y
    z = m.z if z is None else z
    nn = m.nn
    with open(path, "w") as out:
        out.write("# vtk DataFile Version 3.0\n")
        out.write(f"T{nn} mesh output\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        out.write(f"POINTS {len(x)} float\n")
        for i in range(len(x)):
            out.write(f"{x[i]:g} {y[i]:g} {z[i]:g}\n")
        out.write(f"CELLS {m.n_elem} {m.n_elem * (nn + 1)}\n")
        for e in range(m.n_elem):
            out.write(f"{nn} " + " ".join(str(int(v)) for v in m.elems[e]) + " \n")
        out.write(f"CELL_TYPES {m.n_elem}\n")
        ctype = 24 if nn == 10 else 10
        for _ in range(m.n_elem):
            out.write(f"{ctype}\n")


def write_feat_vtk(path: str, m: TetMesh, x, y, z) -> None:
    """Legacy ASCII VTK of the mesh at positions (x, y, z) with the displacement as point
    data, the layout of the reference's ``WriteFEAT10ToVTK`` / ``WriteFEAT4ToVTK``
    (``src/utils/mesh_utils.cc:962-1077``): values with 6 significant digits, as a C++
    stream writes them by default."""
    nn = m.nn
    title = "FEAT10 Tetrahedral Mesh" if nn == 10 else "FEAT4 Tetrahedral Mesh"
    with open(path, "w") as out:
        out.write(f"# vtk DataFile Version 3.0\n{title}\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        out.write(f"POINTS {len(x)} double\n")
        for i in range(len(x)):
            out.write(f"{x[i]:g} {y[i]:g} {z[i]:g}\n")
        out.write(f"\nCELLS {m.n_elem} {m.n_elem * (nn + 1)}\n")
        for e in range(m.n_elem):
            out.write(f"{nn} " + " ".join(str(int(v)) for v in m.elems[e]) + "\n")
        out.write(f"\nCELL_TYPES {m.n_elem}\n")
        ctype = "24" if nn == 10 else "10"
        out.write("".join(f"{ctype}\n" for _ in range(m.n_elem)))
        out.write(f"\nPOINT_DATA {len(x)}\nVECTORS displacement double\n")
        for i in range(len(x)):
            out.write(f"{x[i] - m.x[i]:g} {y[i] - m.y[i]:g} {z[i] - m.z[i]:g}\n")


def read_vtk_legacy(path: str) -> dict:
    """Parses a legacy ASCII VTK unstructured grid of tetrahedra as the writers above and
    the reference's write them: returns ``title``, ``points`` (n x 3), ``point_type``,
    ``cells`` (n_cells x nn), ``cell_types``, and ``point_vectors`` {name: n x 3}."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    out = {"header": lines[0], "title": lines[1], "format": lines[2], "dataset": lines[3], "point_vectors": {}}
    i = 4
    while i < len(lines):
        tok = lines[i].split()
        if not tok:
            i += 1
            continue
        if tok[0] == "POINTS":
            n = int(tok[1])
            out["point_type"] = tok[2]
            out["points"] = np.array([lines[i + 1 + k].split() for k in range(n)], dtype=np.float64).reshape(n, 3)
            i += 1 + n
        elif tok[0] == "CELLS":
            n, total = int(tok[1]), int(tok[2])
            rows = [lines[i + 1 + k].split() for k in range(n)]
            counts = np.array([int(r[0]) for r in rows])
            assert counts.sum() + n == total, "CELLS size field does not match its rows"
            assert np.all(counts == counts[0]), "mixed cell sizes"
            out["cells"] = np.array([r[1:] for r in rows], dtype=np.int64)
            i += 1 + n
        elif tok[0] == "CELL_TYPES":
            n = int(tok[1])
            out["cell_types"] = np.array([int(lines[i + 1 + k]) for k in range(n)])
            i += 1 + n
        elif tok[0] == "POINT_DATA":
            n = int(tok[1])
            vec = lines[i + 1].split()
            assert vec[0] == "VECTORS", "only VECTORS point data is written here"
            out["point_vectors"][vec[1]] = np.array([lines[i + 2 + k].split() for k in range(n)],
                                                    dtype=np.float64).reshape(n, 3)
            i += 2 + n
        else:
            raise ValueError(f"unexpected VTK line: {lines[i]!r}")
    return out

