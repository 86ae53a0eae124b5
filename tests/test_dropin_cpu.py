"""CPU checks of the C++ drop-in layer's host-only parts (no GPU): the TetGen and VTU
readers and the Kuhn-beam generator of tlfea_driver agree exactly with the Python ones."""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import MESHES
from tlfea_amd import build as B
from tlfea_amd import mesh as M


def _dump(spec, tmp_path):
    if not os.path.exists(B.DRIVER):
        pytest.skip("tlfea_driver not built")
    out = str(tmp_path / "mesh.bin")
    r = subprocess.run([B.DRIVER, "mesh_info", "--mesh", spec, "--out", out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout)
    raw = open(out, "rb").read()
    nb = 8 * 3 * info["n_nodes"]
    nodes = np.frombuffer(raw[:nb], dtype=np.float64).reshape(3, -1).T  # column-major n x 3
    elems = np.frombuffer(raw[nb:], dtype=np.int32).reshape(info["nn"], -1).T
    return nodes, elems


@pytest.mark.parametrize("spec,ref", [
    ("kuhn:20,2,2", lambda: M.kuhn_beam(20, 2, 2)),
    ("kuhn4:12,3,2", lambda: M.kuhn_beam(12, 3, 2, tet10=False)),
    ("tetgen:" + os.path.join(MESHES, "cube.1"),
     lambda: M.read_tetgen(os.path.join(MESHES, "cube.1.node"), os.path.join(MESHES, "cube.1.ele"))),
    ("vtu:" + os.path.join(MESHES, "beam.vtu"), lambda: M.read_vtu_ascii(os.path.join(MESHES, "beam.vtu"))),
    ("vtu:" + os.path.join(MESHES, "beam_highres_t4.vtu"),
     lambda: M.read_vtu_ascii(os.path.join(MESHES, "beam_highres_t4.vtu"))),
])
def test_cpp_mesh_io_matches_python(spec, ref, tmp_path):
    nodes, elems = _dump(spec, tmp_path)
    m = ref()
    assert np.array_equal(elems, m.elems)
    # the Kuhn generator uses the same formula; files parse to the same doubles
    assert np.allclose(nodes, m.xyz(), rtol=0, atol=1e-15)


@pytest.mark.parametrize("spec,ref,nn", [
    ("kuhn:20,2,2", lambda: M.kuhn_beam(20, 2, 2), 10),
    ("kuhn4:12,3,2", lambda: M.kuhn_beam(12, 3, 2, tet10=False), 4),
    ("tetgen:" + os.path.join(MESHES, "cube.1"),
     lambda: M.read_tetgen(os.path.join(MESHES, "cube.1.node"), os.path.join(MESHES, "cube.1.ele")), 10),
])
def test_feat_vtk_writer_layout_and_round_trip(spec, ref, nn, tmp_path):
    """ANCFCPUUtils::WriteFEAT10ToVTK / WriteFEAT4ToVTK of the drop-in (mesh_utils.cc:962-1077
    layout) on a sheared mesh: the header lines, POINTS/CELLS/CELL_TYPES/POINT_DATA sections
    with their counts, cell type 24 / 10, the connectivity exactly, positions and the
    displacement vectors to the 6 significant digits the format keeps; and the Python
    writer produces the same bytes."""
    if not os.path.exists(B.DRIVER):
        pytest.skip("tlfea_driver not built")
    path = str(tmp_path / "out.vtk")
    shear = 0.01
    r = subprocess.run([B.DRIVER, "write_vtk", "--mesh", spec, "--vtk-feat", path, "--shear", str(shear)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = ref()
    X, Y, Z = m.x, m.y, m.z
    x, y, z = X + shear * Y * Z, Y - shear * X * Z, Z + shear * X * Y
    v = M.read_vtk_legacy(path)
    assert v["header"] == "# vtk DataFile Version 3.0"
    assert v["title"] == ("FEAT10 Tetrahedral Mesh" if nn == 10 else "FEAT4 Tetrahedral Mesh")
    assert v["format"] == "ASCII" and v["dataset"] == "DATASET UNSTRUCTURED_GRID"
    assert v["point_type"] == "double"
    assert np.array_equal(v["cells"], m.elems)
    assert np.all(v["cell_types"] == (24 if nn == 10 else 10))
    pos = np.stack([x, y, z], 1)
    scale = np.maximum(np.abs(pos), 1e-300)
    assert np.all(np.abs(v["points"] - pos) <= 5.01e-6 * scale)
    disp = pos - m.xyz()
    assert np.all(np.abs(v["point_vectors"]["displacement"] - disp) <= 5.01e-6 * np.abs(disp) + 1e-300)
    py = str(tmp_path / "py.vtk")
    M.write_feat_vtk(py, m, x, y, z)
    assert open(py).read() == open(path).read()
