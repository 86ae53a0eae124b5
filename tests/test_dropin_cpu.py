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
