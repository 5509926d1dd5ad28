"""Probes and point location on the reference's functional case sp_toce
(2,752 quads + 2,048 triangles): the probe file reader
(swSolverParam::readProbes, swSolverParam.cpp:958-1022), the winding-angle
point location (swMesh::getCellFromCoordinates, swMesh.cpp:695-759) and the
probe series format (swSp::init header, swSp.cpp:754-783;
swSp::storeResultProbes, swSp.cpp:1611-1639), against what the compiled
reference wrote on the same case (tests/golden/make_sp_toce_case.py)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, sp_toce_folder


@pytest.fixture(scope="module")
def toce_probe_case(sw, tmp_path_factory):
    d = str(tmp_path_factory.mktemp("toce_probes"))
    sp_toce_folder(sw, d, probe_variant=True)
    return sw.Case(d)


def _ref():
    return np.load(os.path.join(GOLDEN, "sp_toce_input.npz"), allow_pickle=False)


def test_probe_reader_and_header_match_the_reference(toce_probe_case):
    ref = str(_ref()["probes_txt"]).splitlines()
    c = toce_probe_case
    n = len(c.probe_names)
    assert n == 16 and c.probe_names[0] == "probe2"
    hdr = [f"{nm}\t{x:.10e}\t{y:.10e}" for nm, x, y in zip(c.probe_names, c.a["probe_x"], c.a["probe_y"])]
    assert hdr == ref[:n]
    assert ref[n] == "time\t" + "".join(f"{nm}_zb\t{nm}_h\t{nm}_q\t{nm}_r\t" for nm in c.probe_names)


def _line(c, t, h, qx, qy):
    cells = c.a["probe_cell"]
    zb = c.a["cell_zb"]
    return f"{t:.10e}\t" + "".join(f"{zb[k]:.10e}\t{h[k]:.10e}\t{qx[k]:.10e}\t{qy[k]:.10e}\t" for k in cells)


def test_located_cells_reproduce_the_reference_probe_lines(toce_probe_case):
    """The first line (initial state) and the last (the reference's own final
    map) rebuilt from the located cells equal the reference's lines byte for
    byte; the final map differs from cell to cell, so this pins each cell."""
    g = _ref()
    ref = str(g["probes_txt"]).splitlines()
    c = toce_probe_case
    assert ref[17] == _line(c, 0.0, c.a["h0"], c.a["qx0"], c.a["qy0"])
    assert ref[-1] == _line(c, 2.0, g["probes_h"], g["probes_qx"], g["probes_qy"])
    # a quad cell is among them
    from sw2d_amd.cases import Geometry
    geom = Geometry.read(os.path.join(c.folder, "input", "geometry.sw2d"))
    assert np.array_equal(geom.locate(c.a["probe_x"], c.a["probe_y"]), c.a["probe_cell"])


def test_locate_outside_and_probe_outside_the_mesh_is_refused(sw, tmp_path):
    d = str(tmp_path)
    sp_toce_folder(sw, d, probe_variant=True)
    c = sw.Case(d)
    from sw2d_amd.cases import Geometry
    geom = Geometry.read(os.path.join(d, "input", "geometry.sw2d"))
    # the case's first probe (0.331, 7.655) is outside the mesh
    assert geom.locate([0.331, -100.0], [7.655, 0.0]).tolist() == [-1, -1]
    # centroids fall in their own cells
    cx, cy = c.a["cell_x"][::97], c.a["cell_y"][::97]
    assert np.array_equal(geom.locate(cx, cy), np.arange(0, c.n_cell, 97))
    # the reader stops at the first empty line (the case's file ends with one)
    lines = open(os.path.join(d, "input", "probe_locations.txt")).read().splitlines()
    assert lines[-1] == ""
    with open(os.path.join(d, "input", "probe_locations.txt"), "w") as f:
        f.write("\n".join(lines + ["0.331\t7.655\tprobe1"]) + "\n")
    assert len(sw.Case(d).probe_names) == 16
    with open(os.path.join(d, "input", "probe_locations.txt"), "w") as f:
        f.write("\n".join(lines[:-1] + ["0.331\t7.655\tprobe1"]) + "\n")
    with pytest.raises(sw._native.SW2DError, match="out of the mesh"):
        sw.Case(d)
    with open(os.path.join(d, "input", "probe_locations.txt"), "w") as f:
        f.write("x y name\n1.0 2.0\n")
    with pytest.raises(sw._native.SW2DError, match="misformatted"):
        sw.Case(d)
