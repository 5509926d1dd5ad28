"""Host tooling: the reference's case-folder formats and the synthetic meshes."""
import os

import numpy as np
import pytest


def test_result_file_names(sw):
    # names the reference wrote for the shipped twin experiment
    # (configue_fonctionels/sp_toce/reference/ResultHydro_p0000d00h00m10s500.txt ...)
    assert sw.result_name(10.5) == "ResultHydro_p0000d00h00m10s500.txt"
    assert sw.result_name(30.0) == "ResultHydro_p0000d00h00m30s000.txt"
    assert sw.result_name(0.0) == "ResultHydro_p0000d00h00m00s000.txt"
    assert sw.result_name(3 * 86400 + 3661.25) == "ResultHydro_p0003d01h01m01s250.txt"


def test_geometry_round_trip(sw, tmp_path):
    g = sw.Geometry.delaunay(12, 9, 120.0, 90.0, jitter=0.2, seed=3)
    p = str(tmp_path / "geometry.sw2d")
    sw._native.host_check(sw._native.host_lib().sw2dh_geom_write(g._h, p.encode()))
    g2 = sw.Geometry.read(p)
    assert g.sizes() == g2.sizes()
    for k in ("cell_area", "cell_x", "cell_y", "cell_zb", "bnd_cell", "bnd_len", "bnd_nx",
              "bnd_ny", "bnd_code", "int_L", "int_R", "int_len", "int_nx", "int_ny", "int_xm",
              "int_ym", "node_x", "node_y", "cell_tri"):
        np.testing.assert_array_equal(g.array(k), g2.array(k), err_msg=k)


@pytest.mark.parametrize("kind", ["rect", "delaunay", "bricks"])
def test_mesh_geometry_is_consistent(sw, kind):
    g = {"rect": lambda: sw.Geometry.rect(10, 7, 40.0, 21.0),
         "delaunay": lambda: sw.Geometry.delaunay(14, 11, 140.0, 110.0, jitter=0.3, seed=11),
         "bricks": lambda: sw.Geometry.bricks(9, 6, 45.0, 30.0)}[kind]()
    Lx, Ly = {"rect": (40.0, 21.0), "delaunay": (140.0, 110.0), "bricks": (45.0, 30.0)}[kind]
    nc, nb, ni, nn = g.sizes()
    A = g.array("cell_area")
    assert np.isclose(A.sum(), Lx * Ly, rtol=1e-13)
    assert nc + nn - (nb + ni) == 1  # Euler: V - E + F = 1 for a planar disc
    for p in ("bnd", "int"):
        n = np.hypot(g.array(p + "_nx"), g.array(p + "_ny"))
        np.testing.assert_allclose(n, 1.0, rtol=0, atol=1e-14)
    cx, cy = g.array("cell_x"), g.array("cell_y")
    L, R = g.array("int_L"), g.array("int_R")
    assert np.all(g.array("int_nx") * (cx[R] - cx[L]) + g.array("int_ny") * (cy[R] - cy[L]) > 0)
    bc = g.array("bnd_cell")
    assert np.all(g.array("bnd_nx") * (cx[bc] - g.array("bnd_xm")) +
                  g.array("bnd_ny") * (cy[bc] - g.array("bnd_ym")) > 0)
    # closed cells: the outward-normal-weighted edge lengths sum to zero
    sx = np.zeros(nc)
    sy = np.zeros(nc)
    ln, nx, ny = g.array("int_len"), g.array("int_nx"), g.array("int_ny")
    np.add.at(sx, L, ln * nx)
    np.add.at(sy, L, ln * ny)
    np.add.at(sx, R, -ln * nx)
    np.add.at(sy, R, -ln * ny)
    np.add.at(sx, bc, -g.array("bnd_len") * g.array("bnd_nx"))
    np.add.at(sy, bc, -g.array("bnd_len") * g.array("bnd_ny"))
    assert np.max(np.abs(sx)) < 1e-9 and np.max(np.abs(sy)) < 1e-9


def _incircle(a, b, c, d):
    adx, ady = a[0] - d[0], a[1] - d[1]
    bdx, bdy = b[0] - d[0], b[1] - d[1]
    cdx, cdy = c[0] - d[0], c[1] - d[1]
    ad, bd, cd = adx * adx + ady * ady, bdx * bdx + bdy * bdy, cdx * cdx + cdy * cdy
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx)


@pytest.mark.parametrize("jitter,seed", [(0.25, 1), (0.3, 5)])
def test_jittered_mesh_is_delaunay(sw, jitter, seed):
    """Brute force: no vertex lies strictly inside the circumcircle of the
    triangle across any internal edge (the empty-circle property)."""
    g = sw.Geometry.delaunay(25, 20, 250.0, 200.0, jitter=jitter, seed=seed)
    P = np.stack([g.array("node_x"), g.array("node_y")], 1)
    T = g.array("cell_tri").reshape(-1, 3)
    for l, r in zip(g.array("int_L"), g.array("int_R")):
        tl, tr = T[l], T[r]
        d = [v for v in tr if v not in tl]
        assert len(d) == 1
        a, b, c = P[tl]
        if (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]) < 0:
            b, c = c, b
        assert _incircle(a, b, c, P[d[0]]) <= 1e-9 * 250.0 ** 4


def _folder_with_solv(sw, tmp_path, extra_lines):
    sw.dambreak_case(str(tmp_path), nx=4, ny=2, Lx=8.0, Ly=4.0, x_dam=4.0, t_max=1.0)
    p = tmp_path / "input" / "input.solv"
    p.write_text(p.read_text() + extra_lines)
    return str(tmp_path)


def test_input_solv_accepts_assimilation_keys_with_comments(sw, tmp_path):
    f = _folder_with_solv(sw, tmp_path, "alpha_l1     0     # comment\ncheckpoint_stride 30\n")
    c = sw.Case(f)
    assert c.keys["checkpoint_stride"] == "30"


def test_input_solv_rejects_unknown_key(sw, tmp_path):
    f = _folder_with_solv(sw, tmp_path, "no_such_key 3\n")
    with pytest.raises(sw._native.SW2DError, match="not recognized"):
        sw.Case(f)


def test_input_solv_rejects_misformatted_line(sw, tmp_path):
    f = _folder_with_solv(sw, tmp_path, "t_max 1 2\n")
    with pytest.raises(sw._native.SW2DError, match="misformatted"):
        sw.Case(f)


def test_case_load_reads_maps_and_bc(sw, tmp_path):
    from sw2d_amd import cases
    cases.mixed_case(str(tmp_path), friction="manning")
    c = sw.Case(str(tmp_path))
    assert c.n_series == 5 and c.n_times == 4
    assert list(c["series_type"]) == [3, 3, 0, 1, 2]  # q q fr h z
    assert c.params.head_loss == 1 and c.params.friction == 1
    assert len(c["hk1"]) == c.n_cell and np.all(c["phi"] >= 0)
    assert set(np.unique(c["bnd_series"])) <= {1, 2, 3, 4}


def test_uniform_and_distributed_maps(sw, tmp_path):
    from sw2d_amd import cases
    sw.dambreak_case(str(tmp_path), nx=4, ny=2, Lx=8.0, Ly=4.0, x_dam=4.0, t_max=1.0)
    nc = 16
    vals = np.linspace(0.1, 1.0, nc)
    cases.write_distrib_map(str(tmp_path / "input" / "SP_cell_porosity_map.txt"), ["phi"], [vals])
    c = sw.Case(str(tmp_path))
    np.testing.assert_array_equal(c["phi"], vals)
    np.testing.assert_array_equal(c["beta"], np.ones(nc))  # 'Unif 1' map


def test_brick_mesh_cells_are_hexagons(sw):
    """The polygon mesh of sw2dh_geom_cells: interior bricks have 6 nodes and
    6 internal edges (2 above, 2 below, left, right), the rows' end bricks
    fewer; every edge is shared by at most two cells."""
    g = sw.Geometry.bricks(9, 6, 45.0, 30.0)
    nc, nb, ni, nn = g.sizes()
    assert nc == 3 * 9 + 3 * (8 + 2)  # even rows 9 bricks, odd rows 8 + 2 half bricks
    L, R = g.array("int_L"), g.array("int_R")
    deg = np.bincount(np.concatenate([L, R]), minlength=nc)
    cx, cy = g.array("cell_x"), g.array("cell_y")
    inner = (cx > 5.0) & (cx < 40.0) & (cy > 5.0) & (cy < 25.0)
    assert np.all(deg[inner] == 6) and deg.max() == 6
    A = g.array("cell_area")
    np.testing.assert_allclose(A[inner], 5.0 * 5.0, rtol=1e-13)


def test_polygon_mesh_rejects_bad_cells(sw):
    from sw2d_amd import _native as N
    x = np.array([0.0, 1.0, 1.0, 0.0])
    y = np.array([0.0, 0.0, 1.0, 1.0])
    with pytest.raises(N.SW2DError, match="fewer than 3 nodes"):
        sw.Geometry.cells(x, y, np.zeros(4), [[0, 1]])
    with pytest.raises(N.SW2DError, match="outside"):
        sw.Geometry.cells(x, y, np.zeros(4), [[0, 1, 7]])
