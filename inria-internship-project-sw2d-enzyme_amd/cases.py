"""Case folders in the reference's format, and the synthetic workloads.

A case folder is what the reference's ``sw2dSolver <folder>`` reads
(src/sw2dSolver/main.cpp:18-188): ``input/geometry.sw2d`` (v2 binary),
``input/input.solv`` and the parameter / boundary files.  Everything is
written by the C++ host tooling (include/sw2d_host.h); this module only
composes the inputs.

Workloads:
  dam-break   200 x 100 m, 100 x 100 quads split in 2 triangles (20,000 cells),
              flat bed, phi = 1, Manning n = 0.02, h = 2.0 m for x < 100 and
              0.5 m elsewhere, u = v = 0, walls (q = 0), CFL 0.9, 10 s.
  scale       5 x 5 km jittered-lattice Delaunay mesh (~1e6 cells at the full
              size), bed slope 1e-3 plus 20 seeded Gaussian bumps, porosity
              in [0.3, 1] from smoothed seeded noise, partially dry start,
              triangular 0 -> 500 -> 0 m3/s inflow on the west edge, Manning
              0.03.
"""
import ctypes as C
import os

import numpy as np

from . import _native as N


class Geometry:
    """A mesh held by the host library (sw2dh_geom)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.lib = N.host_lib()

    @classmethod
    def rect(cls, nx, ny, Lx, Ly):
        out = C.c_void_p()
        N.host_check(N.host_lib().sw2dh_geom_rect(nx, ny, Lx, Ly, C.byref(out)))
        return cls(out)

    @classmethod
    def delaunay(cls, px, py, Lx, Ly, jitter=0.25, seed=0):
        out = C.c_void_p()
        flips = C.c_int64()
        N.host_check(N.host_lib().sw2dh_geom_delaunay(px, py, Lx, Ly, jitter, seed, C.byref(out),
                                                      C.byref(flips)))
        g = cls(out)
        g.n_flips = flips.value
        return g

    @classmethod
    def cells(cls, x, y, z, cells):
        """Any polygon mesh: node coordinates and elevations, and per cell
        its nodes in order around it (sw2dh_geom_cells)."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        ptr = np.concatenate([[0], np.cumsum([len(c) for c in cells])]).astype(np.int64)
        nodes = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int64) for c in cells]))
        out = C.c_void_p()
        N.host_check(N.host_lib().sw2dh_geom_cells(len(x), N.dptr(x), N.dptr(y), N.dptr(z), len(cells),
                                                   N.i64ptr(ptr), N.i64ptr(nodes), C.byref(out)))
        return cls(out)

    @classmethod
    def bricks(cls, nx, ny, Lx, Ly):
        """Staggered brick wall of nx x ny bricks (rows offset by half a
        brick, half bricks closing alternate rows): every node of a row's
        boundary is a vertex of the bricks on both sides, so an interior
        brick is a 6-node polygon with 6 internal edges (2 above, 2 below,
        left, right) and the half bricks have 4 or 5 edges."""
        w, h = Lx / nx, Ly / ny
        xs = [np.arange(0, 2 * nx + 1) * (0.5 * w)] * (ny + 1)  # every line: nodes every half brick
        idx, x, y = [], [], []
        for j in range(ny + 1):
            idx.append(len(x) + np.arange(len(xs[j])))
            x.extend(xs[j])
            y.extend([j * h] * len(xs[j]))
        cells = []
        for j in range(ny):
            lo, hi = idx[j], idx[j + 1]
            # brick edges at half-brick indices: even rows 0, 2, 4, ...; odd rows 0, 1, 3, 5, ..., 2 nx
            cuts = list(range(0, 2 * nx + 1, 2)) if j % 2 == 0 else [0] + list(range(1, 2 * nx, 2)) + [2 * nx]
            for a, b in zip(cuts[:-1], cuts[1:]):
                ring = [lo[k] for k in range(a, b + 1)] + [hi[k] for k in range(b, a - 1, -1)]
                cells.append(ring)
        x = np.asarray(x)
        y = np.asarray(y)
        return cls.cells(x, y, np.zeros_like(x), cells)

    @classmethod
    def read(cls, path):
        out = C.c_void_p()
        N.host_check(N.host_lib().sw2dh_geom_read(str(path).encode(), C.byref(out)))
        return cls(out)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.sw2dh_geom_free(self._h)
            self._h = C.c_void_p()

    def sizes(self):
        v = [C.c_int64() for _ in range(4)]
        N.host_check(self.lib.sw2dh_geom_sizes(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def array(self, name):
        n = C.c_int64()
        dt = C.c_int32()
        p = self.lib.sw2dh_geom_array(self._h, name.encode(), C.byref(n), C.byref(dt))
        if p is None and n.value != 0:
            raise KeyError(name)
        return _copy(p, n.value, dt.value)

    def set_node_z(self, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        N.host_check(self.lib.sw2dh_geom_set_node_z(self._h, N.dptr(z)))

    def set_bnd_codes(self, codes):
        codes = np.ascontiguousarray(codes, dtype=np.int32)
        N.host_check(self.lib.sw2dh_geom_set_bnd_codes(self._h, N.i32ptr(codes)))

    def locate(self, x, y):
        """Cell of each point (-1 outside): the reference's winding-angle test
        (swMesh::getCellFromCoordinates, swMesh.cpp:695-759) on any polygon."""
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        y = np.ascontiguousarray(np.atleast_1d(y), dtype=np.float64)
        out = np.empty(len(x), dtype=np.int64)
        N.host_check(self.lib.sw2dh_geom_locate(self._h, len(x), N.dptr(x), N.dptr(y), N.i64ptr(out)))
        return out


def _copy(p, n, dt):
    ctype, np_t = {0: (C.c_double, np.float64), 1: (C.c_int64, np.int64),
                   2: (C.c_int32, np.int32)}[dt]
    if n == 0:
        return np.zeros(0, dtype=np_t)
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), shape=(n,)).copy()


def solv_text(scheme="muscl_evr", topo="discontinuous", friction="none", cfl=0.9, t_max=10.0,
              dt_max=10.0, dt_map=1e6, dt_ts=1e6, h_min=1e-5, n_decimal=17, store_flag="default",
              head_loss=False, extra=None):
    """input.solv text (keys of swSolverParam::readInputFile)."""
    lines = [
        "= SP case", "model_hyperb\tsp", f"model_friction\t{friction}",
        f"num_scheme\t{scheme}", f"topo_discretization\t{topo}",
        f"cfl_max\t{cfl!r}", f"t_max\t{t_max!r}", f"dt_max\t{dt_max!r}",
        f"dt_map\t{dt_map!r}", f"dt_time_series\t{dt_ts!r}", f"h_min\t{h_min!r}",
        f"n_decimal\t{n_decimal}", f"store_flag\t{store_flag}", "output_format\ttxt",
    ]
    if head_loss:
        lines.append("model_singular_head_loss\tactive")
    for k, v in (extra or {}).items():
        lines.append(f"{k}\t{v}")
    return "\n".join(lines) + "\n"


def write_case(folder, geom, solv, h0, qx0, qy0, phi=None, beta=None, friction="none", fric=None,
               bc_codes=(1,), bc_types=("q",), bc_t=(0.0,), bc_v=((0.0,),)):
    """Write a complete SP case folder (sw2dh_write_case)."""
    nc = geom.sizes()[0]
    arr = lambda a: None if a is None else np.ascontiguousarray(np.broadcast_to(a, (nc,)), dtype=np.float64)
    h0, qx0, qy0, phi, beta = arr(h0), arr(qx0), arr(qy0), arr(phi), arr(beta)
    fric = arr(fric)
    codes = np.ascontiguousarray(bc_codes, dtype=np.int32)
    t = np.ascontiguousarray(bc_t, dtype=np.float64)
    v = np.ascontiguousarray(np.asarray(bc_v, dtype=np.float64).reshape(len(codes), len(t)))
    os.makedirs(folder, exist_ok=True)
    N.host_check(N.host_lib().sw2dh_write_case(
        str(folder).encode(), geom._h, solv.encode(), N.dptr(h0), N.dptr(qx0), N.dptr(qy0),
        N.dptr(phi), N.dptr(beta), friction.encode(), N.dptr(fric), len(codes), N.i32ptr(codes),
        " ".join(bc_types).encode(), len(t), N.dptr(t), N.dptr(v)))
    return folder


def dambreak_case(folder, nx=100, ny=100, Lx=200.0, Ly=100.0, x_dam=100.0, h_left=2.0,
                  h_right=0.5, scheme="muscl_evr", topo="discontinuous", friction="manning",
                  n_manning=0.02, t_max=10.0, cfl=0.9, dt_map=1e6, dt_ts=1e6, store_flag="default",
                  wind=False, rain=None, rain_where="water_only", infil=False, mdiff=False):
    """Plain SP forward dam-break (20,000 triangles at the default size).
    With wind=True it adds the Smith & Banke wind source: one station on each
    side of the dam, records at 1/6, 1/2 and 5/6 of t_max; rain / infil add
    rain, evaporation and infiltration as in mixed_case; mdiff turns on
    momentum diffusion (all cells wet: h_right > 0)."""
    g = Geometry.rect(nx, ny, Lx, Ly)
    x = g.array("cell_x")
    h0 = np.where(x < x_dam, h_left, h_right)
    solv = solv_text(scheme=scheme, topo=topo, friction=friction, cfl=cfl, t_max=t_max,
                     dt_max=10.0, dt_map=dt_map, dt_ts=dt_ts, store_flag=store_flag,
                     extra=_source_keys(wind, rain, rain_where, infil, mdiff))
    write_case(folder, g, solv, h0, 0.0, 0.0, friction=friction,
               fric=None if friction == "none" else n_manning,
               bc_codes=(1,), bc_types=("q",), bc_t=(0.0, 1e9), bc_v=((0.0, 0.0),))
    if wind:
        write_distrib_map(os.path.join(folder, "input", "wind_station_codes_map.txt"),
                          ["wind_station"], [np.where(x < x_dam, 1, 2)])
        tw = t_max * np.array([1.0, 3.0, 5.0]) / 6.0
        write_wind_series(os.path.join(folder, "input", "wind_time_series.txt"), tw,
                          np.array([[30.0, 40.0, 25.0], [35.0, 20.0, 45.0]]),
                          np.array([[30.0, 120.0, 250.0], [300.0, 10.0, 170.0]]))
    if rain or infil:
        _write_rain(folder, np.where(x < x_dam, 1, 2), t_max, np.random.default_rng(17), rain, infil)
    return folder


def _smooth_noise(rng, n_grid, sigma_cells):
    """Seeded Gaussian-filtered noise on an n_grid^2 lattice, rescaled to [0, 1]."""
    from numpy.fft import fft2, ifft2, fftfreq
    a = rng.standard_normal((n_grid, n_grid))
    k = fftfreq(n_grid)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    filt = np.exp(-2.0 * (np.pi * sigma_cells) ** 2 * (kx ** 2 + ky ** 2))
    s = np.real(ifft2(fft2(a) * filt))
    return (s - s.min()) / (s.max() - s.min())


def scale_case(folder, px=500, py=500, L=5000.0, seed=1234, t_max=3600.0, scheme="muscl_evr",
               topo="discontinuous", n_manning=0.03, q_peak=500.0, wet_fraction=0.4, cfl=0.9,
               dt_map=1e9, dt_ts=1e9, jitter=0.25, hydrograph=None, manning=None,
               friction="manning"):
    """SP forward at scale: jittered-lattice Delaunay mesh of 2*px*py triangles.

    px = py = 707 gives ~1.0e6 cells.  Returns (folder, info dict).
    """
    rng = np.random.default_rng(seed)
    g = Geometry.delaunay(px, py, L, L, jitter=jitter, seed=seed)
    nx_ = g.array("node_x")
    ny_ = g.array("node_y")
    # bed: slope 1e-3 (down towards the east) + 20 Gaussian bumps
    z = 1e-3 * (L - nx_)
    bumps = []
    for _ in range(20):
        cx, cy = rng.uniform(0, L, 2)
        amp = rng.uniform(0.0, 3.0)
        sig = rng.uniform(50.0, 300.0)
        z = z + amp * np.exp(-((nx_ - cx) ** 2 + (ny_ - cy) ** 2) / (2.0 * sig * sig))
        bumps.append((cx, cy, amp, sig))
    g.set_node_z(z)
    # boundary codes: west edge (x = 0) is the inflow (code 2), the rest walls (code 1)
    bxm = g.array("bnd_xm")
    codes = np.where(bxm < 1e-9 * L, 2, 1).astype(np.int32)
    g.set_bnd_codes(codes)
    zb = g.array("cell_zb")
    cx = g.array("cell_x")
    cy = g.array("cell_y")
    # porosity in [0.3, 1] from smoothed noise sampled at the cell centres
    ng = 256
    noise = _smooth_noise(rng, ng, sigma_cells=6.0)
    ix = np.clip((cx / L * (ng - 1)).round().astype(int), 0, ng - 1)
    iy = np.clip((cy / L * (ng - 1)).round().astype(int), 0, ng - 1)
    phi = 0.3 + 0.7 * noise[ix, iy]
    # partially dry start: flat free surface at the wet_fraction quantile of zb
    z0 = float(np.quantile(zb, wet_fraction))
    h0 = np.maximum(z0 - zb, 0.0)
    # west inflow: triangular 0 -> q_peak -> 0 m3/s over t_max, as unit discharge
    west_len = float(np.sum(g.array("bnd_len")[codes == 2]))
    if hydrograph is None:
        tq = np.array([0.0, 0.5 * t_max, t_max, 10 * t_max + 1e6])
        Q = np.array([0.0, q_peak, 0.0, 0.0])
    else:
        tq, Q = hydrograph
        tq = np.asarray(tq, dtype=np.float64)
        Q = np.asarray(Q, dtype=np.float64)
    qunit = Q / west_len
    fric = n_manning if manning is None else manning
    nc, nb, ni, nn = g.sizes()
    info = {"n_cell": nc, "n_bnd": nb, "n_int": ni, "n_node": nn,
            "flips": getattr(g, "n_flips", 0), "west_len": west_len, "z0": z0,
            "wet_fraction": float(np.mean(h0 > 0)), "geometry": g, "bumps": bumps}
    bc_v = np.stack([np.zeros_like(tq), qunit])
    if folder is None:
        # same inputs, kept in memory (a written case reads back bit-identical: %.17e)
        p = N.Params()
        p.scheme = N.MUSCL_EVR if scheme == "muscl_evr" else N.GODUNOV
        p.topo = N.TOPO_BLENDING if topo == "blending" else N.TOPO_DISCONTINUOUS
        p.friction = {"none": N.FRIC_NONE, "manning": N.FRIC_MANNING,
                      "strickler": N.FRIC_STRICKLER, "chezy": N.FRIC_CHEZY}[friction]
        p.head_loss, p.neg_depth_detection = 0, 0
        p.h_min, p.phi_min, p.cfl_max, p.dt_max = 1e-5, 1e-16, cfl, 30.0
        p.dt_map, p.dt_time_series, p.t0, p.t_max, p.v_max, p.fr_max = dt_map, dt_ts, 0.0, t_max, 100.0, 100.0
        arrays = {k: g.array(k) for k in ("cell_area", "cell_x", "cell_y", "cell_zb", "bnd_cell",
                                          "bnd_len", "bnd_nx", "bnd_ny", "bnd_code", "int_L", "int_R",
                                          "int_len", "int_nx", "int_ny", "int_xm", "int_ym")}
        arrays.update(h0=h0, qx0=np.zeros(nc), qy0=np.zeros(nc), phi=phi, beta=np.ones(nc),
                      fc1=np.full(nc, 1.0) * fric, fc2=np.full(nc, 1.0) * fric, falpha=np.zeros(nc),
                      hsing=np.zeros(0), hk1=np.zeros(0), hk2=np.zeros(0), halpha=np.zeros(0),
                      bnd_series=(codes - 1).astype(np.int32),
                      series_type=np.array([N.BC_Q, N.BC_Q], dtype=np.int32),
                      series_t=tq, series_v=np.ascontiguousarray(bc_v.reshape(-1)))
        return MemCase(arrays, p, n_series=2, n_times=len(tq)), info
    solv = solv_text(scheme=scheme, topo=topo, friction=friction, cfl=cfl, t_max=t_max,
                     dt_max=30.0, dt_map=dt_map, dt_ts=dt_ts)
    write_case(folder, g, solv, h0, 0.0, 0.0, phi=phi, friction=friction, fric=fric,
               bc_codes=(1, 2), bc_types=("q", "q"), bc_t=tq, bc_v=bc_v)
    return folder, info


class MemCase:
    """In-memory stand-in for Case (same attributes) for workloads too large
    to be worth a round trip through text maps."""

    def __init__(self, arrays, params, n_series, n_times):
        self.a = arrays
        self.params = params
        self.n_cell = len(arrays["cell_area"])
        self.n_bnd = len(arrays["bnd_len"])
        self.n_int = len(arrays["int_len"])
        self.n_series, self.n_times = n_series, n_times
        self.folder = None

    def __getitem__(self, k):
        return self.a[k]


def write_distrib_map(path, names, cols):
    """A 'Unif 0' parameter map (layout read by swIO::readVectorParameterFile)."""
    cols = [np.asarray(c, dtype=np.float64) for c in cols]
    with open(path, "w") as f:
        f.write("Unif 0\n==== Default Param\n" + "\t".join(names) + "\n")
        f.write("\t".join("0" for _ in names) + "\n==== Distrib\n")
        for row in zip(*cols):
            f.write("\t".join(f"{v:.17e}" for v in row) + "\n")


def mixed_case(folder, px=12, py=10, L=(60.0, 50.0), seed=7, scheme="muscl_evr",
               topo="discontinuous", friction="strickler", head_loss=True, t_max=20.0,
               dt_map=1e6, dt_ts=1e6, cfl=0.9, store_flag="default", closed_fraction=0.04,
               wind=False, rain=None, rain_where="water_only", infil=False, mesh="delaunay"):
    """Small case that exercises every branch of the SP step: Delaunay mesh,
    bumpy bed with wet/dry fronts, porosity field including closed cells
    (phi = 0), all four boundary types (west q hydrograph, east fr, north h,
    south z) and singular head loss.  (With MUSCL-EVR the reference stops on
    an impossible HLLC regime when a dry closed cell borders a dry cell, so
    MUSCL cases use closed_fraction = 0.)  With wind=True it adds the Smith &
    Banke wind source: two stations (west and east halves) whose records start
    after t = 0 and end before t_max, so both clamps and the interpolation
    run.  rain (a model_runoff value) adds rain and evaporation from two
    stations, rain_where is model_porosity_rainfall, infil adds Horton
    infiltration.  mesh "bricks": the staggered brick wall of px x py bricks
    (Geometry.bricks: interior cells are 6-sided polygons) instead of the
    Delaunay triangles."""
    rng = np.random.default_rng(seed)
    Lx, Ly = L
    g = (Geometry.bricks(px, py, Lx, Ly) if mesh == "bricks"
         else Geometry.delaunay(px, py, Lx, Ly, jitter=0.25, seed=seed))
    nx_, ny_ = g.array("node_x"), g.array("node_y")
    z = 0.02 * (Lx - nx_)
    for _ in range(4):
        cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
        z = z + rng.uniform(0.2, 1.0) * np.exp(-((nx_ - cx) ** 2 + (ny_ - cy) ** 2) / (2 * (0.12 * Lx) ** 2))
    g.set_node_z(z)
    bx, by = g.array("bnd_xm"), g.array("bnd_ym")
    eps = 1e-9 * max(Lx, Ly)
    codes = np.where(bx < eps, 2, np.where(bx > Lx - eps, 3, np.where(by > Ly - eps, 4, 5)))
    g.set_bnd_codes(codes.astype(np.int32))
    zb, cx, cy = g.array("cell_zb"), g.array("cell_x"), g.array("cell_y")
    nc = len(zb)
    phi = np.clip(0.3 + 0.7 * rng.random(nc), 0.0, 1.0)
    phi[rng.random(nc) < closed_fraction] = 0.0
    z0 = float(np.quantile(zb, 0.5))
    h0 = np.maximum(z0 - zb, 0.0)
    h0[phi <= 0.0] = 0.0  # closed cells hold no water
    qx0 = 0.05 * h0 * rng.standard_normal(nc)
    qy0 = 0.05 * h0 * rng.standard_normal(nc)
    fric = (20.0 + 30.0 * rng.random(nc) if friction in ("strickler", "chezy")
            else 0.01 + 0.04 * rng.random(nc))
    tq = np.array([0.0, 5.0, 12.0, 100.0])
    bc_v = np.array([
        [0.0, 0.0, 0.0, 0.0],                 # code 1 unused
        [0.0, 2.0, 0.5, 0.5],                 # west q (unit discharge)
        [0.65, 0.65, 0.65, 0.65],             # east fr
        [0.2, 0.4, 0.1, 0.1],                 # north h
        [z0 - 0.3, z0 + 0.2, z0, z0],         # south z
    ])
    solv = solv_text(scheme=scheme, topo=topo, friction=friction, cfl=cfl, t_max=t_max,
                     dt_max=0.5, dt_map=dt_map, dt_ts=dt_ts, head_loss=head_loss,
                     store_flag=store_flag, extra=_source_keys(wind, rain, rain_where, infil))
    write_case(folder, g, solv, h0, qx0, qy0, phi=phi, friction=friction, fric=fric,
               bc_codes=(1, 2, 3, 4, 5), bc_types=("q", "q", "fr", "h", "z"), bc_t=tq, bc_v=bc_v)
    if head_loss:
        hs = 0.05 * rng.random(nc)
        k1 = np.where(rng.random(nc) < 0.3, 0.1 + rng.random(nc), 0.0)
        k2 = 0.5 * k1
        al = rng.uniform(-30, 30, nc)
        write_distrib_map(os.path.join(folder, "input", "singular_head_loss_map.txt"),
                          ["h_sing", "k1", "k2", "alpha"], [hs, k1, k2, al])
    if wind:
        station = np.where(cx < 0.5 * Lx, 1, 2)
        write_distrib_map(os.path.join(folder, "input", "wind_station_codes_map.txt"),
                          ["wind_station"], [station])
        tw = np.array([2.0, 7.0, 11.0, 0.75 * t_max])
        norm = rng.uniform(15.0, 40.0, (2, len(tw)))
        angle = rng.uniform(0.0, 360.0, (2, len(tw)))
        write_wind_series(os.path.join(folder, "input", "wind_time_series.txt"), tw, norm, angle)
    if rain or infil:
        _write_rain(folder, np.where(cy < 0.5 * Ly, 1, 2), t_max, rng, rain, infil)
    return folder


def _source_keys(wind, rain, rain_where, infil, mdiff=False):
    keys = {}
    if mdiff:
        keys["model_momentum_diffusion"] = "velocity_independent"
    if wind:
        keys["model_wind"] = "smithbanke"
    if rain:
        keys["model_runoff"] = rain
        keys["model_porosity_rainfall"] = rain_where
    if infil:
        keys["model_infiltration"] = "horton"
    return keys or None


def _write_rain(folder, station, t_max, rng, rain, infil):
    """Rain and evaporation from two stations (records at 1/10, 7/20, 11/20
    and 3/4 of t_max, so the runs start before the first and end after the
    last), the runoff-model coefficient map and the Horton rate map."""
    inp = os.path.join(folder, "input")
    nc = len(station)
    if rain:
        write_distrib_map(os.path.join(inp, "precipitation_station_codes_map.txt"),
                          ["precipitation_station"], [station])
        tr = t_max * np.array([0.1, 0.35, 0.55, 0.75])
        rate = np.array([[2e-3, 5e-4, -1e-3, 3e-3], [-5e-4, 1e-3, 2e-3, -2e-3]])
        with open(os.path.join(inp, "precipitation_time_series.txt"), "w") as f:
            f.write("# precipitation time series\n# number of stations\n2\n# rate (m/s)\n# time\t1\t2\n")
            for r in range(len(tr)):
                f.write(f"{tr[r]:.17e}\t{rate[0, r]:.17e}\t{rate[1, r]:.17e}\n")
        coeff = rng.uniform(0.0, 1.0, nc)
        if rain == "infiltration_coefficient":
            write_distrib_map(os.path.join(inp, "infiltration_coefficient_map.txt"), ["coeff"], [coeff])
        elif rain == "runoff_coefficient":
            write_distrib_map(os.path.join(inp, "runoff_coefficient_map.txt"), ["coeff"], [coeff])
    if infil:
        write_distrib_map(os.path.join(inp, "infiltration_horton_parameter_map.txt"), ["rate"],
                          [rng.uniform(0.0, 5e-4, nc)])


def write_wind_series(path, t, norm, angle_deg):
    """wind_time_series.txt (layout read by swIO::readTimeSeries2dVectorFile,
    src/common/swIO.cpp:906-1000): per record the time, then (norm, angle
    from North in degrees) for each station."""
    norm, angle_deg = np.atleast_2d(norm), np.atleast_2d(angle_deg)
    with open(path, "w") as f:
        f.write(f"# wind time series\n# number of stations\n{norm.shape[0]}\n")
        f.write("# records\n# time" + "".join(f"\tnorm_{k + 1}\tangle_{k + 1}" for k in range(norm.shape[0])) + "\n")
        for r in range(len(t)):
            f.write(f"{t[r]:.17e}" + "".join(f"\t{norm[k, r]:.17e}\t{angle_deg[k, r]:.17e}"
                                          for k in range(norm.shape[0])) + "\n")


class Case:
    """A loaded case folder (sw2dh_case): arrays in the reference's ordering."""

    ARRAYS = ("cell_area", "cell_x", "cell_y", "cell_zb", "bnd_cell", "bnd_len", "bnd_nx",
              "bnd_ny", "bnd_series", "bnd_code", "int_L", "int_R", "int_len", "int_nx", "int_ny",
              "int_xm", "int_ym", "h0", "qx0", "qy0", "phi", "beta", "fc1", "fc2", "falpha",
              "hsing", "hk1", "hk2", "halpha", "series_type", "series_t", "series_v",
              "wind_code", "wind_t", "wind_x", "wind_y", "rain_code", "rain_t", "rain_v",
              "runoff_coeff", "infil_rate", "source_models", "probe_cell", "probe_x", "probe_y")

    def __init__(self, folder):
        self.folder = str(folder)
        lib = N.host_lib()
        h = C.c_void_p()
        N.host_check(lib.sw2dh_case_load(self.folder.encode(), C.byref(h)))
        try:
            p = N.Params()
            N.host_check(lib.sw2dh_case_params(h, C.byref(p)))
            self.params = p
            nc, nb, ni = C.c_int64(), C.c_int64(), C.c_int64()
            ns, nt = C.c_int32(), C.c_int32()
            N.host_check(lib.sw2dh_case_sizes(h, C.byref(nc), C.byref(nb), C.byref(ni),
                                              C.byref(ns), C.byref(nt)))
            self.n_cell, self.n_bnd, self.n_int = nc.value, nb.value, ni.value
            self.n_series, self.n_times = ns.value, nt.value
            self.a = {}
            for name in self.ARRAYS:
                n = C.c_int64()
                dt = C.c_int32()
                ptr = lib.sw2dh_case_array(h, name.encode(), C.byref(n), C.byref(dt))
                self.a[name] = _copy(ptr, n.value, dt.value)
            self.n_decimal = lib.sw2dh_case_n_decimal(h)
            self.probe_names = [lib.sw2dh_case_probe_name(h, i).decode()
                                for i in range(len(self.a["probe_cell"]))]
            self.keys = {}
            for k in ("store_flag", "checkpoint_stride", "cost_normalize"):
                v = lib.sw2dh_case_key(h, k.encode())
                if v is not None:
                    self.keys[k] = v.decode()
        finally:
            lib.sw2dh_case_free(h)

    def __getitem__(self, k):
        return self.a[k]

    def write_result(self, path, h, qx, qy, n_decimal=None, mode="default"):
        """ResultHydro map of a state (reference order), as the solver writes
        it (swIO::createTxtFileMapDecimal, swIO.cpp:182-223)."""
        lib = N.host_lib()
        hd = C.c_void_p()
        N.host_check(lib.sw2dh_case_load(self.folder.encode(), C.byref(hd)))
        try:
            arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (h, qx, qy)]
            nd = self.n_decimal if n_decimal is None else int(n_decimal)
            N.host_check(lib.sw2dh_write_result(hd, str(path).encode(), *[N.dptr(a) for a in arr], nd,
                                                mode.encode()))
        finally:
            lib.sw2dh_case_free(hd)


def result_name(t):
    buf = C.create_string_buffer(128)
    N.host_check(N.host_lib().sw2dh_result_name(float(t), buf, 128))
    return buf.value.decode()


def read_result_map(path):
    """Read a ResultHydro_p*.txt map: returns dict column -> float64 array."""
    with open(path) as f:
        names = f.readline().split()
    data = np.loadtxt(path, skiprows=1, ndmin=2)
    return {n: data[:, i] for i, n in enumerate(names)}
