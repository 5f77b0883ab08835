"""Host-side scene construction: material-point domains, grids and worlds.

Domain creation stays on the host (it is not part of the GPU step).  The
functions here restate the reference factories so that the points they emit
have the same positions and the same order:

* ``create_domain_auxetic_cell_polygon`` --
  MaterialPoint_Factory_createDomain_AuxeticCell_Polygon.cpp:4-112
* ``create_domain_cuboid`` -- MaterialPoint_Factory_createDomain_Cuboid.cpp:4-57
* ``create_domain_auxetic_cell_swisscheese`` --
  MaterialPoint_Factory_createDomain_AuxeticCell_Swisscheese.cpp:4-101
* ``is_inside_polygon`` / ``is_inside_ellipse`` -- MaterialPoint_Factory_isInside.cpp:4-35
* ``reference_polygon_world`` -- PhysicsEngine_initializeWorld_AuxeticPolygonCell.cpp:4-313
* ``reference_swisscheese_world`` -- PhysicsEngine_initializeWorld_AuxeticSwisscheeseCell.cpp:4-283
* ``describe`` -- the world description both initializers report
* ``from_config`` -- any of the worlds here by name, JSON file or dict

Loop coordinates are accumulated sequentially (``dx += dOffset``) exactly as
the reference loops do, so the FP64 positions match bit for bit.
"""
import math
from dataclasses import dataclass, field

import numpy as np

ELASTIC, NEOHOOKEAN, VISCOELASTIC, PLASTIC, VONMISESHARDENING, GASS = range(6)
FIX_X, FIX_Y, FIX_Z = 1, 2, 4


# ---------------------------------------------------------------- factories
def _accumulate(start, step, stop):
    vals = []
    d = start
    while d <= stop:
        vals.append(d)
        d += step
    return np.array(vals, dtype=np.float64)


def is_inside_polygon(coords, verts):
    """Vectorised MaterialPoint_Factory::isInside(coord, vertices) (isInside.cpp:4-21).

    coords: (n, 2) x/y; verts: list of (x, y, z).  True when the z component of
    cross(v[i+1]-v[i], p-v[i]) is >= 0 for every edge.
    """
    inside = np.ones(len(coords), dtype=bool)
    for i in range(len(verts) - 1):
        v1x = verts[i + 1][0] - verts[i][0]
        v1y = verts[i + 1][1] - verts[i][1]
        v2x = coords[:, 0] - verts[i][0]
        v2y = coords[:, 1] - verts[i][1]
        cz = v1x * v2y - v2x * v1y          # glm::cross(a, b).z = a.x*b.y - b.x*a.y
        inside &= ~(cz < 0.0)
    return inside


def is_inside_ellipse(coords, center, radii):
    """MaterialPoint_Factory::isInside(coord, center, radii) (isInside.cpp:23-35)."""
    lx = coords[:, 0] - center[0]
    ly = coords[:, 1] - center[1]
    a = lx / radii[0]
    b = ly / radii[1]
    return ~((a * a + b * b) > 1.0)


_MIRROR_POLY = [(+1, +1, +1), (-1, +1, +1), (+1, -1, +1), (-1, -1, +1),
                (+1, +1, -1), (-1, +1, -1), (+1, -1, -1), (-1, -1, -1)]
_MIRROR_CUBOID = [(+1, +1, +1), (-1, +1, +1), (-1, -1, +1), (+1, -1, +1),
                  (+1, +1, -1), (-1, +1, -1), (-1, -1, -1), (+1, -1, -1)]


def _emit(center, q, mirrors):
    """For each quarter coordinate (in loop order) emit the mirrored points in
    the reference order: result shape (len(q)*len(mirrors), 3)."""
    out = np.empty((len(q), len(mirrors), 3), dtype=np.float64)
    for k, (sx, sy, sz) in enumerate(mirrors):
        out[:, k, 0] = center[0] + sx * q[:, 0]
        out[:, k, 1] = center[1] + sy * q[:, 1]
        out[:, k, 2] = center[2] + sz * q[:, 2]
    return out.reshape(-1, 3)


def _quarter_grid(xs, ys, zs):
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")   # x outer, z inner
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def polygon_cell_vertices(dims, dent, thickness):
    """Quarter-cell polygons of createDomain_AuxeticCell_Polygon (…Polygon.cpp:8-39)."""
    Dx, Dy = dims[0], dims[1]
    l_inc = math.sqrt((0.5 * Dx - 2.0 * dent) ** 2 + (0.5 * Dy) ** 2)
    s = 0.5 * Dy / l_inc
    c = math.sqrt(1.0 - s ** 2)
    t = s / c
    total = [(0.0, 0.0, 0.0), (0.5 * Dx, 0.0, 0.0), (0.5 * Dx, 0.5 * Dy, 0.0), (0.0, 0.5 * Dy, 0.0), (0.0, 0.0, 0.0)]
    right = [(0.5 * Dx, 0.5 * thickness, 0.0), (0.5 * Dx, 0.5 * Dy, 0.0),
             (0.5 * Dx - dent + 0.5 * thickness / s, 0.5 * Dy, 0.0),
             (dent + 0.5 * thickness / s + 0.5 * thickness / t, 0.5 * thickness, 0.0),
             (0.5 * Dx, 0.5 * thickness, 0.0)]
    left = [(0.0, 0.0, 0.0), (dent - 0.5 * thickness / s, 0.0, 0.0),
            (0.5 * Dx - dent - 0.5 * thickness / s - 0.5 * thickness / t, 0.5 * Dy - 0.5 * thickness, 0.0),
            (0.0, 0.5 * Dy - 0.5 * thickness, 0.0), (0.0, 0.0, 0.0)]
    return total, right, left


def create_domain_auxetic_cell_polygon(center, dims, dent, thickness, offset):
    """MaterialPoint_Factory::createDomain_AuxeticCell_Polygon (…Polygon.cpp:4-112)."""
    total, right, left = polygon_cell_vertices(dims, dent, thickness)
    xs = _accumulate(0.5 * offset, offset, 0.5 * dims[0])
    ys = _accumulate(0.5 * offset, offset, 0.5 * dims[1])
    zs = _accumulate(0.5 * offset, offset, 0.5 * dims[2])
    q = _quarter_grid(xs, ys, zs)
    keep = is_inside_polygon(q, total) & ~is_inside_polygon(q, right) & ~is_inside_polygon(q, left)
    return _emit(center, q[keep], _MIRROR_POLY)


def create_domain_cuboid(center, dims, offset):
    """MaterialPoint_Factory::createDomain_Cuboid (…Cuboid.cpp:4-57)."""
    single = dims[2] <= offset
    xs = _accumulate(0.5 * offset, offset, 0.5 * dims[0])
    ys = _accumulate(0.5 * offset, offset, 0.5 * dims[1])
    zs = _accumulate(0.0 if single else 0.5 * offset, offset, 0.5 * dims[2])
    q = _quarter_grid(xs, ys, zs)
    return _emit(center, q, _MIRROR_CUBOID[:4] if single else _MIRROR_CUBOID)


def create_domain_auxetic_cell_swisscheese(center, dims, spacing, radii, offset):
    """MaterialPoint_Factory::createDomain_AuxeticCell_Swisscheese (…Swisscheese.cpp:4-101)."""
    centers, rad = [], []
    ix = 0
    while ix < dims[0] / spacing[0]:
        iy = 0
        while iy < dims[1] / spacing[1] - 2:
            centers.append((ix * spacing[0], iy * spacing[1]))
            rad.append(tuple(radii) if (ix + iy) % 2 == 0 else (radii[1], radii[0]))
            iy += 1
        ix += 1
    single = dims[2] <= offset
    xs = _accumulate(0.5 * offset, offset, 0.5 * dims[0])
    ys = _accumulate(0.5 * offset, offset, 0.5 * dims[1])
    zs = _accumulate(0.0 if single else 0.5 * offset, offset, 0.5 * dims[2])
    q = _quarter_grid(xs, ys, zs)
    keep = np.ones(len(q), dtype=bool)
    for c, r in zip(centers, rad):
        keep &= ~is_inside_ellipse(q, c, r)
    return _emit(center, q[keep], _MIRROR_POLY[:4] if single else _MIRROR_POLY)


# -------------------------------------------------------------------- scene
@dataclass
class Scene:
    """Everything an engine needs: grid, node BCs, points, solver parameters."""
    length: tuple
    cells: tuple
    fixed: np.ndarray                      # uint8 per node, FIX_* bits
    x: np.ndarray                          # (n, 3)
    mass: np.ndarray
    vol0: np.ndarray
    E: np.ndarray
    nu: np.ndarray
    sy: np.ndarray
    C0: np.ndarray
    C1: np.ndarray
    type: np.ndarray                       # int32
    dc: np.ndarray                         # uint8
    v: np.ndarray = None
    fext: np.ndarray = None
    surface: np.ndarray = None
    dt_max: float = 1e-9
    time_end: float = 1.0
    damping: float = 0.0
    mass_min: float = 0.0
    console_interval: float = 0.0
    timeline_t: np.ndarray = None
    timeline_v: np.ndarray = None
    neohookean: int = 0
    node_group: np.ndarray = None          # uint16 per node (batched worlds), or None
    n_groups: int = 0
    info: dict = field(default_factory=dict)

    def __post_init__(self):
        n = len(self.x)
        if self.v is None:
            self.v = np.zeros((n, 3))
        if self.fext is None:
            self.fext = np.zeros((n, 3))
        if self.surface is None:
            self.surface = np.zeros(n, dtype=np.uint8)
        if self.timeline_t is None:
            self.timeline_t = np.zeros(0)
            self.timeline_v = np.zeros((0, 3))

    @property
    def n(self):
        return len(self.x)

    @property
    def n_deformable(self):
        return int(np.count_nonzero(self.dc == 0))

    @property
    def nodes(self):
        return tuple(c + 1 for c in self.cells)

    @property
    def h(self):
        return tuple(self.length[a] / self.cells[a] for a in range(3))

    def copy(self):
        import copy
        return copy.deepcopy(self)


def node_positions(length, cells):
    """GridPoint_Factory::createGrid node positions (GridPoint_Factory.h:33-67),
    x fastest: returns (nx*ny*nz, 3)."""
    nn = [c + 1 for c in cells]
    ax = [np.arange(nn[a], dtype=np.float64) * length[a] / cells[a] for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def reference_bcs(length, cells, tol_frac=0.1):
    """Node BCs of initializeWorld_AuxeticPolygonCell (…PolygonCell.cpp:42-79):
    x=0 / x=L fixed in x, y=0 fixed in x,y,z, z=0 / z=L fixed in z."""
    p = node_positions(length, cells)
    h = length[0] / cells[0]
    tol = tol_frac * h
    f = np.zeros(len(p), dtype=np.uint8)
    f[np.abs(p[:, 0] - 0.0) < tol] |= FIX_X
    f[np.abs(p[:, 0] - length[0]) < tol] |= FIX_X
    f[np.abs(p[:, 1] - 0.0) < tol] |= FIX_X | FIX_Y | FIX_Z
    f[np.abs(p[:, 2] - 0.0) < tol] |= FIX_Z
    f[np.abs(p[:, 2] - length[2]) < tol] |= FIX_Z
    return f


def _material_arrays(n, E, nu, sy, C0, C1, mtype, rho, vol):
    return dict(
        mass=np.full(n, rho * vol), vol0=np.full(n, vol), E=np.full(n, E), nu=np.full(n, nu),
        sy=np.full(n, sy), C0=np.full(n, C0), C1=np.full(n, C1), type=np.full(n, mtype, dtype=np.int32))


def _assemble(length, cells, fixed, bodies, **params):
    xs, cols, dcs = [], {k: [] for k in ("mass", "vol0", "E", "nu", "sy", "C0", "C1", "type")}, []
    for pts, mat, dc in bodies:
        xs.append(pts)
        for k in cols:
            cols[k].append(mat[k])
        dcs.append(np.full(len(pts), 1 if dc else 0, dtype=np.uint8))
    return Scene(length=tuple(length), cells=tuple(cells), fixed=fixed, x=np.concatenate(xs),
                 dc=np.concatenate(dcs), **{k: np.concatenate(v) for k, v in cols.items()}, **params)


def p_wave_speed(E, nu, rho):
    """c = sqrt((lambda + 2 mu) / rho)."""
    return math.sqrt(E * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu)) / rho)


# ------------------------------------------------------------------- worlds
def reference_polygon_world():
    """PhysicsEngine::initializeWorld_AuxeticPolygonCell (…PolygonCell.cpp:4-313)."""
    L = (0.06, 0.03, 0.002)
    cells = (120, 60, 2)
    h = tuple(L[a] / cells[a] for a in range(3))
    fixed = reference_bcs(L, cells)
    off = 1.0 / 2.0 * h[0]
    thick = 2.0 * off
    dims_cell = (0.050346, 0.02, thick)
    center = [0.5 * L[0], 0.5 * L[1], 0.5 * L[2]]
    center[1] = 0.5 * dims_cell[1] + 1.0 * h[1]
    center[2] = 0.5 * thick
    dent = 0.166666666 * dims_cell[0]
    cell = create_domain_auxetic_cell_polygon(center, dims_cell, dent, 1.0 * 0.0006, off)
    vol = off * off * off
    al = _material_arrays(len(cell), 70.0e9, 0.33, 1.0 * 150.0e6, 4.0e1, 50.0e6, PLASTIC, 2700.0, vol)
    top_c = [center[0], center[1] + 0.5 * dims_cell[1] + 0.5 * h[1], center[2]]
    top = create_domain_cuboid(top_c, (dims_cell[0], 2.0 * off, dims_cell[2]), off)
    steel_t = _material_arrays(len(top), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    bot_c = [center[0], 0.5 * h[1], center[2]]
    bot = create_domain_cuboid(bot_c, (dims_cell[0], 2.0 * off, dims_cell[2]), off)
    steel_b = _material_arrays(len(bot), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    sc = _assemble(L, cells, fixed, [(cell, al, False), (top, steel_t, True), (bot, steel_b, False)],
                   dt_max=1.0e-9, time_end=0.010, damping=0.001, mass_min=0.001 * 2700.0 * vol,
                   console_interval=1.0e-6,
                   timeline_t=np.array([0.0, 0.1]), timeline_v=np.array([[0.0, -1.0, 0.0], [0.0, -1.0, 0.0]]))
    sc.info = dict(name="reference_polygon_world", offset=off)
    return sc


def reference_swisscheese_world():
    """PhysicsEngine::initializeWorld_AuxeticSwisscheeseCell
    (…SwisscheeseCell.cpp:4-283): a 50 x 50 mm swiss-cheese cell (elliptical
    holes 7.5 x 5.5 mm on a 14 mm pitch), perfectly plastic aluminium,
    between a displacement-controlled top platen and a free bottom platen;
    the top and bottom node planes fixed, the TimeLine ramping the platen to
    1 m/s over 10 us."""
    L = (0.06, 0.06, 0.001)
    cells = (120, 120, 2)
    h = tuple(L[a] / cells[a] for a in range(3))
    p = node_positions(L, cells)
    tol = 0.1 * h[0]
    fixed = np.zeros(len(p), dtype=np.uint8)
    fixed[np.abs(p[:, 0] - 0.0) < tol] |= FIX_X
    fixed[np.abs(p[:, 0] - L[0]) < tol] |= FIX_X
    fixed[np.abs(p[:, 1] - 0.0) < tol] |= FIX_X | FIX_Y | FIX_Z
    fixed[np.abs(p[:, 1] - L[1]) < tol] |= FIX_X | FIX_Y | FIX_Z
    fixed[np.abs(p[:, 2] - 0.0) < tol] |= FIX_Z
    off = 1.0 / 2.0 * h[0]
    thick = 2.0 * off
    dims_cell = (0.05, 0.05, thick)
    center = [0.5 * L[0], 0.5 * dims_cell[1] + 1.5 * h[1], 0.5 * thick]
    cell = create_domain_auxetic_cell_swisscheese(center, dims_cell, (0.014, 0.014), (0.0075, 0.0055), off)
    vol = off * off * off
    al = _material_arrays(len(cell), 70.0e9, 0.33, 280.0e6, 4.0e1, 50.0e6, PLASTIC, 2700.0, vol)
    plat_dims = (1.1 * dims_cell[0], 2.0 * off, dims_cell[2])
    top = create_domain_cuboid([center[0], center[1] + 0.5 * dims_cell[1] + 1.0 * h[1], center[2]], plat_dims, off)
    steel_t = _material_arrays(len(top), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    bot = create_domain_cuboid([center[0], center[1] - 0.5 * dims_cell[1] - 1.0 * h[1], center[2]], plat_dims, off)
    steel_b = _material_arrays(len(bot), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    sc = _assemble(L, cells, fixed, [(cell, al, False), (top, steel_t, True), (bot, steel_b, False)],
                   dt_max=2.0e-9, time_end=1.0e-3, damping=0.0, mass_min=0.001 * 2700.0 * vol,
                   console_interval=1.0e-6, timeline_t=np.array([0.0, 1.0e-5, 1.0e6]),
                   timeline_v=np.array([[0.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, -1.0, 0.0]]))
    sc.info = dict(name="reference_swisscheese_world", offset=off, describe=dict(threads=False, timeline_probe=1.0))
    return sc


def _script(v, p):
    """Script(double, p) of Definitions.h:43-56: scientific, explicit sign."""
    return f"{v:+.{p}e}"


def describe(sc, started=None):
    """The world description initializeWorld_* hands to reportConsole
    (…PolygonCell.cpp:275-309, …SwisscheeseCell.cpp:253-280); ``started`` is
    the "Process started on" time (a ``time.struct_time``; now by default).
    Pass it to ``PhysicsEngine.reportConsole`` for the first log line."""
    import time as _time
    fmt = sc.info.get("describe", dict(threads=True, timeline_probe=1.0e-4))
    started = _time.localtime() if started is None else started
    bar = "-" * 61 + "\n"
    mass = 0.0
    for m in np.asarray(sc.mass, dtype=np.float64).tolist():   # in point order, as the reference loop
        mass += m
    tl = np.asarray(sc.timeline_t, dtype=np.float64)
    tv = np.asarray(sc.timeline_v, dtype=np.float64).reshape(-1, 3)
    v_probe = 0.0
    for i in range(len(tl) - 1):   # TimeLine::getVelocity (TimeLine.h:33-46)
        t = fmt["timeline_probe"]
        if tl[i] <= t < tl[i + 1]:
            w = (t - tl[i]) / (tl[i + 1] - tl[i])
            v_probe = (1.0 - w) * tv[i, 1] + w * tv[i + 1, 1]
    out = bar + bar + "Process started on: " + _time.strftime("%d-%m-%Y %H:%M:%S", started) + "\n" + bar
    if fmt["threads"]:
        out += "Number of threads: 1\n"
    out += "Time increment: " + _script(sc.dt_max, 6) + "\n"
    out += "Material Point count: " + str(sc.n) + "\n"
    out += "Mass: " + _script(mass, 6) + "\n" + bar
    out += "Grid Resolution: (%d,%d,%d)\n" % tuple(sc.cells)
    out += "Kernel Resolution: (%d,%d,%d)\n" % tuple(sc.cells)
    out += "Material Point offset: " + _script(sc.info.get("offset", 0.5 * sc.h[0]), 3) + "\n"
    out += "Timeline Speed: " + _script(v_probe, 3) + " m/s\n"
    out += "Yield: " + _script(float(sc.sy[0]), 3) + " N/m^2\n"
    out += "Modulus: " + _script(float(sc.E[0]), 3) + " N/m^2\n"
    out += "Global Damping: " + _script(sc.damping, 3) + "\n"
    out += "Non-slip contact\n"
    return out


def _grid_for(extent, h):
    cells = tuple(int(math.ceil(extent[a] / h[a] - 1e-9)) for a in range(3))
    return tuple(cells[a] * h[a] for a in range(3)), cells


def polygon_mesh_world(nx=1, ny=1, cell_w=0.010, cell_h=0.010, wall=0.0008, thickness=0.001, h=0.000125,
                       material="elastic", E=1.0e9, nu=0.3, rho=1200.0, sy=250.0e6, H=700.0e6,
                       platen_speed=0.1, strain_target=None, steps=2000, cfl=0.3, margin_cells=2,
                       damping=0.001, dent_frac=1.0 / 6.0, ramp_steps=0):
    """A tiled mesh of reference polygon cells compressed by a displacement-
    controlled top platen; bottom node plane fixed.

    Each unit cell is createDomain_AuxeticCell_Polygon with dent = W/6 (as
    …PolygonCell.cpp:103), points at 2x2x2 per grid cell (offset h/2,
    …PolygonCell.cpp:89).  The platen is a createDomain_Cuboid slab of two
    point layers on top of the mesh, moving at -platen_speed in y; with
    ``strain_target`` the speed is set so the platen travels
    strain_target * mesh height in ``steps`` steps of dt = cfl*h/c.  With
    ``ramp_steps`` the TimeLine ramps the platen from rest over that many steps.
    """
    hx = h
    off = 0.5 * hx
    mesh_w, mesh_h = nx * cell_w, ny * cell_h
    extent = (mesh_w + 2 * margin_cells * hx, mesh_h + (2 + margin_cells) * hx, thickness)
    L, cells = _grid_for(extent, (hx, hx, hx))
    L = (L[0], L[1], thickness)
    cells = (cells[0], cells[1], max(1, int(round(thickness / hx))))
    x0 = 0.5 * (L[0] - mesh_w)
    pts = []
    dims = (cell_w, cell_h, thickness)
    for j in range(ny):
        for i in range(nx):
            c = (x0 + (i + 0.5) * cell_w, (j + 0.5) * cell_h, 0.5 * thickness)
            pts.append(create_domain_auxetic_cell_polygon(c, dims, dent_frac * cell_w, wall, off))
    body = np.concatenate(pts)
    vol = off ** 3
    if material == "elastic":
        mat = _material_arrays(len(body), E, nu, 0.0, 0.0, 0.0, ELASTIC, rho, vol)
    elif material == "j2":
        C1 = 1.0e15
        mat = _material_arrays(len(body), E, nu, sy, H / C1, C1, VONMISESHARDENING, rho, vol)
    elif material == "plastic":
        mat = _material_arrays(len(body), E, nu, sy, 0.0, 0.0, PLASTIC, rho, vol)
    elif material == "neohookean":
        mat = _material_arrays(len(body), E, nu, 0.0, 0.0, 0.0, NEOHOOKEAN, rho, vol)
    else:
        raise ValueError(material)
    top_c = (0.5 * L[0], mesh_h + 0.5 * hx, 0.5 * thickness)
    top = create_domain_cuboid(top_c, (mesh_w, 2.0 * off, thickness), off)
    plat = _material_arrays(len(top), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    fixed = reference_bcs(L, cells)
    c = p_wave_speed(E, nu, rho)
    dt = cfl * hx / c
    speed = platen_speed
    if strain_target is not None:
        speed = strain_target * mesh_h / (steps * dt)
    if ramp_steps:
        tl_t = np.array([0.0, ramp_steps * dt, 1.0e3])
        tl_v = np.array([[0.0, 0.0, 0.0], [0.0, -speed, 0.0], [0.0, -speed, 0.0]])
    else:
        tl_t = np.array([0.0, 1.0e3])
        tl_v = np.array([[0.0, -speed, 0.0], [0.0, -speed, 0.0]])
    sc = _assemble(L, cells, fixed, [(body, mat, False), (top, plat, True)],
                   dt_max=dt, time_end=steps * dt, damping=damping, mass_min=0.001 * rho * vol,
                   timeline_t=tl_t, timeline_v=tl_v,
                   neohookean=1 if material == "neohookean" else 0)
    sc.info = dict(name="polygon_mesh", nx=nx, ny=ny, h=hx, dt=dt, wave_speed=c, platen_speed=speed,
                   steps=steps, material=material, cell_w=cell_w, cell_h=cell_h, wall=wall)
    return sc


def config1_world(platen="literal", steps=2000):
    """one 10 x 10 x 1 mm re-entrant polygon cell, wall
    0.8 mm, h = 0.125 mm, linear elastic E = 1 GPa, nu = 0.3, rho = 1200."""
    return polygon_mesh_world(1, 1, 0.010, 0.010, 0.0008, 0.001, 0.000125, "elastic", 1.0e9, 0.3, 1200.0,
                              platen_speed=0.1, strain_target=(0.2 if platen == "strain" else None),
                              steps=steps)


def config2_world(material="j2", scale=0.21, steps=5000, platen="ramp", speed=1.0):
    """10 x 10 tiling of the polygon cell at h = 0.1 mm,
    J2 hardening E = 70 GPa, nu = 0.33, sigma_y = 250 MPa, H = 700 MPa.
    The config-1 cell (10 mm, wall 0.8 mm) is scaled by ``scale`` in-plane
    (thickness 1 mm kept) so the mesh holds 1.0M points.

    platen="ramp" (default): ``speed`` m/s reached over 100 steps.  The
    literal "30% strain in 5000 steps" (platen="strain", ~250 m/s) and even
    2 m/s blow up under the reference's explicit update within a few hundred
    steps (the FP64 CPU oracle does too), so the benchmark scene is driven at
    1 m/s; platen="literal" uses the stated 0.1 m/s."""
    if platen == "strain":
        return polygon_mesh_world(10, 10, 0.010 * scale, 0.010 * scale, 0.0008 * scale, 0.001, 0.0001, material,
                                  70.0e9, 0.33, 2700.0, sy=250.0e6, H=700.0e6, strain_target=0.3, steps=steps)
    return polygon_mesh_world(10, 10, 0.010 * scale, 0.010 * scale, 0.0008 * scale, 0.001, 0.0001, material,
                              70.0e9, 0.33, 2700.0, sy=250.0e6, H=700.0e6,
                              platen_speed=(0.1 if platen == "literal" else speed), steps=steps,
                              ramp_steps=(100 if platen == "ramp" else 0))


def small_world(nx=1, ny=1, material="elastic", h=0.0005, cell=0.004, wall=0.0008, thickness=0.001,
                steps=100, speed=2.0):
    """A few-thousand-point polygon world for quick parity tests.  Under the
    reference update the elastic variant stays bounded for ~150 steps at
    5 m/s and the J2 variant for > 1500 steps at 2 m/s."""
    return polygon_mesh_world(nx, ny, cell, cell, wall, thickness, h, material, 1.0e9 if material != "j2" else 70e9,
                              0.3, 1200.0, sy=5.0e6, H=50.0e6, platen_speed=speed, steps=steps)


def translation_world(v=(31.7, -19.3, 10.9), h=0.001, block=(4, 4, 2), cells=(40, 40, 20), dt=1.0e-6, seed=5):
    """A stress-free elastic block in free flight through the grid (no BCs,
    no platen), crossing many cells and bricks so that it exercises
    rebinning.  Positions are jittered and the velocity is not commensurate
    with the lattice: points sitting exactly on a node plane hit the
    reference basis' sign(0) / trunc discontinuity (GridPoint_Mediator.h:35,85),
    where a one-ulp difference changes the gradient."""
    L = tuple(c * h for c in cells)
    off = h / 2
    o = (6, 16, 4)
    ax = [[(o[a] + i) * h + (k + 0.5) * off for i in range(block[a]) for k in range(2)] for a in range(3)]
    X = np.array([(x, y, z) for x in ax[0] for y in ax[1] for z in ax[2]])
    rng = np.random.default_rng(seed)
    X = X + rng.uniform(-0.2, 0.2, X.shape) * off
    n = len(X)
    vol = off ** 3
    mat = _material_arrays(n, 1.0e7, 0.3, 0.0, 0.0, 0.0, ELASTIC, 1000.0, vol)
    sc = Scene(length=L, cells=cells, fixed=np.zeros(int(np.prod([c + 1 for c in cells])), np.uint8), x=X,
               dc=np.zeros(n, np.uint8), v=np.tile(np.asarray(v, float), (n, 1)), dt_max=dt, time_end=1.0,
               damping=0.0, mass_min=0.0, **mat)
    sc.info = dict(name="translation")
    return sc


# ------------------------------------------------------------ batched worlds
def splitmix64(seed):
    """SplitMix64 (Steele, Lea and Flood, 2014): an infinite generator of
    uniform doubles in [0, 1) from a 64-bit seed (the same recurrence in C++
    and Python, so a seed fixes a batch everywhere)."""
    m = (1 << 64) - 1
    state = seed & m
    while True:
        state = (state + 0x9E3779B97F4A7C15) & m
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        z ^= z >> 31
        yield (z >> 11) * (1.0 / (1 << 53))


def dent_for_angle(width, height, angle_deg):
    """Dent of the polygon cell whose inclined walls make ``angle_deg`` with
    the horizontal: tan(angle) = (H/2) / (W/2 - 2 dent)."""
    return 0.5 * (0.5 * width - 0.5 * height / math.tan(math.radians(angle_deg)))


def swisscheese_cell_world(width, wall, thickness, h, material="elastic", E=1.0e9, nu=0.3, rho=1200.0,
                           sy=250.0e6, H=700.0e6, platen_speed=0.1, steps=2000, cfl=0.3, margin_cells=2,
                           damping=0.001):
    """One square swiss-cheese cell (createDomain_AuxeticCell_Swisscheese with
    elliptic holes on a W/4 pitch, ligaments ``wall`` wide) between a fixed
    bottom node plane and a top platen, as polygon_mesh_world builds it."""
    off = 0.5 * h
    extent = (width + 2 * margin_cells * h, width + (2 + margin_cells) * h, thickness)
    L, cells = _grid_for(extent, (h, h, h))
    L = (L[0], L[1], thickness)
    cells = (cells[0], cells[1], max(1, int(round(thickness / h))))
    pitch = 0.25 * width
    r = 0.5 * (pitch - wall)
    body = create_domain_auxetic_cell_swisscheese((0.5 * L[0], 0.5 * width, 0.5 * thickness),
                                                  (width, width, thickness), (pitch, pitch), (1.3 * r, 0.7 * r), off)
    return _cell_world(L, cells, body, width, width, thickness, h, material, E, nu, rho, sy, H, platen_speed, steps,
                       cfl, damping, "swisscheese")


def _cell_world(L, cells, body, mesh_w, mesh_h, thickness, h, material, E, nu, rho, sy, H, platen_speed, steps,
                cfl, damping, name):
    off = 0.5 * h
    vol = off ** 3
    if material == "elastic":
        mat = _material_arrays(len(body), E, nu, 0.0, 0.0, 0.0, ELASTIC, rho, vol)
    else:
        C1 = 1.0e15
        mat = _material_arrays(len(body), E, nu, sy, H / C1, C1, VONMISESHARDENING, rho, vol)
    top = create_domain_cuboid((0.5 * L[0], mesh_h + 0.5 * h, 0.5 * thickness), (mesh_w, 2.0 * off, thickness), off)
    plat = _material_arrays(len(top), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    c = p_wave_speed(E, nu, rho)
    dt = cfl * h / c
    sc = _assemble(L, cells, reference_bcs(L, cells), [(body, mat, False), (top, plat, True)], dt_max=dt,
                   time_end=steps * dt, damping=damping, mass_min=0.001 * rho * vol,
                   timeline_t=np.array([0.0, 1.0e3]),
                   timeline_v=np.array([[0.0, -platen_speed, 0.0], [0.0, -platen_speed, 0.0]]))
    sc.info = dict(name=name, h=h, dt=dt, wave_speed=c, platen_speed=platen_speed, steps=steps, material=material,
                   cell_w=mesh_w, cell_h=mesh_h)
    return sc


def config4_worlds(n=256, seed=20240607, h=0.0002, thickness=0.001, platen_speed=0.1, steps=2000,
                   width_range=(0.004, 0.0105)):
    """``n`` independent single cells drawn from ``seed``
    with SplitMix64: re-entrant angle uniform in [55, 80] degrees, wall
    thickness uniform in [0.5, 1.2] mm, polygon or swiss-cheese topology,
    elastic (E = 1 GPa) or J2 (aluminium) material, cell width uniform in
    ``width_range`` rounded to a multiple of 2h (sizes 5k-80k points at
    h = 0.2 mm)."""
    rnd = splitmix64(seed)
    worlds = []
    for _ in range(n):
        angle = 55.0 + 25.0 * next(rnd)
        wall = 0.0005 + 0.0007 * next(rnd)
        poly = next(rnd) < 0.5
        elastic = next(rnd) < 0.5
        width = width_range[0] + (width_range[1] - width_range[0]) * next(rnd)
        # on the grid (a multiple of 2h) as in the reference worlds: the point
        # lattice then sits h/4 off the node planes instead of arbitrarily
        # close to them (where the linear-hat update loses a node's mass)
        width = 2 * h * max(2, round(width / (2 * h)))
        mat = dict(material="elastic", E=1.0e9, nu=0.3, rho=1200.0) if elastic else \
            dict(material="j2", E=70.0e9, nu=0.33, rho=2700.0, sy=250.0e6, H=700.0e6)
        if poly:
            sc = polygon_mesh_world(1, 1, width, width, wall, thickness, h, platen_speed=platen_speed, steps=steps,
                                    dent_frac=dent_for_angle(width, width, angle) / width, **mat)
        else:
            sc = swisscheese_cell_world(width, wall, thickness, h, platen_speed=platen_speed, steps=steps, **mat)
        sc.info.update(angle=angle, wall=wall, topology="polygon" if poly else "swisscheese")
        worlds.append(sc)
    return worlds


def config3_world(h=2.0e-3 / 21, n_holes=8, pitch=2.0e-3, radius_frac=0.4, thickness=2.0e-3, mu=0.5e6,
                  kappa=5.0e6, rho=1000.0, platen_speed=0.25, steps=10000, cfl=0.3, margin_cells=12,
                  mass_min_frac=0.3):
    """a swiss-cheese block, ``n_holes`` x ``n_holes``
    circular holes of radius ``radius_frac`` x ``pitch`` through the
    ``thickness``, Neo-Hookean (shear modulus ``mu``, bulk modulus ``kappa``:
    E = 9 kappa mu / (3 kappa + mu), nu = (3 kappa - 2 mu) / (2 (3 kappa + mu))),
    points at 2 x 2 x 2 per cell (about 2.35M at the default h = 2 mm / 21),
    on the fixed bottom node plane, compressed by a separate point platen
    moving down at ``platen_speed`` (stored velocity, as the contact step moves
    platens).  Points within one point spacing of a hole or of the outer
    boundary, and the platen's lower layer, are flagged as surface for the
    contact step.  Run it on the contact path (friction 0.2 in the benchmark).

    The node mass gate is ``mass_min_frac`` x the point mass.  The reference
    worlds use 0.001, and with it the reference update diverges on this block
    within ~170 steps at 0.5 m/s (~1270 at 0.05 m/s; the FP64 oracle diverges
    in the same 10-step window): nodes that hold a sliver of a point's mass
    at the hole surfaces take huge accelerations.  0.1 keeps the plain path
    stable, and the contact path, whose two fields split node masses further,
    needs 0.3.  Even so the outer z layers at the free top corners under the
    platen edge pick up growing z velocities: at 0.5 m/s a point leaves the
    grid after ~8000 steps; at the default 0.25 m/s the 10000 steps complete
    (about 6% strain).  ``margin_cells`` leaves room for the lateral bulge."""
    off = 0.5 * h
    W = n_holes * pitch
    extent = (W + 2 * margin_cells * h, W + 4 * h, thickness)
    L, cells = _grid_for(extent, (h, h, h))
    L = (L[0], L[1], thickness)
    cells = (cells[0], cells[1], max(1, int(round(thickness / h))))
    x0 = off * round(0.5 * (L[0] - W) / off)   # on the point lattice of the cells
    xs = x0 + off * (0.5 + np.arange(int(round(W / off))))
    ys = off * (0.5 + np.arange(int(round(W / off))))
    zs = off * (0.5 + np.arange(int(round(thickness / off))))
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    R = radius_frac * pitch
    ix = np.clip(((pts[:, 0] - x0) // pitch).astype(int), 0, n_holes - 1)
    iy = np.clip((pts[:, 1] // pitch).astype(int), 0, n_holes - 1)
    d = np.hypot(pts[:, 0] - (x0 + pitch * (ix + 0.5)), pts[:, 1] - pitch * (iy + 0.5))
    keep = d > R
    body, d = pts[keep], d[keep]
    edge = np.minimum.reduce([body[:, 0] - x0, x0 + W - body[:, 0], body[:, 1], W - body[:, 1]])
    surf_body = ((d - R < off) | (edge < off)).astype(np.uint8)
    E = 9.0 * kappa * mu / (3.0 * kappa + mu)
    nu = (3.0 * kappa - 2.0 * mu) / (2.0 * (3.0 * kappa + mu))
    vol = off ** 3
    mat = _material_arrays(len(body), E, nu, 0.0, 0.0, 0.0, NEOHOOKEAN, rho, vol)
    top = create_domain_cuboid((0.5 * L[0], W + 0.5 * h, 0.5 * thickness), (W, 2.0 * off, thickness), off)
    plat = _material_arrays(len(top), 210.0e9, 0.33, 200.0e6, 0.0, 0.0, ELASTIC, 7800.0, vol)
    dt = cfl * h / p_wave_speed(E, nu, rho)
    sc = _assemble(L, cells, reference_bcs(L, cells), [(body, mat, False), (top, plat, True)],
                   dt_max=dt, time_end=steps * dt, damping=0.0, mass_min=mass_min_frac * rho * vol,
                   timeline_t=np.array([0.0, 1.0e3]),
                   timeline_v=np.array([[0.0, -platen_speed, 0.0], [0.0, -platen_speed, 0.0]]), neohookean=1)
    sc.v[sc.dc == 1] = (0.0, -platen_speed, 0.0)
    sc.surface = np.concatenate([surf_body, (top[:, 1] < W + 0.5 * h).astype(np.uint8)])
    sc.info = dict(name="config3", h=h, dt=dt, E=E, nu=nu, holes=n_holes, pitch=pitch, radius=R,
                   platen_speed=platen_speed, steps=steps)
    return sc


def log_uniform_moduli(n, seed, lo=0.5e9, hi=2.0e9):
    """``n`` Young's moduli log-uniform in [lo, hi] from SplitMix64(seed)."""
    rnd = splitmix64(seed)
    u = np.fromiter((next(rnd) for _ in range(n)), dtype=np.float64, count=n)
    return np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))


def config5_world(seed=20240611, nx=4, ny=4, h=0.00025, thickness=0.00075, steps=400, platen_speed=1.0,
                  ramp_steps=100, E_range=(0.5e9, 2.0e9)):
    """an ``nx`` x ``ny`` re-entrant polygon mesh (about
    160k points at h = 0.25 mm, 0.75 mm thick), elastic, with a per-point
    Young's modulus log-uniform in ``E_range`` drawn from ``seed``, the time
    step set by the stiffest modulus, and a platen ramped to ``platen_speed``
    over ``ramp_steps`` steps."""
    sc = polygon_mesh_world(nx, ny, h=h, thickness=thickness, material="elastic", E=E_range[1],
                            platen_speed=platen_speed, steps=steps, ramp_steps=ramp_steps)
    keep = sc.dc == 0
    sc.E = sc.E.copy()
    sc.E[keep] = log_uniform_moduli(int(keep.sum()), seed, *E_range)
    sc.info.update(name="config5", seed=seed, E_range=E_range)
    return sc


def batch_worlds(worlds, gap_cells=2, row_cells=None):
    """Pack independent worlds into one grid for a single launch per step.

    All worlds share the cell size, time step rule and TimeLine; each keeps its
    own node BCs, and worlds sit ``gap_cells`` (>= 2) empty cells apart so no
    node is touched by two of them.  Shelf packing in x-y rows of about
    ``row_cells`` cells.  Returns the combined Scene with ``node_group`` (world
    id per node), ``info["world"]`` (world id per point) and the offsets."""
    if gap_cells < 2:
        raise ValueError("worlds need >= 2 empty cells between them")
    h = [worlds[0].length[a] / worlds[0].cells[a] for a in range(3)]
    for w in worlds:
        hw = [w.length[a] / w.cells[a] for a in range(3)]
        if not np.allclose(hw, h, rtol=1e-12, atol=0.0):
            raise ValueError("worlds must share the cell size")
        if not (np.array_equal(w.timeline_t, worlds[0].timeline_t) and np.array_equal(w.timeline_v, worlds[0].timeline_v)):
            raise ValueError("worlds must share the TimeLine")
        if w.neohookean != worlds[0].neohookean or w.damping != worlds[0].damping:
            raise ValueError("worlds must share damping and the NeoHookean dispatch flag")
    if row_cells is None:
        row_cells = int(math.ceil(math.sqrt(sum((w.cells[0] + gap_cells) * (w.cells[1] + gap_cells) for w in worlds))))
    offs, ox, oy, row_h, width = [], 0, 0, 0, 0
    for w in worlds:
        if ox > 0 and ox + w.cells[0] > row_cells:
            ox, oy, row_h = 0, oy + row_h + gap_cells, 0
        offs.append((ox, oy))
        ox += w.cells[0] + gap_cells
        row_h = max(row_h, w.cells[1])
        width = max(width, ox - gap_cells)
    cells = (width, oy + row_h, max(w.cells[2] for w in worlds))
    L = tuple(cells[a] * h[a] for a in range(3))
    nn = [c + 1 for c in cells]
    fixed = np.zeros(nn[0] * nn[1] * nn[2], np.uint8)
    group = np.zeros(len(fixed), np.uint16)
    cols = {k: [] for k in ("x", "mass", "vol0", "E", "nu", "sy", "C0", "C1", "type", "dc", "v", "fext", "surface")}
    world_of = []
    for i, (w, (ox, oy)) in enumerate(zip(worlds, offs)):
        wn = [c + 1 for c in w.cells]
        iz, iy, ix = np.meshgrid(np.arange(wn[2]), np.arange(wn[1]), np.arange(wn[0]), indexing="ij")
        dst = (ox + ix) + nn[0] * ((oy + iy) + nn[1] * iz)
        fixed[dst.ravel()] = w.fixed
        group[dst.ravel()] = i
        shift = np.array([ox * h[0], oy * h[1], 0.0])
        for k in cols:
            a = getattr(w, k)
            cols[k].append(a + shift if k == "x" else a)
        world_of.append(np.full(w.n, i, np.int32))
    w0 = worlds[0]
    sc = Scene(length=L, cells=cells, fixed=fixed, **{k: np.concatenate(v) for k, v in cols.items()},
               dt_max=min(w.dt_max for w in worlds), time_end=max(w.time_end for w in worlds), damping=w0.damping,
               mass_min=min(w.mass_min for w in worlds), timeline_t=w0.timeline_t, timeline_v=w0.timeline_v,
               neohookean=w0.neohookean, node_group=group, n_groups=len(worlds))
    sc.info = dict(name="batch", world=np.concatenate(world_of), offsets=offs, n_worlds=len(worlds))
    return sc


def effective_poisson(x0, x1, keep):
    """Effective Poisson's ratio of one cell from point positions before and
    after loading: -(lateral strain) / (axial strain), with the lateral
    extent measured on the points of the middle half of the height and the
    axial extent on all points, both as 2%-98% quantile spans."""
    y0 = x0[keep, 1]
    lo, hi = np.quantile(y0, [0.25, 0.75])
    mid = keep.copy()
    mid[keep] = (y0 >= lo) & (y0 <= hi)

    def span(x, sel, a):
        q = np.quantile(x[sel, a], [0.02, 0.98])
        return q[1] - q[0]
    eps_ax = span(x1, keep, 1) / span(x0, keep, 1) - 1.0
    eps_lat = span(x1, mid, 0) / span(x0, mid, 0) - 1.0
    return -eps_lat / eps_ax


# ------------------------------------------------------------ config-driven
def worlds():
    """Name -> builder of every world this module makes (the
    initializeWorld_* of the reference and the benchmark configurations)."""
    return {
        "reference_polygon": reference_polygon_world,
        "reference_swisscheese": reference_swisscheese_world,
        "polygon_mesh": polygon_mesh_world,
        "swisscheese_cell": swisscheese_cell_world,
        "small": small_world,
        "translation": translation_world,
        "config1": config1_world,
        "config2": config2_world,
        "config3": config3_world,
        "config4": lambda **kw: batch_worlds(config4_worlds(**kw)),
        "config5": config5_world,
    }


_SCENE_FIELDS = ("dt_max", "time_end", "damping", "mass_min", "console_interval", "neohookean")


def from_config(spec):
    """Build a scene from a spec: a world name, a path to a JSON file, or a
    dict ``{"world": name, <builder keyword arguments>..., "set": {field:
    value}}`` where ``set`` overrides solver fields (dt_max, time_end,
    damping, mass_min, console_interval, neohookean) after the build."""
    import json
    if isinstance(spec, str):
        if spec.endswith(".json"):
            with open(spec) as f:
                spec = json.load(f)
        else:
            spec = {"world": spec}
    spec = dict(spec)
    name = spec.pop("world")
    over = spec.pop("set", {})
    table = worlds()
    if name not in table:
        raise ValueError(f"unknown world {name!r}; known: {sorted(table)}")
    sc = table[name](**spec)
    for k, v in over.items():
        if k not in _SCENE_FIELDS:
            raise ValueError(f"cannot set {k!r}; settable: {_SCENE_FIELDS}")
        setattr(sc, k, type(getattr(sc, k))(v))
    sc.info.setdefault("spec", dict(world=name, **spec, set=over))
    return sc

