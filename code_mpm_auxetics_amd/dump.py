"""Text dumps of material points and grid nodes in the reference's column
format.

MaterialPoint_Factory::getScript (MaterialPoint_Factory_getScript.cpp:4-28)
writes ``ID mass position_x position_y position_z``; GridPoint_Factory::
getScript (GridPoint_Factory.h:69-100) writes ``ID mass fixed_x fixed_y fixed_z
position_x position_y position_z``, its ID column always 0.  Numbers use the
Script() formats of Definitions.h:26-58: integers in decimal, doubles in
scientific notation with a sign and 8 digits after the point (``%+.8e``).
Called with no point, each returns its header line.  Every row ends with a
tab and a newline, as the reference's do.

MaterialPoint_Factory::saveLatex is declared (MaterialPoint_Factory.h:41)
and called from the contact step (…_MP_Contact.cpp:656-661) but never
defined.  ``save_latex`` defines it as the getScript table written to
``prefix + postfix``: whitespace-separated columns under one header row, the
table form pgfplots' addplot-table command reads.

The step keeps no grid between steps (nodes live in on-chip memory inside
one launch), so the node mass a grid dump shows is the P2G mass of the
points' current positions (…_MP.cpp:106-145 with the hats of
GridPoint_Mediator.h:31-88): the mass the next step scatters.  The
reference's GridPoint holds the mass of the step before, scattered from the
positions at that step's start.
"""
import numpy as np

from . import scene as _scene

MP_HEADER = "ID\tmass\t\t\tposition_x\t\tposition_y\t\tposition_z\t\t\n"
GP_HEADER = "ID\tmass\t\t\tfixed_x\tfixed_y\tfixed_z\tposition_x\t\tposition_y\t\tposition_z\t\n"


def _e8(v):
    return "%+.8e" % float(v)


def material_point_script(point_id=None, mass=0.0, position=(0.0, 0.0, 0.0)):
    """One row of MaterialPoint_Factory::getScript, or its header when
    point_id is None."""
    if point_id is None:
        return MP_HEADER
    return "%d\t%s\t%s\t%s\t%s\t\n" % (int(point_id), _e8(mass), *(_e8(c) for c in position))


def grid_point_script(mass=None, fixed=(False, False, False), position=(0.0, 0.0, 0.0)):
    """One row of GridPoint_Factory::getScript, or its header when mass is
    None."""
    if mass is None:
        return GP_HEADER
    return "0\t%s\t%d\t%d\t%d\t%s\t%s\t%s\t\n" % (_e8(mass), *(int(bool(f)) for f in fixed),
                                                   *(_e8(c) for c in position))


def material_points_script(x, mass, ids=None):
    """Header plus one row per point.  ``ids`` defaults to 1 for every point,
    the ID the reference worlds give every point
    (PhysicsEngine_initializeWorld_AuxeticPolygonCell.cpp:115,161,207)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    mass = np.asarray(mass, dtype=np.float64).reshape(-1)
    ids = np.ones(len(x), dtype=np.int64) if ids is None else np.asarray(ids).reshape(-1)
    return MP_HEADER + "".join(material_point_script(i, m, p) for i, m, p in zip(ids, mass, x))


def node_masses(length, cells, x, mass, dc):
    """P2G node mass of points at ``x`` (…_MP.cpp:106-145): hats of
    GridPoint_Mediator.h:31-88 (base node by truncation, 8 nodes, nodes off the
    grid skipped, weights outside [0, 1] zeroed), displacement-controlled
    points left out.  Returns (nx*ny*nz,) with x the fastest index."""
    nn = np.array([c + 1 for c in cells])
    h = np.array([length[a] / cells[a] for a in range(3)])
    keep = np.asarray(dc).reshape(-1) == 0
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)[keep]
    m = np.asarray(mass, dtype=np.float64).reshape(-1)[keep]
    base = np.trunc(x / h).astype(np.int64)
    out = np.zeros(int(np.prod(nn)))
    for ox in (0, 1):
        for oy in (0, 1):
            for oz in (0, 1):
                node = base + np.array([ox, oy, oz])
                ok = np.all((node >= 0) & (node < nn), axis=1)
                s = 1.0 - np.abs(x - node * h) / h
                s = np.where((s < 0.0) | (s > 1.0), 0.0, s)
                w = s[:, 0] * s[:, 1] * s[:, 2]
                idx = node[:, 0] + node[:, 1] * nn[0] + node[:, 2] * nn[0] * nn[1]
                out += np.bincount(idx[ok], weights=(w * m)[ok], minlength=len(out))
    return out


def grid_points_script(sc, x=None, active_only=False, mass=None):
    """Header plus one row per node of the scene's grid (x fastest).  Node
    masses are ``mass`` when given (a solver's grid readback), else the P2G
    mass of points at ``x`` (default: the scene's initial positions) -- the
    host-side factory view before any solver exists; ``active_only`` keeps
    the nodes with mass only."""
    pos = _scene.node_positions(sc.length, sc.cells)
    m = np.asarray(mass, np.float64) if mass is not None else node_masses(sc.length, sc.cells,
                                                                           sc.x if x is None else x, sc.mass, sc.dc)
    fx = np.asarray(sc.fixed).reshape(-1)
    rows = [GP_HEADER]
    for i in (np.flatnonzero(m > 0.0) if active_only else range(len(pos))):
        f = int(fx[i])
        rows.append(grid_point_script(m[i], (f & _scene.FIX_X, f & _scene.FIX_Y, f & _scene.FIX_Z), pos[i]))
    return "".join(rows)


def save_latex(x, mass, prefix, postfix, ids=None):
    """MaterialPoint_Factory::saveLatex(points, strPrefix, strPostfix): the
    getScript table of the points, written to ``prefix + postfix``; returns
    the path."""
    path = str(prefix) + str(postfix)
    with open(path, "w") as f:
        f.write(material_points_script(x, mass, ids))
    return path
