"""Initial conditions and reinitialisation G2P of the 3D driver
(3D/init_conditions.py:10-22,238-273) with the reference's names."""
from __future__ import annotations

import math

import torch

from .. import core

_last_C = None


def last_init_C():
    """Rows most recently written by init_particles_imp_grad_m (the reference
    keeps them in the module-global init_C_x/y/z fields)."""
    return _last_C


def init_particles_pos_uniform(particles_pos, X, res_x, res_y, particles_per_cell, dx,
                               particles_per_cell_axis, dist_between_neighbor, *, res_z=None, seed=-1,
                               order=0):
    """3D/init_conditions.py:238-252.  ``seed >= 0`` adds seeded jitter and
    ``order=1`` numbers particles cell by cell (extensions; defaults give the
    reference lattice)."""
    if res_z is None:
        n = particles_pos.shape[1]
        per = particles_per_cell_axis ** 3
        res_z = n // (per * res_x * res_y)
    core.seed_particles((res_x, res_y, res_z), particles_per_cell_axis, dist_between_neighbor, particles_pos,
                        seed=seed, order=order)


def init_particles_imp(particles_init_imp, particles_pos, u_x, u_y, u_z, C_x, C_y, C_z, dx):
    """3D/init_conditions.py:254-263 (C is not written, as in the reference)."""
    res = core.res_from_faces(u_x)
    core.g2p(res, dx, particles_pos, (u_x, u_y, u_z), particles_init_imp, None)


def init_particles_imp_grad_m(particles_imp, particles_pos, u_x, u_y, u_z, C_x, C_y, C_z, dx):
    """3D/init_conditions.py:265-273."""
    global _last_C
    res = core.res_from_faces(u_x)
    core.g2p(res, dx, particles_pos, (u_x, u_y, u_z), particles_imp, (C_x, C_y, C_z))
    _last_C = (C_x, C_y, C_z)


def init_particles_smoke(particles_smoke, particles_grad_smoke, particles_pos, smoke, dx):
    """3D/init_conditions.py:275-279 (smoke (4, *res); particle smoke (4, n),
    gradient (12, n) row-major 3x4)."""
    core.g2p_smoke(tuple(smoke.shape[1:]), dx, particles_pos, smoke, particles_smoke, particles_grad_smoke)


def add_vortex_ring(w, rad, delta, c, unit_x, unit_y, pf, vf, num_samples, *, dx=None):
    """3D/init_conditions.py:10-22: adds one smoothed Biot-Savart ring into the
    cell-centred velocity ``vf`` (3, rx, ry, rz).  ``pf`` (cell centres) is
    implied by the grid and only used for its shape."""
    res = tuple(vf.shape[1:])
    core.vortex_ring(res, core.default_dx(res) if dx is None else dx, vf, w, rad, delta, c, unit_x, unit_y,
                     num_samples)


def add_vortex_ring_and_smoke(w, rad, delta, c, unit_x, unit_y, pf, vf, smokef, color, num_samples, *,
                              dx=None):
    """3D/init_conditions.py:31-56: the ring's velocity into ``vf`` (3, *res)
    as add_vortex_ring, its smoke into ``smokef`` (4, *res): density into
    channel 3, then (color, 1) where it exceeds 0.002 and 0 elsewhere."""
    res = tuple(vf.shape[1:])
    d = core.default_dx(res) if dx is None else dx
    core.vortex_ring(res, d, vf, w, rad, delta, c, unit_x, unit_y, num_samples)
    core.vortex_ring_smoke(res, d, smokef, rad, delta, c, unit_x, unit_y, color, num_samples)


def coaxial_rings(res, rings, num_samples=2000, device="cuda"):
    """Cell-centred velocity of several rings (list of dicts with w, rad,
    delta, c, unit_x, unit_y) summed as the reference's init_vorts_* do
    (3D/init_conditions.py:81-114)."""
    vc = torch.zeros((3,) + tuple(res), dtype=torch.float32, device=device)
    for r in rings:
        add_vortex_ring(r["w"], r["rad"], r["delta"], r["c"], r["unit_x"], r["unit_y"], None, vc, num_samples)
    return vc


def _unit(v):
    n = math.sqrt(sum(x * x for x in v))
    return [x / n for x in v]


def _ring_and_smoke(u, smoke, w, rad, delta, c, unit_x, unit_y, color, num_samples):
    add_vortex_ring_and_smoke(w=w, rad=rad, delta=delta, c=c, unit_x=unit_x, unit_y=unit_y, pf=None, vf=u,
                              smokef=smoke, color=color, num_samples=num_samples)


def init_vorts_leapfrog(X, u, smoke1, smoke2):
    """3D/init_conditions.py:81-114 (the "front and back revised" pair, case 0):
    two coaxial rings of radius 0.21 travelling along +x, core 0.08 r,
    strength 0.1 r, centred at x = 0.16 and 0.16 + 0.625 r (y = z = 0.5), each
    with its own smoke colour (red, blue); smoke1 ends as the sum of both.
    u (3, *res) is the cell-centred velocity the rings are added to, smoke1/2
    (4, *res); X (the cell centres) is implied by the grid.  The reference's
    smoke2 is a fresh (zero) field."""
    radius = 0.21
    x_gap = 0.625 * radius
    x_start = 0.16
    delta = 0.08 * radius
    w = radius * 0.1
    ex, ey = [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]
    _ring_and_smoke(u, smoke1, w, radius, delta, [x_start, 0.5, 0.5], ex, ey, [1.0, 0.0, 0.0], 2000)
    _ring_and_smoke(u, smoke2, w, radius, delta, [x_start + x_gap, 0.5, 0.5], ex, ey, [0.0, 0.0, 1.0], 2000)
    core.add_fields(smoke1, smoke2, smoke1, 1.0)


def init_vorts_eight(X, u, smoke1, smoke2):
    """3D/init_conditions.py:119-180 (case 1): eight rings of radius 0.08, core
    0.024, strength +-0.02 (500 samples each) at the corners of the cube
    0.5 +- 0.16, each in the plane normal to its corner's diagonal, with one
    smoke colour per ring; smoke1 accumulates every ring's smoke, smoke2 is
    the per-ring scratch (it ends holding the last ring's smoke)."""
    smoke1.zero_()
    smoke2.zero_()
    size = 0.08
    cos45 = math.cos(math.pi / 4)
    tan_theta = math.sqrt(2) / 2
    cot_theta = math.sqrt(2)
    off = (0.16, 0.16, 0.16 / cos45 * tan_theta)
    c45 = cos45
    # (w sign, corner signs (x, y, z), unit_x, unit_y, colour) per ring, in the reference's order
    rings = [
        (+1, (-1, -1, -1), (-c45, c45, 0.0), (-c45, -c45, cot_theta), (1, 0, 0)),
        (-1, (+1, -1, -1), (c45, c45, 0.0), (c45, -c45, cot_theta), (0, 1, 0)),
        (+1, (-1, +1, -1), (c45, c45, 0.0), (-c45, c45, cot_theta), (0, 0, 1)),
        (-1, (+1, +1, -1), (-c45, c45, 0.0), (c45, c45, cot_theta), (0, 0, 0)),
        (+1, (-1, -1, +1), (-c45, c45, 0.0), (c45, c45, cot_theta), (0, 1, 1)),
        (-1, (+1, -1, +1), (c45, c45, 0.0), (-c45, c45, cot_theta), (1, 0, 1)),
        (+1, (-1, +1, +1), (c45, c45, 0.0), (c45, -c45, cot_theta), (1, 1, 0)),
        (-1, (+1, +1, +1), (-c45, c45, 0.0), (-c45, -c45, cot_theta), (1, 1, 1)),
    ]
    for k, (sgn, corner, ex, ey, color) in enumerate(rings):
        c = [0.5 + s * o for s, o in zip(corner, off)]
        target = smoke1 if k == 0 else smoke2
        _ring_and_smoke(u, target, sgn * 2.0e-2, size, 0.024, c, _unit(ex), _unit(ey), [float(x) for x in color],
                        500)
        if k > 0:
            core.add_fields(smoke1, smoke2, smoke1, 1.0)
            if k < len(rings) - 1:  # the reference leaves the last ring's smoke in smoke2
                smoke2.zero_()
