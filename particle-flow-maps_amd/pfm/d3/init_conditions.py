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
