"""Initial conditions and reinitialisation G2P of the 2D driver
(2D/init_conditions.py:5-60,139-210) with the reference's names."""
from __future__ import annotations

import torch

from .. import core


def init_particles_pos_uniform(particles_pos, X, res_x, particles_per_cell, dx, particles_per_cell_axis,
                               dist_between_neighbor, *, res_y=None, seed=-1, order=0):
    """2D/init_conditions.py:139-166 (+ seeded jitter / cell order extensions)."""
    if res_y is None:
        n = particles_pos.shape[1]
        res_y = n // (particles_per_cell_axis ** 2 * res_x)
    core.seed_particles((res_x, res_y), particles_per_cell_axis, dist_between_neighbor, particles_pos,
                        seed=seed, order=order)


def init_particles_imp(particles_imp, particles_init_imp, particles_pos, u_x, u_y, C_x, C_y, dx):
    """2D/init_conditions.py:183-191: imp = init_imp = u(pos)."""
    res = core.res_from_faces(u_x)
    core.g2p(res, dx, particles_pos, (u_x, u_y), particles_imp, None)
    particles_init_imp.copy_(particles_imp)


def init_particles_imp_grad_m(particles_imp, particles_pos, u_x, u_y, C_x, C_y, dx):
    """2D/init_conditions.py:203-210."""
    res = core.res_from_faces(u_x)
    core.g2p(res, dx, particles_pos, (u_x, u_y), particles_imp, (C_x, C_y))


def point_vortex_vel_func(vf, centers, weights, rad=0.02, strength=-0.01, *, dx=None):
    """leapfrog_vel_func (2D/init_conditions.py:29-60) for any list of vortex
    centres and weights: vf (2, rx, ry) cell-centred velocity."""
    res = tuple(vf.shape[1:])
    core.point_vortices(res, core.default_dx(res) if dx is None else dx, centers, weights, vf, rad, strength)


def leapfrog_vel_func(vf, pf=None):
    """2D/init_conditions.py:29-60 with the reference's four vortices."""
    point_vortex_vel_func(vf, [(0.25, 0.62), (0.25, 0.38), (0.25, 0.74), (0.25, 0.26)], [-0.5, 0.5, -0.5, 0.5])
