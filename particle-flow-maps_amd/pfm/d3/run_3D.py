"""Step kernels of the 3D driver with the reference's names and argument order
(3D/run_3D.py:95-552).  Fields are torch CUDA tensors (see ``pfm.core``);
every call runs on the current torch stream.

The reference reads ``dx``, ``res_*`` and ``init_C_*`` as module globals; here
the grid comes from the face fields' shapes (dx = 1/res_y) unless ``dx=`` is
given, and P2G takes ``init_C=(C_x, C_y, C_z)`` or, like the reference, the
rows most recently written by ``init_particles_imp_grad_m``.
"""
from __future__ import annotations

import torch

from .. import core
from .. import _lib as L
from . import init_conditions as _ic

# calc_max_speed output (the reference's `max_speed` field, 3D/run_3D.py:89)
max_speed = None


def _res3(u_x):
    return core.res_from_faces(u_x)


def _dx(res, dx):
    return core.default_dx(res) if dx is None else dx


def calc_max_speed(u_x, u_y, u_z):
    """3D/run_3D.py:95-103; result in ``max_speed`` (device tensor [1])."""
    global max_speed
    res = _res3(u_x)
    max_speed = core.max_speed(res, (u_x, u_y, u_z), out=max_speed)
    return max_speed


def reset_to_identity_grid(psi_x, psi_y, psi_z, T_x, T_y, T_z, *, res=None, dx=None):
    """3D/run_3D.py:106-120 (psi = face coordinates, T = identity columns)."""
    res = res or (psi_y.shape[1], psi_x.shape[2], psi_x.shape[3])
    dx = _dx(res, dx)
    for psi, T in ((psi_x, T_x), (psi_y, T_y), (psi_z, T_z)):
        core.reset_identity_grid(res, dx, psi, T)


def reset_to_identity(psi, T_x, T_y, T_z, *, particles_pos=None):
    """3D/run_3D.py:122-131 (psi = particles_pos, T = I)."""
    if particles_pos is not None:
        psi.copy_(particles_pos)
    core.reset_T_identity((T_x, T_y, T_z))


def reset_T_to_identity(T_x, T_y, T_z):
    """3D/run_3D.py:133-140."""
    core.reset_T_identity((T_x, T_y, T_z))


def RK4_grid(psi_x, T_x, u_x0, u_y0, u_z0, dt, *, dx=None):
    """3D/run_3D.py:144-173 (backward RK4 of one face axis, in place)."""
    res = _res3(u_x0)
    core.rk4_grid(res, _dx(res, dx), psi_x, T_x, (u_x0, u_y0, u_z0), dt)


def RK4_T_forward(psi, T_x, T_y, T_z, u_x0, u_y0, u_z0, dt, *, dx=None):
    """3D/run_3D.py:175-219 (forward RK4 of particles and T, in place)."""
    res = _res3(u_x0)
    core.rk4_particles(res, _dx(res, dx), psi, (T_x, T_y, T_z), (u_x0, u_y0, u_z0), dt, True)


def stretch_T_and_advect_particles(particles_pos, T_x, T_y, T_z, u_x, u_y, u_z, dt):
    """3D/run_3D.py:336-337."""
    RK4_T_forward(particles_pos, T_x, T_y, T_z, u_x, u_y, u_z, dt)


def advect_u(u_x0, u_y0, u_z0, u_x1, u_y1, u_z1, T_x, T_y, T_z, psi_x, psi_y, psi_z, dx):
    """3D/run_3D.py:296-312."""
    res = _res3(u_x0)
    core.advect_u(res, dx, (u_x0, u_y0, u_z0), (u_x1, u_y1, u_z1), (T_x, T_y, T_z), (psi_x, psi_y, psi_z))


def advect_smoke(smoke0, smoke1, psi_x, psi_y, psi_z, dx):
    """3D/run_3D.py:315-323 (smoke as (4, rx, ry, rz), psi as (3, face shape))."""
    core.advect_smoke(tuple(smoke0.shape[1:]), dx, smoke0, smoke1, (psi_x, psi_y, psi_z), mode=1)


def clamp_smoke(smoke0, smoke1, psi_x, psi_y, psi_z, dx):
    """3D/run_3D.py:325-334."""
    core.advect_smoke(tuple(smoke0.shape[1:]), dx, smoke0, smoke1, (psi_x, psi_y, psi_z), mode=2)


def P2G_smoke(particles_smoke, particles_init_grad_smoke, particles_pos, smoke, T_x_grad_m, T_y_grad_m,
              T_z_grad_m, T_x_init, T_y_init, T_z_init, p2g_weight, *, use_APIC_smoke=True, dx=None):
    """3D/run_3D.py:554-593 (smoke (4, *res), p2g_weight res, particle smoke
    (4, n), gradient (12, n)); deterministic gather over the cell sort."""
    res = tuple(p2g_weight.shape)
    d = _dx(res, dx)
    b = core.binning_for(res, particles_pos.shape[1], particles_pos.device)
    b.bin(particles_pos, d)
    core.p2g_smoke(res, d, particles_pos, particles_smoke, particles_init_grad_smoke,
                   (T_x_grad_m, T_y_grad_m, T_z_grad_m), (T_x_init, T_y_init, T_z_init), b, smoke, p2g_weight,
                   use_apic=use_APIC_smoke)


def clamp_smoke_particle(cell_particles_id, cell_particle_num, particles_init_pos, smoke, init_smoke, *, dx=None):
    """3D/run_3D.py:595-614 over the lists of the last
    get_particles_id_in_every_cell (its cell sort, at most cell_max per cell)."""
    res = tuple(cell_particle_num.shape)
    b = core.binning_for(res, particles_init_pos.shape[1], particles_init_pos.device)
    if b._key is None:
        raise L.PfmError("clamp_smoke_particle needs get_particles_id_in_every_cell first")
    core.clamp_smoke_particle(res, _dx(res, dx), b, cell_particles_id.shape[-1], particles_init_pos, init_smoke,
                              smoke)


def get_particles_id_in_every_cell(cell_particles_id, cell_particle_num, particles_pos, *, dx=None):
    """3D/run_3D.py:342-351, deterministic: ids sorted ascending inside a cell."""
    res = tuple(cell_particle_num.shape)
    b = core.binning_for(res, particles_pos.shape[1], particles_pos.device)
    d = _dx(res, dx)
    b.bin(particles_pos, d)
    b.cell_lists(d, cell_particles_id, cell_particle_num)
    return b


def update_T(T_x_init, T_y_init, T_z_init, T_x_grad_m, T_y_grad_m, T_z_grad_m):
    """3D/run_3D.py:430-439: T_init <- cols(T_grad_m) @ cols(T_init)."""
    Tin = (T_x_init, T_y_init, T_z_init)
    core.update_T((T_x_grad_m, T_y_grad_m, T_z_grad_m), Tin, Tin)


def P2G(particles_init_imp, particles_init_imp_grad_m, particles_pos, u_x, u_y, u_z, p2g_weight_x,
        p2g_weight_y, p2g_weight_z, T_x_grad_m, T_y_grad_m, T_z_grad_m, T_x_init, T_y_init, T_z_init,
        cell_particles_id=None, cell_particle_num=None, *, init_C=None, dx=None):
    """3D/run_3D.py:441-552 as a deterministic gather (cell lists are unused,
    as in the reference)."""
    res = _res3(u_x)
    d = _dx(res, dx)
    C = init_C if init_C is not None else _ic.last_init_C()
    if C is None:
        raise L.PfmError("P2G needs init_C (call init_particles_imp_grad_m first or pass init_C=)")
    b = core.binning_for(res, particles_pos.shape[1], particles_pos.device)
    W = (p2g_weight_x, p2g_weight_y, p2g_weight_z) if p2g_weight_x is not None else None
    b.p2g(d, particles_pos, particles_init_imp, C, (u_x, u_y, u_z), W=W,
          Tgm=(T_x_grad_m, T_y_grad_m, T_z_grad_m), Tin=(T_x_init, T_y_init, T_z_init))
