"""Step kernels of the 2D driver with the reference's names and argument order
(2D/run_2D.py:101-611).  Fields are torch CUDA tensors (see ``pfm.core``).

The reference reads ``dx``, ``res_*`` and ``C_x``/``C_y`` as module globals;
here the grid comes from the face fields (dx = 1/res_y) unless ``dx=`` is
given, and P2G takes ``C=(C_x, C_y)`` (the rows from
``update_particles_grad_m``) or the ones it last wrote.
"""
from __future__ import annotations

from .. import core
from .. import _lib as L

max_speed = None
_last_C = None


def _res2(u_x):
    return core.res_from_faces(u_x)


def _dx(res, dx):
    return core.default_dx(res) if dx is None else dx


def calc_max_speed(u_x, u_y):
    """2D/run_2D.py:101-108."""
    global max_speed
    max_speed = core.max_speed(_res2(u_x), (u_x, u_y), out=max_speed)
    return max_speed


def reset_to_identity_grid(psi_x, psi_y, T_x, T_y, *, dx=None):
    """2D/run_2D.py:120-129."""
    res = (psi_y.shape[1], psi_x.shape[2])
    d = _dx(res, dx)
    core.reset_identity_grid(res, d, psi_x, T_x)
    core.reset_identity_grid(res, d, psi_y, T_y)


def reset_to_identity(psi, T_x, T_y, *, particles_pos=None):
    """2D/run_2D.py:132-139 (psi = particles_pos, T = I)."""
    if particles_pos is not None:
        psi.copy_(particles_pos)
    core.reset_T_identity((T_x, T_y))


def reset_T_to_identity(T_x, T_y):
    """2D/run_2D.py:141-146."""
    core.reset_T_identity((T_x, T_y))


def RK4_grid(psi_x, T_x, u_x0, u_y0, dt, *, dx=None):
    """2D/run_2D.py:233-261."""
    res = _res2(u_x0)
    core.rk4_grid(res, _dx(res, dx), psi_x, T_x, (u_x0, u_y0), dt)


def RK4_T_forward(psi, T_x, T_y, u_x0, u_y0, dt, advect_psi, *, dx=None):
    """2D/run_2D.py:304-341."""
    res = _res2(u_x0)
    core.rk4_particles(res, _dx(res, dx), psi, (T_x, T_y), (u_x0, u_y0), dt, bool(advect_psi))


def stretch_T_and_advect_particles(particles_pos, T_x, T_y, u_x, u_y, dt):
    """2D/run_2D.py:220-222."""
    RK4_T_forward(particles_pos, T_x, T_y, u_x, u_y, dt, 1)


def stretch_T(particles_pos, T_x, T_y, u_x, u_y, dt):
    """2D/run_2D.py:224-225."""
    RK4_T_forward(particles_pos, T_x, T_y, u_x, u_y, dt, 0)


def advect_u(u_x0, u_y0, u_x1, u_y1, T_x, T_y, psi_x, psi_y, dx):
    """2D/run_2D.py:429-441."""
    res = _res2(u_x0)
    core.advect_u(res, dx, (u_x0, u_y0), (u_x1, u_y1), (T_x, T_y), (psi_x, psi_y))


def advect_smoke(smoke0, smoke1, psi_x, psi_y, dx):
    """2D/run_2D.py:444-450 (smoke as (3, rx, ry), psi as (2, face shape))."""
    core.advect_smoke(tuple(smoke0.shape[1:]), dx, smoke0, smoke1, (psi_x, psi_y), mode=1)


def update_particles_imp(particles_imp, particles_init_imp, T_x, T_y):
    """2D/run_2D.py:517-524."""
    core.update_particles_imp((T_x, T_y), particles_init_imp, particles_imp)


def update_particles_grad_m(C_x, C_y, init_C_x, init_C_y, T_x, T_y):
    """2D/run_2D.py:526-536."""
    global _last_C
    core.update_particles_grad_m((C_x, C_y), (init_C_x, init_C_y), (T_x, T_y))
    _last_C = (C_x, C_y)


def update_T(T_x, T_y, T_x_init, T_y_init, T_x_grad_m, T_y_grad_m):
    """2D/run_2D.py:538-546: T <- cols(T_grad_m) @ cols(T_init)."""
    core.update_T((T_x_grad_m, T_y_grad_m), (T_x_init, T_y_init), (T_x, T_y))


def get_particles_id_in_every_cell(cell_particles_id, cell_particle_num, particles_pos, *, dx=None):
    """Binning as in 3D/run_3D.py:342-351 (the 2D driver has no caller)."""
    res = tuple(cell_particle_num.shape)
    b = core.binning_for(res, particles_pos.shape[1], particles_pos.device)
    d = _dx(res, dx)
    b.bin(particles_pos, d)
    b.cell_lists(d, cell_particles_id, cell_particle_num)
    return b


def P2G(particles_imp, particles_pos, u_x, u_y, psi=None, psi_x_grid=None, psi_y_grid=None,
        p2g_weight_x=None, p2g_weight_y=None, *, C=None, dx=None):
    """2D/run_2D.py:548-611 as a deterministic gather.  The reference's psi
    splat into psi_x_grid/psi_y_grid is dead work (overwritten before any
    read) and is not performed."""
    res = _res2(u_x)
    d = _dx(res, dx)
    Cr = C if C is not None else _last_C
    if Cr is None:
        raise L.PfmError("P2G needs C (call update_particles_grad_m first or pass C=)")
    b = core.binning_for(res, particles_pos.shape[1], particles_pos.device)
    W = (p2g_weight_x, p2g_weight_y) if p2g_weight_x is not None else None
    b.p2g(d, particles_pos, particles_imp, Cr, (u_x, u_y), W=W)
