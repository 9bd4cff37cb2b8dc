"""The drop-in surface on the GPU: the reference's step loops written with
the pfm.d3 / pfm.d2 names, argument order and field shapes, as the
reference's main loops write them (3D/run_3D.py:716-822, 2D/run_2D.py:
772-909), stepped next to the oracle's restatement of the same loop.

The loop bodies below are the reference's sequence of calls on torch
tensors; the only edits are the field allocation (torch instead of
ti.field) and the scalar read of the max speed."""
import math

import numpy as np
import pytest
import torch

from helpers import rel_l2, smooth_mac_field

pytestmark = pytest.mark.gpu

DEV = "cuda"


def N(t):
    return t.detach().cpu().numpy()


def Z(*shape, dtype=torch.float32):
    return torch.zeros(shape, dtype=dtype, device=DEV)


def test_reference_3d_loop_with_pfm_names(oracle, pfm):
    from pfm.d3 import init_conditions as IC
    from pfm.d3 import run_3D as R
    from pfm.d3 import taichi_utils as TU
    from pfm.d3.mgpcg import MGPCG_3

    res_x = res_y = res_z = 16
    dx = 1.0 / res_y
    CFL, visualize_dt = 0.5, 0.1
    reinit_every, reinit_every_grad_m = 4, 2
    particles_per_cell, particles_per_cell_axis = 8, 2
    dist_between_neighbor = dx / particles_per_cell_axis
    n = res_x * res_y * res_z * particles_per_cell
    cell_max = int(1.6 * particles_per_cell)
    u = Z(3, res_x, res_y, res_z)
    u_x, u_y, u_z = Z(res_x + 1, res_y, res_z), Z(res_x, res_y + 1, res_z), Z(res_x, res_y, res_z + 1)
    p2g_weight_x, p2g_weight_y, p2g_weight_z = torch.zeros_like(u_x), torch.zeros_like(u_y), torch.zeros_like(u_z)
    psi_x_grid, psi_y_grid, psi_z_grid = Z(3, *u_x.shape), Z(3, *u_y.shape), Z(3, *u_z.shape)
    T_x_grid, T_y_grid, T_z_grid = Z(3, *u_x.shape), Z(3, *u_y.shape), Z(3, *u_z.shape)
    particles_pos, particles_init_imp, particles_init_imp_grad_m = Z(3, n), Z(3, n), Z(3, n)
    T_x_init, T_y_init, T_z_init = Z(3, n), Z(3, n), Z(3, n)
    T_x_grad_m, T_y_grad_m, T_z_grad_m = Z(3, n), Z(3, n), Z(3, n)
    init_C_x, init_C_y, init_C_z = Z(3, n), Z(3, n), Z(3, n)
    cell_particles_id = Z(res_x, res_y, res_z, cell_max, dtype=torch.int32)
    cell_particle_num = Z(res_x, res_y, res_z, dtype=torch.int32)
    X = None
    solver = MGPCG_3([[2, 2], [2, 2], [2, 2]], N=[res_x, res_y, res_z], base_level=3)

    res = (res_x, res_y, res_z)
    vc = 0.05 * oracle.get_central(res, smooth_mac_field(res, seed=36))
    u.copy_(torch.from_numpy(vc))
    TU.split_central_vector(u, u_x, u_y, u_z)
    solver.Poisson(u_x, u_y, u_z)
    osim = oracle.OracleSim(res, reinit_every=reinit_every, reinit_every_grad_m=reinit_every_grad_m, CFL=CFL,
                            visualize_dt=visualize_dt, particles_per_cell=8, base_level=3)
    osim.set_velocity_from_central(vc)

    sub_t, frame_idx, i, ik = 0.0, 0, -1, 0
    for _ in range(6):
        i += 1
        j = i % reinit_every
        R.calc_max_speed(u_x, u_y, u_z)
        curr_dt = CFL * dx / float(R.max_speed[0])
        if sub_t + curr_dt >= visualize_dt:
            curr_dt = visualize_dt - sub_t
            sub_t = 0.0
            frame_idx += 1
        else:
            sub_t += curr_dt
        if j == 0:
            IC.init_particles_pos_uniform(particles_pos, X, res_x, res_y, particles_per_cell, dx,
                                          particles_per_cell_axis, dist_between_neighbor)
            ik = i
            IC.init_particles_imp(particles_init_imp, particles_pos, u_x, u_y, u_z, init_C_x, init_C_y, init_C_z,
                                  dx)
            R.reset_T_to_identity(T_x_init, T_y_init, T_z_init)
            R.reset_T_to_identity(T_x_grad_m, T_y_grad_m, T_z_grad_m)
        k = (i - ik) % reinit_every_grad_m
        if k == 0:
            IC.init_particles_imp_grad_m(particles_init_imp_grad_m, particles_pos, u_x, u_y, u_z,
                                         init_C_x, init_C_y, init_C_z, dx)
            R.update_T(T_x_init, T_y_init, T_z_init, T_x_grad_m, T_y_grad_m, T_z_grad_m)
            R.reset_T_to_identity(T_x_grad_m, T_y_grad_m, T_z_grad_m)
        R.reset_to_identity_grid(psi_x_grid, psi_y_grid, psi_z_grid, T_x_grid, T_y_grid, T_z_grid)
        R.RK4_grid(psi_x_grid, T_x_grid, u_x, u_y, u_z, 0.5 * curr_dt)
        R.RK4_grid(psi_y_grid, T_y_grid, u_x, u_y, u_z, 0.5 * curr_dt)
        R.RK4_grid(psi_z_grid, T_z_grid, u_x, u_y, u_z, 0.5 * curr_dt)
        TU.copy_to(u_x, p2g_weight_x)
        TU.copy_to(u_y, p2g_weight_y)
        TU.copy_to(u_z, p2g_weight_z)
        R.advect_u(p2g_weight_x, p2g_weight_y, p2g_weight_z, u_x, u_y, u_z, T_x_grid, T_y_grid, T_z_grid,
                   psi_x_grid, psi_y_grid, psi_z_grid, dx)
        solver.Poisson(u_x, u_y, u_z)
        R.stretch_T_and_advect_particles(particles_pos, T_x_grad_m, T_y_grad_m, T_z_grad_m, u_x, u_y, u_z, curr_dt)
        R.get_particles_id_in_every_cell(cell_particles_id, cell_particle_num, particles_pos)
        R.P2G(particles_init_imp, particles_init_imp_grad_m, particles_pos, u_x, u_y, u_z, p2g_weight_x,
              p2g_weight_y, p2g_weight_z, T_x_grad_m, T_y_grad_m, T_z_grad_m, T_x_init, T_y_init, T_z_init,
              cell_particles_id, cell_particle_num)
        solver.Poisson(u_x, u_y, u_z)

        dt_o = osim.step()
        assert abs(curr_dt - dt_o) <= 1e-6 * dt_o, (i, curr_dt, dt_o)
        for a, ug in enumerate((u_x, u_y, u_z)):
            err = rel_l2(N(ug), osim.u[a])
            assert err < 1e-5 * (i + 1), (i, a, err)
        assert rel_l2(N(particles_pos), osim.pos) < 1e-6 * (i + 1)
        # the reference's cell lists of the advected positions (first cell_max ids per cell)
        _, counts, start, perm = oracle.bin_particles(res, dx, np.ascontiguousarray(N(particles_pos)))
        np.testing.assert_array_equal(N(cell_particle_num).ravel(), counts)
        ids = oracle.cell_lists(res, counts, perm, start, cell_max)
        np.testing.assert_array_equal(N(cell_particles_id), ids)


def test_reference_2d_loop_with_pfm_names(oracle, pfm):
    from pfm.d2 import init_conditions as IC
    from pfm.d2 import run_2D as R
    from pfm.d2 import taichi_utils as TU
    from pfm.d2.mgpcg import MGPCG_2

    res_x = res_y = 32
    dx = 1.0 / res_y
    CFL, visualize_dt = 0.5, 0.1
    reinit_every, reinit_every_grad_m = 5, 3
    particles_per_cell = 4
    particles_per_cell_axis = int(math.sqrt(particles_per_cell))
    dist_between_neighbor = dx / particles_per_cell_axis
    n = res_x * res_y * particles_per_cell
    u = Z(2, res_x, res_y)
    u_x, u_y = Z(res_x + 1, res_y), Z(res_x, res_y + 1)
    tmp_u_x, tmp_u_y = torch.zeros_like(u_x), torch.zeros_like(u_y)
    p2g_weight_x, p2g_weight_y = torch.zeros_like(u_x), torch.zeros_like(u_y)
    psi_x_grid, psi_y_grid = Z(2, *u_x.shape), Z(2, *u_y.shape)
    T_x_grid, T_y_grid = Z(2, *u_x.shape), Z(2, *u_y.shape)
    particles_pos, psi = Z(2, n), Z(2, n)
    particles_imp, particles_init_imp, particles_init_imp_grad_m = Z(2, n), Z(2, n), Z(2, n)
    T_x, T_y, T_x_init, T_y_init = Z(2, n), Z(2, n), Z(2, n), Z(2, n)
    T_x_grad_m, T_y_grad_m = Z(2, n), Z(2, n)
    C_x, C_y, init_C_x, init_C_y = Z(2, n), Z(2, n), Z(2, n), Z(2, n)
    X = None
    solver = MGPCG_2([[2, 2], [2, 2]], N=[res_x, res_y], base_level=3)

    res = (res_x, res_y)
    vc = 0.05 * oracle.get_central(res, smooth_mac_field(res, seed=37))
    u.copy_(torch.from_numpy(vc))
    TU.split_central_vector(u, u_x, u_y)
    solver.Poisson(u_x, u_y)
    # particle set-up before the loop (2D/run_2D.py:731-738)
    IC.init_particles_pos_uniform(particles_pos, X, res_x, particles_per_cell, dx, particles_per_cell_axis,
                                  dist_between_neighbor)
    IC.init_particles_imp_grad_m(particles_init_imp_grad_m, particles_pos, u_x, u_y, init_C_x, init_C_y, dx)
    R.reset_T_to_identity(T_x_grad_m, T_y_grad_m)
    IC.init_particles_imp(particles_imp, particles_init_imp, particles_pos, u_x, u_y, init_C_x, init_C_y, dx)
    osim = oracle.OracleSim(res, reinit_every=reinit_every, reinit_every_grad_m=reinit_every_grad_m, CFL=CFL,
                            visualize_dt=visualize_dt, particles_per_cell=4, base_level=3)
    osim.set_velocity_from_central(vc)
    osim.initialize_2d_particles()

    sub_t, frame_idx, i, ik = 0.0, 0, -1, 0
    for _ in range(7):
        i += 1
        j = i % reinit_every
        R.calc_max_speed(u_x, u_y)
        curr_dt = CFL * dx / float(R.max_speed[0])
        if sub_t + curr_dt >= visualize_dt:
            curr_dt = visualize_dt - sub_t
            sub_t = 0.0
            frame_idx += 1
        else:
            sub_t += curr_dt
        if j == 0:
            IC.init_particles_pos_uniform(particles_pos, X, res_x, particles_per_cell, dx,
                                          particles_per_cell_axis, dist_between_neighbor)
            ik = i
            IC.init_particles_imp(particles_imp, particles_init_imp, particles_pos, u_x, u_y, init_C_x, init_C_y,
                                  dx)
            R.reset_to_identity(psi, T_x, T_y)
            R.reset_T_to_identity(T_x_init, T_y_init)
            TU.get_central_vector(u_x, u_y, u)
        k = (i - ik) % reinit_every_grad_m
        if k == 0:
            IC.init_particles_imp_grad_m(particles_init_imp_grad_m, particles_pos, u_x, u_y, init_C_x, init_C_y,
                                         dx)
            R.reset_T_to_identity(T_x_grad_m, T_y_grad_m)
            TU.copy_to(T_x, T_x_init)
            TU.copy_to(T_y, T_y_init)
        R.reset_to_identity_grid(psi_x_grid, psi_y_grid, T_x_grid, T_y_grid)
        R.RK4_grid(psi_x_grid, T_x_grid, u_x, u_y, 0.5 * curr_dt)
        R.RK4_grid(psi_y_grid, T_y_grid, u_x, u_y, 0.5 * curr_dt)
        TU.copy_to(u_x, tmp_u_x)
        TU.copy_to(u_y, tmp_u_y)
        R.advect_u(tmp_u_x, tmp_u_y, u_x, u_y, T_x_grid, T_y_grid, psi_x_grid, psi_y_grid, dx)
        solver.Poisson(u_x, u_y)
        R.stretch_T_and_advect_particles(particles_pos, T_x_grad_m, T_y_grad_m, u_x, u_y, curr_dt)
        R.update_particles_grad_m(C_x, C_y, init_C_x, init_C_y, T_x_grad_m, T_y_grad_m)
        R.update_T(T_x, T_y, T_x_init, T_y_init, T_x_grad_m, T_y_grad_m)
        R.update_particles_imp(particles_imp, particles_init_imp, T_x, T_y)
        R.P2G(particles_imp, particles_pos, u_x, u_y, psi, psi_x_grid, psi_y_grid, p2g_weight_x, p2g_weight_y)
        solver.Poisson(u_x, u_y)

        dt_o = osim.step()
        assert abs(curr_dt - dt_o) <= 1e-6 * dt_o, (i, curr_dt, dt_o)
        for a, ug in enumerate((u_x, u_y)):
            err = rel_l2(N(ug), osim.u[a])
            assert err < 1e-5 * (i + 1), (i, a, err)
        assert rel_l2(N(particles_pos), osim.pos) < 1e-6 * (i + 1)
