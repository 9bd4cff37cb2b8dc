"""GPU parity of the full step (reference schedule) against the oracle's
restatement of the main loop: grid velocity after one step within 1e-5
relative L2, the reinit schedule, and the 2D kinetic-energy trajectory."""
import math

import numpy as np
import pytest
import torch

from helpers import rel_l2, smooth_mac_field

pytestmark = pytest.mark.gpu

DEV = "cuda"


def G(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def N(t):
    return t.detach().cpu().numpy()


def _pair(oracle, res, **kw):
    from pfm.simulation import Simulation, StepConfig

    cfg = StepConfig(**kw)
    sim = Simulation(res, cfg)
    osim = oracle.OracleSim(res, reinit_every=cfg.reinit_every, reinit_every_grad_m=cfg.reinit_every_grad_m,
                            CFL=cfg.CFL, visualize_dt=cfg.visualize_dt, particles_per_cell=cfg.particles_per_cell,
                            base_level=cfg.base_level, max_iters=cfg.max_iters, tol=cfg.tol,
                            tol_mode=0 if cfg.tol_mode == "abs" else 1, seed=cfg.seed,
                            particle_order=cfg.particle_order)
    return sim, osim


@pytest.mark.parametrize("res,kw", [
    ((32, 32), dict(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=8, CFL=0.5, tol=1e-6, tol_mode="rel")),
    ((16, 16, 16), dict(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, tol=1e-6,
                        tol_mode="rel")),
    ((16, 16, 16), dict(particles_per_cell=8, reinit_every=3, reinit_every_grad_m=2, CFL=0.5, tol=1e-6,
                        tol_mode="rel", seed=1, particle_order=1)),
])
def test_steps_match_oracle(oracle, pfm, res, kw):
    sim, osim = _pair(oracle, res, **kw)
    d = len(res)
    vcc = oracle.get_central(res, smooth_mac_field(res, seed=30))  # smooth; projected first
    sim.set_velocity_from_central(G(vcc))
    osim.set_velocity_from_central(vcc)
    for a in range(d):
        assert rel_l2(N(sim.u[a]), osim.u[a]) < 1e-5
    if d == 2:
        sim.initialize_2d_particles()
        osim.initialize_2d_particles()
    for step in range(4):
        dt_g = sim.step()
        dt_o = osim.step()
        assert abs(dt_g - dt_o) <= 1e-6 * dt_o
        for a in range(d):
            err = rel_l2(N(sim.u[a]), osim.u[a])
            assert err < 1e-5 * (step + 1), (step, a, err)
        assert rel_l2(N(sim.pos), osim.pos) < 1e-6 * (step + 1)


def test_ke_trajectory_2d_leapfrog(oracle, pfm):
    """2D 128^2, 4 particles/cell, four Gaussian vortices, reinit 12: kinetic
    energy within 1e-4 relative of the oracle over 100 steps."""
    from pfm import core

    res = (128, 128)
    kw = dict(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=8, CFL=0.5, base_level=4, tol=1e-6,
              tol_mode="rel", seed=0, visualize_dt=1e9)
    sim, osim = _pair(oracle, res, **kw)
    cen = [(0.25, 0.45), (0.25, 0.55), (0.35, 0.45), (0.35, 0.55)]
    w = [1.0, -1.0, 1.0, -1.0]
    vc = oracle.point_vortices(res, 1 / 128, cen, w)
    sim.set_velocity_from_central(G(vc))
    osim.set_velocity_from_central(vc)
    sim.initialize_2d_particles()
    osim.initialize_2d_particles()
    worst = 0.0
    for i in range(100):
        sim.step()
        osim.step()
        eg = sim.kinetic_energy()
        eo = oracle.kinetic_energy(res, osim.u)
        worst = max(worst, abs(eg - eo) / eo)
    assert worst < 1e-4, worst


@pytest.mark.parametrize("res,kw", [
    ((48, 40), dict(particles_per_cell=4, reinit_every=4, reinit_every_grad_m=3, CFL=0.5, tol=1e-6, tol_mode="rel",
                    seed=3, particle_order=1, visualize_dt=0.004)),
    ((16, 20, 24), dict(particles_per_cell=8, reinit_every=3, reinit_every_grad_m=2, CFL=0.5, tol=1e-6,
                        tol_mode="rel", seed=1, particle_order=1, write_imp_grad_m=False, visualize_dt=0.01)),
    ((16, 16, 16), dict(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, tol=1e-6,
                        mg_precision=1, tol_mode="rel")),
])
@pytest.mark.parametrize("fused", [True, False])
def test_native_step_equals_python_step(oracle, pfm, res, kw, fused):
    """pfm_scene_step runs the same entry points with the same arguments as
    the Python driver: fields, schedule state and iteration counts are equal
    bit for bit over steps that cross long and short reinits and frame clips."""
    from pfm.simulation import Simulation, StepConfig

    d = len(res)
    if d == 2 and not fused:
        pytest.skip("2D has no fused path")
    vcc = G(oracle.get_central(res, smooth_mac_field(res, seed=33)))
    sims = []
    for native in (False, True):
        sim = Simulation(res, StepConfig(**kw), native=native)
        sim.fused_p2g = fused and d == 3
        sim.set_velocity_from_central(vcc.clone())
        if d == 2:
            sim.initialize_2d_particles()
        sims.append(sim)
    py, nat = sims
    for step in range(7):
        assert py.step() == nat.step()
        assert (py.i, py.ik, py.frame_idx, py.sub_t) == (nat.i, nat.ik, nat.frame_idx, nat.sub_t)
    assert py.poisson_iters == nat.poisson_iters
    fields = ["u", "pos", "init_imp", "T_init", "T_grad_m", "init_C"] + (["imp", "T", "C"] if d == 2 else [])
    for name in fields:
        a, b = getattr(py, name), getattr(nat, name)
        for x, y in (zip(a, b) if isinstance(a, list) else [(a, b)]):
            assert torch.equal(x, y), name
    # a given dt leaves sub_t alone on both
    assert py.step(1e-3) == nat.step(1e-3) == 1e-3
    assert py.sub_t == nat.sub_t
    for x, y in zip(py.u, nat.u):
        assert torch.equal(x, y)


def test_fold_T_equals_separate_update(oracle, pfm):
    """The RK4 + record kernel that applies the next update_T itself gives the
    same step as update_T + reset_T_to_identity run at the next short
    reinitialisation: equal velocities, positions and impulse state, and its
    T_init equals update_T of the unfolded state, bit for bit."""
    from pfm import core
    from pfm.simulation import Simulation, StepConfig

    res = (16, 18, 20)
    kw = dict(particles_per_cell=8, reinit_every=4, reinit_every_grad_m=1, CFL=0.5, tol=1e-6, tol_mode="rel",
              seed=2, particle_order=1, write_imp_grad_m=False)
    vcc = G(oracle.get_central(res, smooth_mac_field(res, seed=34)))
    sims = []
    for fold in (False, True):
        sim = Simulation(res, StepConfig(fold_T=fold, **kw))
        sim.set_velocity_from_central(vcc.clone())
        sims.append(sim)
    a, b = sims
    assert not a.folds_T and b.folds_T
    for step in range(6):
        assert a.step() == b.step()
        for x, y in zip(a.u, b.u):
            assert torch.equal(x, y), step
        assert torch.equal(a.pos, b.pos)
        for x, y in zip(a.init_C, b.init_C):
            assert torch.equal(x, y)
        want = [torch.empty_like(t) for t in a.T_init]
        core.update_T(a.T_grad_m, a.T_init, want)
        for x, y in zip(want, b.T_init):
            assert torch.equal(x, y), step
        for m in range(3):
            eye = torch.zeros_like(b.T_grad_m[m])
            eye[m] = 1.0
            assert torch.equal(b.T_grad_m[m], eye)
    assert a.poisson_iters == b.poisson_iters
