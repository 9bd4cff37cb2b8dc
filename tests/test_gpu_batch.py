"""Batched 2D scenes (pfm.batch.Batch2D) against single-scene runs.

Each scene of a batch is checked against its own single-instance
Simulation on the same initial velocity and seed: the phases before a
projection share their per-item arithmetic with the single-scene kernels
and must match bit for bit; a projection sums its dot products in a
different order, so velocities after one agree to the solver tolerance
(relative L2 <= 1e-5).  A scene's result must not depend on the other
scenes of its batch, on graph replay or on the tuned launch shapes (bit for
bit)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from pfm.batch import KINDS, Batch2D, random_vortex_scenes  # noqa: E402
from pfm.simulation import Simulation, StepConfig  # noqa: E402

RES = [(64, 64), (80, 72), (97, 97), (128, 128), (150, 100), (40, 48)]


def _cfg(**kw):
    base = dict(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=4, CFL=0.5, base_level=4,
                max_iters=400, tol=1e-6, tol_mode="rel", visualize_dt=0.1)
    base.update(kw)
    return StepConfig(**base)


def _vortices(s):
    rng = np.random.default_rng(100 + s)
    k = 8 + s
    return rng.uniform(0.2, 0.8, (k, 2)), rng.normal(0.0, 1.0, k)


def _batch(res, cfg, seeds, graphs=True):
    b = Batch2D(res, cfg, seeds=seeds, graphs=graphs)
    for s in range(len(res)):
        c, w = _vortices(s)
        b.set_point_vortices(s, c, w, rad=1.5 / res[s][1])
    b.project()
    return b


def _single(b, s, cfg):
    """A Simulation of scene s starting from the batch's projected velocity."""
    sim = Simulation(b.res[s], StepConfig(**{**cfg.__dict__, "seed": b.seeds[s]}))
    for a in range(2):
        sim.u[a].copy_(b.u_view(s, a))
    return sim


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b).clamp_min(1e-30))


def test_batch_projection_matches_single():
    cfg = _cfg()
    b = Batch2D(RES, cfg, seeds=list(range(len(RES))))
    sims = []
    for s, res in enumerate(RES):
        c, w = _vortices(s)
        b.set_point_vortices(s, c, w, rad=1.5 / res[1])
        sim = Simulation(res, cfg)
        for a in range(2):
            sim.u[a].copy_(b.u_view(s, a))
        sims.append(sim)
    b.project()
    info = b.scene_info()
    for s, sim in enumerate(sims):
        it = sim.project()
        for a in range(2):
            assert _rel(b.u_view(s, a), sim.u[a]) <= 1e-5, (s, a)
        assert abs(int(info["iters"][s, 0]) - it) <= 1, (s, info["iters"][s, 0], it)
        assert info["converged"][s, 0] == 1


def test_batch_step_phases_bit_exact_without_midpoint():
    cfg = _cfg(use_midpoint_vel=False)
    b = _batch(RES, cfg, seeds=[7 * s + 1 for s in range(len(RES))])
    sims = [_single(b, s, cfg) for s in range(len(RES))]
    b.init_particles()
    for sim in sims:
        sim.initialize_2d_particles()
    b.step()
    info = b.scene_info()
    for s, sim in enumerate(sims):
        dt = sim.step()
        assert info["dt"][s] == dt, s
        for name in ("pos", "imp", "init_imp"):
            assert torch.equal(b.particle_view(s, getattr(b, name)), getattr(sim, name)), (s, name)
        for a in range(2):
            for bt, st in ((b.T_grad_m, sim.T_grad_m), (b.T, sim.T), (b.T_init, sim.T_init), (b.Cm, sim.C),
                           (b.init_C, sim.init_C)):
                assert torch.equal(b.particle_view(s, bt[a]), st[a]), (s, a)
            assert _rel(b.u_view(s, a), sim.u[a]) <= 1e-5, (s, a)


def test_batch_midpoint_step_matches_single():
    cfg = _cfg()
    b = _batch(RES, cfg, seeds=[3 * s for s in range(len(RES))])
    sims = [_single(b, s, cfg) for s in range(len(RES))]
    b.init_particles()
    for sim in sims:
        sim.initialize_2d_particles()
    b.step()
    info = b.scene_info()
    for s, sim in enumerate(sims):
        dt = sim.step()
        assert info["dt"][s] == dt, s
        # the face flow maps come from the unprojected velocity: bit for bit
        for a in range(2):
            o, n = b.f_off[a][s], b.nf[s][a]
            assert torch.equal(b.psi_grid[a][:, o:o + n], sim.psi_grid[a].reshape(2, -1)), (s, a)
            assert torch.equal(b.T_grid[a][:, o:o + n], sim.T_grid[a].reshape(2, -1)), (s, a)
            assert _rel(b.u_view(s, a), sim.u[a]) <= 1e-5, (s, a)
        assert all(abs(int(info["iters"][s, k]) - sim.poisson_iters[-2 + k]) <= 1 for k in range(2)), s


def test_batch_long_run_tracks_single():
    """13 steps: the long reinitialisation at step 12 and the short ones."""
    res = [(32, 32), (40, 48), (64, 40)]
    cfg = _cfg()
    b = _batch(res, cfg, seeds=[11, 12, 13])
    sims = [_single(b, s, cfg) for s in range(len(res))]
    b.init_particles()
    for sim in sims:
        sim.initialize_2d_particles()
    for _ in range(13):
        b.step()
        for sim in sims:
            sim.step()
    for s, sim in enumerate(sims):
        for a in range(2):
            assert _rel(b.u_view(s, a), sim.u[a]) <= 1e-4, (s, a)


def _run(res, seeds, steps, graphs=True, tune=False):
    cfg = _cfg(reinit_every_grad_m=2)
    b = _batch(res, cfg, seeds=seeds, graphs=graphs)
    if tune:
        b.tune()
    b.init_particles()
    for _ in range(steps):
        b.step()
    return b


def test_batch_result_independent_of_batch_graphs_and_tuning():
    seeds = [5, 6, 7, 8]
    res = RES[:4]
    full = _run(res, seeds, 3)
    alone = Batch2D([res[2]], _cfg(reinit_every_grad_m=2), seeds=[seeds[2]])
    c, w = _vortices(2)
    alone.set_point_vortices(0, c, w, rad=1.5 / res[2][1])
    alone.project()
    alone.init_particles()
    for _ in range(3):
        alone.step()
    for a in range(2):
        assert torch.equal(full.u_view(2, a), alone.u_view(0, a)), a
    assert torch.equal(full.particle_view(2, full.pos), alone.particle_view(0, alone.pos))
    eager = _run(res, seeds, 3, graphs=False)
    tuned = _run(res, seeds, 3, tune=True)
    shapes = tuned.tuning()
    assert set(KINDS) <= set(shapes) and shapes["coarse_threads"] in (256, 512, 1024)
    for other in (eager, tuned):
        for a in range(2):
            assert torch.equal(full.u[a], other.u[a]), a
        assert torch.equal(full.pos, other.pos)


def test_random_vortex_workload_steps():
    scenes = random_vortex_scenes(24, seed=2, lo=64, hi=256)
    cfg = _cfg()
    b = Batch2D([sc[0] for sc in scenes], cfg, seeds=list(range(24)))
    for s, (res, c, w, rad) in enumerate(scenes):
        b.set_point_vortices(s, c, w, rad)
    b.project()
    b.init_particles()
    for _ in range(2):
        b.step()
    info = b.scene_info()
    assert (info["converged"] == 1).all()
    for a in range(2):
        assert torch.isfinite(b.u[a]).all()


@pytest.mark.parametrize("iters", [1, 2, 3])
def test_batch_pcg_iterations_match_single(iters):
    """A fixed number of PCG iterations: the V-cycle of every level (tiled
    down/up launches and the per-scene coarse kernel) reproduces the
    single-scene solver's, so the velocities agree far below the solver
    tolerance (only alpha/beta may round differently)."""
    cfg = _cfg(max_iters=iters, tol=1e-30)
    b = Batch2D(RES, cfg, seeds=list(range(len(RES))))
    sims = []
    for s, res in enumerate(RES):
        c, w = _vortices(s)
        b.set_point_vortices(s, c, w, rad=1.5 / res[1])
        sim = Simulation(res, cfg)
        for a in range(2):
            sim.u[a].copy_(b.u_view(s, a))
        sims.append(sim)
    b.project()
    for s, sim in enumerate(sims):
        sim.project()
        for a in range(2):
            assert _rel(b.u_view(s, a), sim.u[a]) <= 1e-6, (s, a, _rel(b.u_view(s, a), sim.u[a]))


def test_large_batch_bit_reproducible():
    """Many scenes with many tiles each (the V-cycle launches then run tiles
    of one level in many waves): the batch gives the same bits with and
    without graph replay, and a scene gives the same bits alone."""
    scenes = random_vortex_scenes(48, seed=4, lo=128, hi=512)
    cfg = _cfg(reinit_every_grad_m=2, particle_order=1)

    def run(idx, graphs=True):
        b = Batch2D([scenes[k][0] for k in idx], cfg, seeds=list(idx), graphs=graphs)
        for j, k in enumerate(idx):
            res, c, w, rad = scenes[k]
            b.set_point_vortices(j, c, w, rad)
        b.project()
        b.init_particles()
        for _ in range(2):
            b.step()
        return b

    full = run(list(range(48)))
    eager = run(list(range(48)), graphs=False)
    for a in range(2):
        assert torch.equal(full.u[a], eager.u[a]), a
    big = max(range(48), key=lambda k: scenes[k][0][0])
    alone = run([big])
    for a in range(2):
        assert torch.equal(full.u_view(big, a), alone.u_view(0, a)), a
