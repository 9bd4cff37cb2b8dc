"""GPU checks at the benchmark workloads' own sizes.

* 3D 128^3, two coaxial rings, 8 particles/cell (16.8 M), seed 1, reinit
  20/1, 5 multigrid levels at a 1e-7 relative residual -- the reference main
  loop 3D/run_3D.py:716-831 at that size: two steps against the oracle's
  restatement of the loop (velocity, dt, positions per step; the cell sort
  of the advected positions bit for bit), then 50 steps on the GPU alone with
  size-independent properties.
* 3D 256^3 obstacle Poisson (free-slip walls, solid box [0.4, 0.6]^3, RHS the
  divergence of a 64-mode Fourier velocity, seed 0) for fp32, fp16 and bf16
  level storage against the FP64 oracle solve (3D/mgpcg.py:202-280); one
  256^3 x 8 particles/cell full step: reproducible bit for bit from the same
  state, divergence-free to the solver tolerance, its cell sort equal to a
  host stable argsort.
* 512 independent 2D scenes 64^2-512^2 (seed 2) advanced as one batch for two
  steps: a size-stratified sample against single-scene runs, the two smallest
  scenes against the oracle.
* A served job stream drawn at the stream's own sizes (2D 128^2-1024^2, 3D
  64^3-192^3, seed 3) with step counts capped: every job's fields are the
  same alone and in the mix, in the per-job slots and in the batched 2D lane.
"""
import math

import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def N(t):
    return t.detach().cpu().numpy()


def rings_3d():
    """Two coaxial rings: major radius 0.15 / 0.2 at z = 0.3 / 0.4, core 0.02,
    unit circulation (add_vortex_ring, 3D/init_conditions.py:10-22)."""
    return [dict(w=1.0, rad=0.15, delta=0.02, c=[0.5, 0.5, 0.3], unit_x=[1.0, 0, 0], unit_y=[0, 1.0, 0]),
            dict(w=1.0, rad=0.20, delta=0.02, c=[0.5, 0.5, 0.4], unit_x=[1.0, 0, 0], unit_y=[0, 1.0, 0])]


def _ring_velocity(res):
    from pfm import core

    vc = torch.zeros((3,) + tuple(res), dtype=torch.float32, device=DEV)
    for r in rings_3d():
        core.vortex_ring(res, 1.0 / res[1], vc, r["w"], r["rad"], r["delta"], r["c"], r["unit_x"], r["unit_y"],
                         2000)
    return vc


def _divergence(u):
    """Undivided MAC divergence (3D/mgpcg.py:380-390) in float64."""
    d = len(u)
    div = None
    for a in range(d):
        x = u[a].double()
        t = x.narrow(a, 1, x.shape[a] - 1) - x.narrow(a, 0, x.shape[a] - 1)
        div = t if div is None else div + t
    return div


CFG2 = dict(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, base_level=3, tol=1e-7,
            tol_mode="rel", seed=1, particle_order=1, visualize_dt=0.1)


@pytest.mark.timeout(900)
def test_rings_128_two_steps_vs_oracle(oracle, pfm):
    from pfm.simulation import Simulation, StepConfig

    res = (128, 128, 128)
    print("rings 128^3 vs oracle: set-up", flush=True)
    cfg = StepConfig(**CFG2)
    sim = Simulation(res, cfg)
    assert sim.solver.n_mg_levels == 5
    osim = oracle.OracleSim(res, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, visualize_dt=0.1,
                            particles_per_cell=8, base_level=3, tol=1e-7, tol_mode=1, seed=1, particle_order=1)
    vc = _ring_velocity(res)
    sim.set_velocity_from_central(vc)
    osim.set_velocity_from_central(N(vc))
    for a in range(3):
        assert rel_l2(N(sim.u[a]), osim.u[a]) < 1e-5, a
    for step in range(2):
        print(f"rings 128^3 vs oracle: step {step}", flush=True)
        dt_g = sim.step()
        dt_o = osim.step()
        assert abs(dt_g - dt_o) <= 1e-6 * dt_o, (step, dt_g, dt_o)
        assert sim.poisson_iters[-1] >= 1 and sim.solver.converged
        for a in range(3):
            err = rel_l2(N(sim.u[a]), osim.u[a])
            assert err < 1e-5 * (step + 1), (step, a, err)
        pos = N(sim.pos)
        assert rel_l2(pos, osim.pos) < 2e-6 * (step + 1)
        # the cell sort of the advected positions that P2G used, bit for bit
        # with the oracle's binning of the same positions
        _, counts, start, perm = oracle.bin_particles(res, 1.0 / 128, np.ascontiguousarray(pos))
        np.testing.assert_array_equal(N(sim.binning.perm), perm)
        np.testing.assert_array_equal(N(sim.binning.cell_start), start)


@pytest.mark.timeout(300)
def test_rings_128_fifty_steps_properties(pfm):
    """The 50-step run of the workload on the GPU alone: finite fields, every
    projection converged and divergence-free to its tolerance, particles in
    the domain, kinetic energy bounded by its initial value (the flow has no
    forcing), and the flow-map schedule (two long reinitialisations)."""
    from pfm.simulation import Simulation, StepConfig

    res = (128, 128, 128)
    sim = Simulation(res, StepConfig(**CFG2))
    sim.set_velocity_from_central(_ring_velocity(res))
    ke0 = sim.kinetic_energy()
    for step in range(50):
        sim.step()
        if step % 10 == 9:
            print(f"rings 128^3 step {step + 1} KE {sim.kinetic_energy():.6e}", flush=True)
        assert sim.solver.converged, step
    assert sim.i == 49 and sim.ik == 40
    for a in range(3):
        assert torch.isfinite(sim.u[a]).all()
    div = _divergence(sim.u)
    # |div| relative to the face differences that make it up
    scale = sum(float(torch.linalg.vector_norm(x.double())) for x in sim.u)
    assert float(torch.linalg.vector_norm(div)) < 1e-5 * scale
    # the reference neither clamps nor reflects particles at the walls; the
    # wall-normal velocity interpolated near a wall is small, so particles stay
    # within half a cell (the CFL bound of one step) of the domain
    dx = 1.0 / 128
    assert float(sim.pos.min()) >= -0.5 * dx and float(sim.pos.max()) <= 1.0 + 0.5 * dx
    ke = sim.kinetic_energy()
    assert 0.8 * ke0 < ke < 1.05 * ke0, (ke0, ke)


# ---- 256^3 obstacle Poisson and full step ---------------------------------
@pytest.fixture(scope="module")
def obstacle_256(oracle):
    from pfm import workloads

    res = (256, 256, 256)
    print("256^3 obstacle Poisson: right-hand side and FP64 oracle solve", flush=True)
    u = workloads.fourier_velocity(res, seed=0, device=DEV)
    solid = workloads.box_mask(res, device=DEV)
    sol = N(solid).astype(bool)
    mg_ref = oracle.MG(res, base_level=3)
    ub = [N(x).copy() for x in u]
    mg_ref.apply_bc(ub)
    mg_ref.set_solid(sol)
    mg_ref.apply_solid_bc(ub)
    mg_ref.divergence(ub)
    rhs = mg_ref.field("u_div", 0).astype(np.float64).copy()
    del mg_ref, ub
    x64, it, conv = oracle.poisson_fp64_solve(rhs, base_level=3, tol=1e-8, solid=sol)
    fl = ~sol
    x64 = x64[fl] - x64[fl].mean()
    return res, u, solid, fl, x64


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_obstacle_poisson_256_vs_fp64(pfm, obstacle_256, precision):
    res, u0, solid, fl, x64 = obstacle_256
    print(f"256^3 obstacle Poisson precision {precision}", flush=True)
    mg = pfm.MGPCG([[2, 2]] * 3, list(res), dim=3, base_level=3, precision=precision)
    mg.set_solid(solid)
    u = [x.clone() for x in u0]
    it = mg.Poisson(*u, tol=1e-6, tol_mode="rel")
    assert mg.converged, it
    p = N(mg.p).astype(np.float64)
    assert np.all(p[~fl] == 0)
    pf = p[fl] - p[fl].mean()
    err = float(np.linalg.norm(pf - x64) / np.linalg.norm(x64))
    print(f"  iterations {it}, pressure rel L2 vs FP64 {err:.3e}", flush=True)
    assert err <= 1e-5, err
    # projected: discrete divergence at the solver tolerance, obstacle faces zero
    div = _divergence(u)
    scale = sum(float(torch.linalg.vector_norm(x.double())) for x in u)
    assert float(torch.linalg.vector_norm(div)) < 1e-5 * scale


@pytest.mark.timeout(600)
def test_full_step_256_reproducible(pfm):
    """One 256^3 x 8 particles/cell step (134 M particles) from the same state
    twice: equal bits; the result is divergence-free to the solver tolerance;
    the cell sort P2G used equals a host stable argsort of the cell ids."""
    from pfm.simulation import Simulation, StepConfig

    res = (256, 256, 256)
    cfg = StepConfig(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, tol=1e-6,
                     tol_mode="rel", seed=1, particle_order=1, write_imp_grad_m=False)
    vc = _ring_velocity(res)
    out = []
    for k in range(2):
        print(f"256^3 full step, run {k}", flush=True)
        sim = Simulation(res, cfg)
        sim.set_velocity_from_central(vc)
        sim.step()
        sim.step()
        torch.cuda.synchronize()
        out.append(([x.clone() for x in sim.u], sim.pos.clone(), list(sim.poisson_iters)))
        if k == 0:
            cell = N(sim.binning.cell_of).astype(np.int64)
            pos = sim.pos
            ci = torch.clamp((pos * 256).to(torch.int32), 0, 255).to(torch.int64)
            want_cell = N((ci[0] * 256 + ci[1]) * 256 + ci[2])
            np.testing.assert_array_equal(cell, want_cell)
            np.testing.assert_array_equal(N(sim.binning.perm), np.argsort(want_cell, kind="stable"))
            div = _divergence(sim.u)
            scale = sum(float(torch.linalg.vector_norm(x.double())) for x in sim.u)
            assert float(torch.linalg.vector_norm(div)) < 1e-5 * scale
            for a in range(3):
                assert torch.isfinite(sim.u[a]).all()
        del sim
        torch.cuda.empty_cache()
    (ua, pa, ia), (ub, pb, ib) = out
    assert ia == ib
    assert torch.equal(pa, pb)
    for x, y in zip(ua, ub):
        assert torch.equal(x, y)


# ---- 512 batched 2D scenes --------------------------------------------------
@pytest.mark.timeout(600)
def test_batch_512_scenes_vs_single_and_oracle(oracle, pfm):
    from pfm.batch import Batch2D, random_vortex_scenes
    from pfm.simulation import Simulation, StepConfig

    n_scenes = 512
    cfg = StepConfig(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=8, CFL=0.5, base_level=4,
                     tol=1e-6, tol_mode="rel", write_imp_grad_m=False, visualize_dt=1e9, particle_order=1)
    scenes = random_vortex_scenes(n_scenes, seed=2)
    print("512-scene batch: set-up", flush=True)
    b = Batch2D([sc[0] for sc in scenes], cfg, seeds=list(range(n_scenes)))
    for k, (res, c, w, rad) in enumerate(scenes):
        b.set_point_vortices(k, c, w, rad)
    b.project()
    sizes = [sc[0][0] for sc in scenes]
    order = np.argsort(sizes, kind="stable")
    picked = [int(order[i]) for i in np.linspace(0, n_scenes - 1, 16).round().astype(int)]
    small = [int(order[0]), int(order[1])]
    ic = {k: [b.u_view(k, a).clone() for a in range(2)] for k in set(picked) | set(small)}
    b.init_particles()
    for _ in range(2):
        b.step()
    torch.cuda.synchronize()
    info = b.scene_info()
    assert (info["converged"] == 1).all()
    for k in picked:
        sim = Simulation(scenes[k][0], StepConfig(**{**cfg.__dict__, "seed": k}))
        for a in range(2):
            sim.u[a].copy_(ic[k][a])
        sim.initialize_2d_particles()
        for _ in range(2):
            sim.step()
        for a in range(2):
            err = rel_l2(N(b.u_view(k, a)), N(sim.u[a]))
            assert err <= 1e-5, (k, sizes[k], a, err)
        del sim
    for k in small:
        res = scenes[k][0]
        print(f"512-scene batch: scene {k} {res} vs oracle", flush=True)
        osim = oracle.OracleSim(res, reinit_every=12, reinit_every_grad_m=8, CFL=0.5, visualize_dt=1e9,
                                particles_per_cell=4, base_level=4, tol=1e-6, tol_mode=1, seed=k, particle_order=1)
        osim.u = [N(x).copy() for x in ic[k]]
        osim.initialize_2d_particles()
        for _ in range(2):
            osim.step()
        for a in range(2):
            err = rel_l2(N(b.u_view(k, a)), osim.u[a])
            assert err <= 2e-5, (k, a, err)


# ---- served stream at its own sizes --------------------------------------
@pytest.mark.timeout(600)
def test_stream_jobs_same_alone_and_in_the_mix(pfm):
    from pfm import serving
    from pfm.serving import Job, Server

    drawn = serving.job_stream(24, rate=50.0, seed=3)
    # the largest and a mid-size 2D and 3D jobs of the draw: 2D 784^2, 680^2,
    # 512^2, 344^2 and 3D 168^3, 144^3, 88^3, 10 steps each
    picked = [drawn[i] for i in (13, 0, 23, 1, 11, 9, 5)]
    assert [j.res[0] for j in picked] == [784, 680, 512, 344, 168, 144, 88]
    jobs = [Job(id=j.id, res=j.res, steps=10, seed=j.seed, arrival=0.0) for j in picked]
    print(f"stream jobs {[j.res for j in jobs]}", flush=True)
    for batch_2d in (False, True):
        mixed = Server(slots=8, keep_fields=True, batch_2d=batch_2d).run(jobs, realtime=False)
        assert mixed.summary()["failed"] == 0, [(r.job.id, r.error) for r in mixed.results if r.error]
        for job in jobs:
            print(f"  batch_2d={batch_2d} job {job.id} {job.res} alone", flush=True)
            alone = Server(slots=1, keep_fields=True, batch_2d=batch_2d).run([job], realtime=False)
            a = alone.results[0]
            m = next(r for r in mixed.results if r.job == job)
            assert a.error is None and m.error is None
            assert a.poisson_iters == m.poisson_iters, job
            for ua, um in zip(a.fields, m.fields):
                assert torch.equal(ua, um), f"job {job.id} differs between alone and mixed"
