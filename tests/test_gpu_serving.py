"""Serving on the GPU: every job's result is the same alone and in the mix
(bit for bit: the step has no float atomics), admission under a tight memory
budget runs jobs one at a time, and the footprint bounds what a job allocates."""
import pytest
import torch

from pfm import serving
from pfm.serving import Job, Server

pytestmark = pytest.mark.gpu

MIX = [
    Job(id=0, res=(64, 64), steps=7, seed=11),
    Job(id=1, res=(24, 24, 24), steps=5, seed=12),
    Job(id=2, res=(96, 96), steps=6, seed=13),
    Job(id=3, res=(32, 32, 32), steps=4, seed=14),
    Job(id=4, res=(48, 48), steps=9, seed=15),
]


def _serve(slots, jobs=MIX, check=True, **kw):
    print(f"serve slots={slots} jobs={[j.res for j in jobs]} {kw}", flush=True)
    rep = Server(slots=slots, keep_fields=True, **kw).run(jobs, realtime=False)
    for r in rep.results if check else ():
        assert r.error is None, r.error
        assert len(r.step_ms) == r.job.steps
    return rep


@pytest.mark.parametrize("prec", [1, 0])
def test_same_result_alone_and_in_the_mix(prec):
    alone = _serve(1, mg_precision=prec)
    mixed = _serve(len(MIX), mg_precision=prec)
    # the mix really overlapped: some job started before another finished
    spans = sorted((r.started, r.finished) for r in mixed.results)
    assert any(b[0] < a[1] for a, b in zip(spans, spans[1:]))
    for a, m in zip(alone.results, mixed.results):
        assert a.job == m.job
        assert a.poisson_iters == m.poisson_iters
        for ua, um in zip(a.fields, m.fields):
            assert torch.equal(ua, um), f"job {a.job.id} differs between alone and mixed"
        assert a.kinetic_energy == m.kinetic_energy
    s = mixed.summary()
    assert s["jobs"] == len(MIX) and s["failed"] == 0
    assert s["step_ms_p99"] >= s["step_ms_p50"] > 0


def test_result_matches_a_plain_simulation():
    job = MIX[2]
    rep = _serve(2, jobs=[job, MIX[1]])
    sim = serving.build_scene(job, 1)
    for _ in range(job.steps):
        sim.step()
    got = next(r for r in rep.results if r.job == job)
    assert len(got.fields) == len(sim.u) == 2
    for ua, ub in zip(got.fields, sim.u):
        assert torch.equal(ua, ub)


def test_tight_budget_runs_one_at_a_time():
    jobs = [Job(id=i, res=(64, 64), steps=3 + i, seed=20 + i) for i in range(3)]
    budget = 3 * serving.footprint_bytes(jobs[0], 1) // 2  # room for one job, not two
    rep = _serve(4, jobs=jobs, mg_precision=1, mem_budget=budget)
    spans = sorted((r.started, r.finished) for r in rep.results)
    assert all(b[0] >= a[1] - 1e-3 for a, b in zip(spans, spans[1:]))
    over = _serve(2, jobs=[Job(id=9, res=(64, 64), steps=1, seed=1)], check=False, mem_budget=1 << 20)
    assert over.results[0].error.startswith("rejected")


@pytest.mark.parametrize("job", [MIX[2], MIX[3]], ids=["2d", "3d"])
def test_footprint_bounds_torch_allocations(job):
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    sim = serving.build_scene(job, 1)
    torch.cuda.synchronize()
    used = torch.cuda.memory_allocated() - before
    fp = serving.footprint_bytes(job, 1) - serving.mg_bytes(job.res, 3, 1)
    assert 0 < used <= fp
    del sim


def test_batched_lane_same_result_alone_and_in_the_mix():
    """2D jobs served by the batched lane: each job's fields are the same when
    it is served alone and together with the others (a batch's result for a
    scene does not depend on the other scenes; finished scenes are frozen)."""
    mixed = _serve(2, batch_2d=True)
    assert mixed.summary()["batch_2d"] is True
    for job in MIX:
        alone = _serve(1, jobs=[job], batch_2d=True)
        a = alone.results[0]
        m = next(r for r in mixed.results if r.job == job)
        assert a.poisson_iters == m.poisson_iters, job
        for ua, um in zip(a.fields, m.fields):
            assert torch.equal(ua, um), f"job {job.id} differs between alone and mixed"
        assert a.kinetic_energy == m.kinetic_energy


def test_batched_lane_matches_plain_simulation():
    """The batched lane's result for a 2D job agrees with a single-scene
    Simulation of the same job and knobs (only the PCG dot-product order
    differs)."""
    import dataclasses

    from pfm.simulation import Simulation

    job = MIX[2]
    rep = _serve(1, jobs=[job, MIX[4]], batch_2d=True)
    got = next(r for r in rep.results if r.job == job)
    cfg = dataclasses.replace(serving.batch_step_config(), seed=job.seed % (2**31))
    sim = Simulation(job.res, cfg)
    c, w, rad = serving.point_vortex_ic(job)
    sim.init_leapfrog_2d(c, w, rad=rad)
    sim.initialize_2d_particles()
    for _ in range(job.steps):
        sim.step()
    for ua, ub in zip(got.fields, sim.u):
        assert float(torch.linalg.vector_norm(ua - ub) / torch.linalg.vector_norm(ub)) <= 1e-5
