"""GPU: curl bit-exact against the oracle, and checkpoint/resume reproducing
the uninterrupted run bit for bit."""
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


@pytest.mark.parametrize("res", [(16, 12, 10), (20, 16)])
def test_curl_bit_exact(oracle, pfm, res):
    from pfm import core

    d = len(res)
    dx = 1.0 / res[1]
    vc = np.random.default_rng(5).normal(size=(d,) + res).astype(np.float32)
    out = torch.empty((3,) + res if d == 3 else res, device=DEV)
    core.curl(res, dx, G(vc), out)
    np.testing.assert_array_equal(N(out), oracle.curl(res, dx, vc))


@pytest.mark.parametrize("res,kw", [
    ((32, 32), dict(particles_per_cell=4, reinit_every=3, reinit_every_grad_m=2, tol=1e-6, tol_mode="rel",
                    seed=0, particle_order=1)),
    ((16, 16, 16), dict(particles_per_cell=8, reinit_every=3, reinit_every_grad_m=1, tol=1e-6, tol_mode="rel",
                        seed=1, particle_order=1, write_imp_grad_m=False)),
])
def test_resume_is_bit_identical(oracle, pfm, tmp_path, res, kw):
    from pfm import checkpoint
    from pfm.simulation import Simulation, StepConfig

    d = len(res)
    sim = Simulation(res, StepConfig(**kw))
    vcc = oracle.get_central(res, smooth_mac_field(res, seed=31))
    sim.set_velocity_from_central(G(vcc))
    if d == 2:
        sim.initialize_2d_particles()
    for _ in range(4):
        sim.step()
    checkpoint.save_state(sim, str(tmp_path / "state"))
    checkpoint.save_velocity(str(tmp_path / "ckpt"), 4, sim.u)
    saved = (sim.i, sim.ik, sim.sub_t)
    u_saved = [t.clone() for t in sim.u]
    for _ in range(3):
        sim.step()
    torch.cuda.synchronize()
    back = checkpoint.load_state(str(tmp_path / "state"))
    assert (back.i, back.ik, back.sub_t) == saved
    for _ in range(3):
        back.step()
    torch.cuda.synchronize()
    for a in range(d):
        assert torch.equal(back.u[a], sim.u[a])
    assert torch.equal(back.pos, sim.pos)
    # the reference's velocity files restore the saved velocity exactly
    u2 = [torch.zeros_like(t) for t in sim.u]
    checkpoint.load_velocity(str(tmp_path / "ckpt"), 4, u2)
    for a in range(d):
        assert torch.equal(u2[a], u_saved[a])


@pytest.mark.parametrize("res,kw", [
    ((32, 32), dict(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=8, tol=1e-6, tol_mode="rel",
                    seed=0)),
    ((16, 16, 16), dict(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, tol=1e-6, tol_mode="rel",
                        seed=1, particle_order=1)),
])
def test_resume_from_frame_vs_oracle(oracle, pfm, tmp_path, res, kw):
    """from_frame > 0 as the reference does it (3D/run_3D.py:676-680, 708-713;
    2D/run_2D.py:726-738): velocity and smoke from the frame's files, flow
    maps restarted.  The GPU and the oracle start from the same files and
    step on together."""
    from pfm import checkpoint
    from pfm.simulation import Simulation, StepConfig

    d = len(res)
    sim = Simulation(res, StepConfig(**kw))
    sim.set_velocity_from_central(G(oracle.get_central(res, smooth_mac_field(res, seed=35))))
    if d == 2:
        sim.initialize_2d_particles()
    for _ in range(3):
        sim.step()
    ck = str(tmp_path / "ckpt")
    checkpoint.save_velocity(ck, 7, sim.u)
    smoke = torch.rand((4,) + res, generator=torch.Generator(device="cpu").manual_seed(1)).to(DEV)
    checkpoint.save_smoke(ck, 7, smoke)
    # the files are the reference's: (*face shape) float32 and (*res, 4) smoke
    for a, path in enumerate(checkpoint.velocity_files(ck, 7, d)):
        arr = np.load(path, allow_pickle=False)
        assert arr.dtype == np.float32 and arr.shape == tuple(sim.u[a].shape)
        np.testing.assert_array_equal(arr, N(sim.u[a]))
    sm = np.load(checkpoint.smoke_file(ck, 7), allow_pickle=False)
    assert sm.shape == res + (4,)
    np.testing.assert_array_equal(np.moveaxis(sm, -1, 0), N(smoke))

    back = Simulation(res, StepConfig(**kw))
    smoke2 = torch.zeros_like(smoke)
    checkpoint.resume_from_frame(back, ck, 7, smoke=smoke2)
    assert torch.equal(smoke2, smoke)
    assert (back.i, back.ik, back.sub_t, back.frame_idx) == (-1, 0, 0.0, 7)
    osim = oracle.OracleSim(res, reinit_every=kw["reinit_every"], reinit_every_grad_m=kw["reinit_every_grad_m"],
                            particles_per_cell=kw["particles_per_cell"], tol=1e-6, tol_mode=1, seed=kw["seed"],
                            particle_order=kw.get("particle_order", 0))
    osim.u = [np.load(p, allow_pickle=False) for p in checkpoint.velocity_files(ck, 7, d)]
    if d == 2:
        osim.initialize_2d_particles()
    for step in range(3):
        dt_g, dt_o = back.step(), osim.step()
        assert abs(dt_g - dt_o) <= 1e-6 * dt_o
        for a in range(d):
            err = rel_l2(N(back.u[a]), osim.u[a])
            assert err < 1e-5 * (step + 1), (step, a, err)
    with pytest.raises(ValueError):
        checkpoint.resume_from_frame(back, ck, 0)
