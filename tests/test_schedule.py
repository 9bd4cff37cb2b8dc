"""Row 30: the driver's reinit schedule equals the oracle's restatement of
3D/run_3D.py:716-776 and 2D/run_2D.py:772-856 (integer logic, bit-exact)."""
import pytest


@pytest.mark.parametrize("re,rg,dim,reseed", [(20, 1, 3, True), (12, 8, 2, True), (12, 4, 3, True),
                                              (20, 8, 2, True), (12, 5, 3, False), (7, 3, 2, False)])
def test_schedule_matches_oracle(oracle, re, rg, dim, reseed):
    from pfm.simulation import StepConfig, reinit_decision

    cfg = StepConfig(reinit_every=re, reinit_every_grad_m=rg, reinit_particle_pos=reseed)
    want = oracle.reinit_schedule(200, re, rg, reseed, dim)
    ik = 0
    for i in range(200):
        lr, rs, sr, ik = reinit_decision(i, ik, cfg, dim)
        assert (lr, rs, sr) == want[i], i
