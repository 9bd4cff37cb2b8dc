"""The reference's named 3D scenes (3D/init_conditions.py:81-180) built on
the GPU through pfm.d3: each against the oracle's ring kernels called with
the scene's rings, plus the geometry that fixes the scene (every ring's
self-induced velocity points along its axis: the two leapfrog rings travel
together along +x, the eight rings all travel toward the cube's centre)."""
import math

import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def N(t):
    return t.detach().cpu().numpy()


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _oracle_scene(oracle, res, rings, num_samples):
    """u (3, *res) and smoke1 (4, *res) of rings added one by one, each ring's
    smoke thresholded in its own scratch field and summed into smoke1."""
    dx = 1.0 / res[1]
    u = np.zeros((3,) + res, np.float32)
    s1 = np.zeros((4,) + res, np.float32)
    for w, rad, delta, c, ex, ey, color in rings:
        oracle.vortex_ring(res, dx, u, w, rad, delta, c, ex, ey, num_samples)
        s = np.zeros((4,) + res, np.float32)
        oracle.vortex_ring_smoke(res, dx, s, rad, delta, c, ex, ey, color, num_samples)
        s1 += s
    return u, s1


def _velocity_at(u, res, p):
    """The cell-centred velocity of the cell containing p."""
    dx = 1.0 / res[1]
    i = [min(max(int(x / dx), 0), n - 1) for x, n in zip(p, res)]
    return u[:, i[0], i[1], i[2]].astype(np.float64)


def _check_smoke(got, want):
    # the 0.002 density threshold may flip within rounding of it (device and
    # host exp/sin/cos differ in the last bits)
    diff = np.any(got != want, axis=0)
    assert diff.sum() <= max(2, diff.size // 1000), diff.sum()


def test_init_vorts_leapfrog(oracle, pfm):
    from pfm.d3 import hyperparameters as hp
    from pfm.d3 import init_conditions as ic

    assert hp.knobs(0)["reinit_every"] == 20 and hp.knobs(0)["res_x"] == 2 * hp.knobs(0)["res_y"]
    res = (64, 32, 32)  # case 0's 2:1:1 box, [0, 2] x [0, 1]^2
    u = torch.zeros((3,) + res, device=DEV)
    s1 = torch.zeros((4,) + res, device=DEV)
    s2 = torch.zeros((4,) + res, device=DEV)
    ic.init_vorts_leapfrog(None, u, s1, s2)
    r = 0.21
    ex, ey = [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]
    rings = [(0.1 * r, r, 0.08 * r, [0.16, 0.5, 0.5], ex, ey, [1, 0, 0]),
             (0.1 * r, r, 0.08 * r, [0.16 + 0.625 * r, 0.5, 0.5], ex, ey, [0, 0, 1])]
    wu, ws = _oracle_scene(oracle, res, rings, 2000)
    assert rel_l2(N(u), wu) < 1e-5
    _check_smoke(N(s1), ws)
    # both rings' axis is e_x (ex x ey); their centres move along +x
    assert np.allclose(np.cross(ex, ey), [1, 0, 0])
    for ring in rings:
        v = _velocity_at(N(u), res, ring[3])
        assert v[0] > 0 and abs(v[0]) > 3 * max(abs(v[1]), abs(v[2])), v


def test_init_vorts_eight(oracle, pfm):
    from pfm.d3 import hyperparameters as hp
    from pfm.d3 import init_conditions as ic

    k1 = hp.knobs(1)
    assert (k1["res_x"], k1["reinit_every"], k1["reinit_every_grad_m"]) == (128, 12, 4)
    res = (48, 48, 48)
    u = torch.zeros((3,) + res, device=DEV)
    s1 = torch.full((4,) + res, 7.0, device=DEV)  # zeroed by the scene
    s2 = torch.full((4,) + res, 7.0, device=DEV)
    ic.init_vorts_eight(None, u, s1, s2)
    c = math.cos(math.pi / 4)
    ct = math.sqrt(2)
    spec = [(+1, (-1, -1, -1), (-c, c, 0), (-c, -c, ct), (1, 0, 0)),
            (-1, (+1, -1, -1), (c, c, 0), (c, -c, ct), (0, 1, 0)),
            (+1, (-1, +1, -1), (c, c, 0), (-c, c, ct), (0, 0, 1)),
            (-1, (+1, +1, -1), (-c, c, 0), (c, c, ct), (0, 0, 0)),
            (+1, (-1, -1, +1), (-c, c, 0), (c, c, ct), (0, 1, 1)),
            (-1, (+1, -1, +1), (c, c, 0), (-c, c, ct), (1, 0, 1)),
            (+1, (-1, +1, +1), (c, c, 0), (c, -c, ct), (1, 1, 0)),
            (-1, (+1, +1, +1), (-c, c, 0), (-c, -c, ct), (1, 1, 1))]
    rings = []
    for sgn, corner, ex, ey, color in spec:
        centre = [0.5 + 0.16 * s for s in corner]
        rings.append((0.02 * sgn, 0.08, 0.024, centre, list(_unit(ex)), list(_unit(ey)), list(color)))
    wu, ws = _oracle_scene(oracle, res, rings, 500)
    assert rel_l2(N(u), wu) < 1e-5
    _check_smoke(N(s1), ws)
    # smoke2 keeps the last ring's own smoke (the reference does not clear it)
    last = np.zeros((4,) + res, np.float32)
    oracle.vortex_ring_smoke(res, 1 / 48, last, *rings[-1][1:4], *rings[-1][4:6], rings[-1][6], 500)
    _check_smoke(N(s2), last)
    # every ring lies in the plane normal to its corner's diagonal, and its
    # self-induced velocity (along sign(w) ex x ey) points to the cube centre
    for sgn, corner, ex, ey, _ in spec:
        centre = np.array([0.5 + 0.16 * s for s in corner])
        inward = _unit(0.5 - centre)
        axis = sgn * _unit(np.cross(_unit(ex), _unit(ey)))
        assert abs(np.dot(_unit(ex), _unit(ey))) < 1e-12
        assert np.dot(axis, inward) > 1 - 1e-12
        v = _velocity_at(N(u), res, centre)
        assert np.dot(_unit(v), inward) > 0.8, (corner, v)


def test_add_fields(pfm):
    from pfm.d3.taichi_utils import add_fields

    g = torch.Generator(device="cpu").manual_seed(3)
    a = torch.randn((4, 9, 7, 5), generator=g).to(DEV)
    b = torch.randn((4, 9, 7, 5), generator=g).to(DEV)
    out = torch.empty_like(a)
    add_fields(a, b, out, 0.75)
    # the product rounded, then the sum: the same two roundings as torch's
    assert torch.equal(out, a + 0.75 * b)
    a0 = a.clone()
    add_fields(a, b, a, 1.0)  # dest aliasing f1, as init_vorts_* use it
    assert torch.equal(a, a0 + b)
