"""Pins of the CPU oracle against closed forms, invariants and independent
library routines (CPU only).  Each test is chosen so that a plausible slip in
the restatement (wrong sign, index, transposed operand, dropped term) fails."""
import math

import numpy as np
import pytest

from helpers import affine_mac_field, face_coords, neumann_laplacian, polyP, rel_l2, smooth_mac_field


# ---- row 1: quadratic B-spline -------------------------------------------
def test_bspline_values(oracle):
    assert oracle.N2(0.0) == 0.75
    assert oracle.N2(0.5) == 0.5 and oracle.N2(-0.5) == 0.5
    assert oracle.N2(1.0) == 0.125 and oracle.N2(-1.0) == 0.125
    assert oracle.N2(1.5) == 0.0 and oracle.N2(2.0) == 0.0
    assert oracle.dN2(0.0) == 0.0
    assert oracle.dN2(1.0) == -0.5 and oracle.dN2(-1.0) == 0.5
    assert oracle.dN2(0.25) == -0.5 and oracle.dN2(-0.25) == 0.5


@pytest.mark.parametrize("x", np.random.default_rng(0).uniform(0, 1, 25))
def test_bspline_partition_moment_derivative(oracle, x):
    js = range(-2, 4)
    w = [oracle.N2(x - j) for j in js]
    dw = [oracle.dN2(x - j) for j in js]
    assert abs(sum(w) - 1.0) < 1e-6
    assert abs(sum((x - j) * wj for j, wj in zip(js, w))) < 1e-6
    assert abs(sum(dw)) < 1e-6
    # derivative of the published quadratic B-spline, by central differences in double
    for j in js:
        t = x - j
        h = 1e-4

        def N(s):
            a = abs(s)
            return 0.75 - a * a if a < 0.5 else (0.5 * (1.5 - a) ** 2 if a < 1.5 else 0.0)

        if min(abs(abs(t) - 0.5), abs(abs(t) - 1.5)) > 2 * h:
            assert abs(oracle.dN2(t) - (N(t + h) - N(t - h)) / (2 * h)) < 1e-4


# ---- rows 3-5, 7-8: interpolation reproduces affine fields -----------------
@pytest.mark.parametrize("res", [(16, 16), (12, 16, 10)])
def test_g2p_affine_reproduction(oracle, res):
    rng = np.random.default_rng(1)
    d = len(res)
    dx = 1.0 / res[1]
    a0 = rng.normal(size=d)
    B = rng.normal(size=(d, d))
    u = affine_mac_field(res, a0, B)
    n = 200
    # interior points: at least 2 cells away from every wall
    lo = 2.5 * dx
    hi = np.array([(r - 2.5) * dx for r in res])
    pos = (lo + rng.uniform(size=(d, n)) * (hi[:, None] - lo)).astype(np.float32)
    imp = np.zeros((d, n), np.float32)
    C = np.zeros((d, d, n), np.float32)
    oracle.g2p(res, dx, pos, u, imp, C)
    want = a0[:, None] + B @ pos.astype(np.float64)
    assert np.max(np.abs(imp - want)) < 2e-5 * (1 + np.abs(want).max())
    for a in range(d):
        for c in range(d):
            assert np.max(np.abs(C[a, c] - B[a, c])) < 2e-3 * (1 + abs(B).max())


# ---- row 14: particle RK4 on a linear field ------------------------------
@pytest.mark.parametrize("res", [(16, 16), (16, 16, 16)])
def test_rk4_particles_linear_field(oracle, res):
    rng = np.random.default_rng(2)
    d = len(res)
    dx = 1.0 / res[1]
    A = rng.normal(size=(d, d)) * 0.5
    u = affine_mac_field(res, np.zeros(d), A)
    n = 64
    pos = (0.45 + 0.1 * rng.uniform(size=(d, n))).astype(np.float32)
    T = rng.normal(size=(d, d, n)).astype(np.float32)
    pos0 = pos.astype(np.float64).copy()
    T0 = T.astype(np.float64).copy()
    dt = 0.05
    oracle.rk4_particles(res, dx, pos, T, u, dt)
    Px = polyP(dt * A)
    PT = polyP(-dt * A.T)
    assert np.max(np.abs(pos - Px @ pos0)) < 1e-5
    for m in range(d):
        assert np.max(np.abs(T[m] - PT @ T0[m])) < 1e-4


def test_rk4_particles_uniform_field(oracle):
    res = (16, 16, 16)
    dx = 1.0 / 16
    vel = np.array([0.3, -0.2, 0.1])
    u = affine_mac_field(res, vel, np.zeros((3, 3)))
    n = 50
    rng = np.random.default_rng(3)
    pos = (0.3 + 0.4 * rng.uniform(size=(3, n))).astype(np.float32)
    T = np.tile(np.eye(3, dtype=np.float32)[:, :, None], (1, 1, n)).copy()
    p0 = pos.copy()
    oracle.rk4_particles(res, dx, pos, T, u, 0.1)
    assert np.max(np.abs(pos - (p0 + 0.1 * vel[:, None]))) < 1e-6
    assert np.max(np.abs(T - np.eye(3)[:, :, None])) < 1e-5


# ---- row 12: backward grid RK4 on a linear field -------------------------
@pytest.mark.parametrize("res", [(16, 16), (16, 16, 16)])
def test_rk4_grid_linear_field(oracle, res):
    rng = np.random.default_rng(4)
    d = len(res)
    dx = 1.0 / res[1]
    A = rng.normal(size=(d, d)) * 0.5
    u = affine_mac_field(res, np.zeros(d), A)
    dt = 0.04
    for axis in range(d):
        psi, T = oracle.reset_identity_grid(res, dx, axis)
        # identity reset: face coordinates and unit columns
        np.testing.assert_array_equal(psi, face_coords(res, axis, dx).astype(np.float32))
        p0 = psi.astype(np.float64).reshape(d, -1).copy()
        oracle.rk4_grid(res, dx, axis, psi, T, u, dt)
        P = polyP(-dt * A)
        p1 = psi.reshape(d, -1)
        X = p0
        inner = np.all([(X[c] > 2.5 * dx) & (X[c] < (res[c] - 2.5) * dx) for c in range(d)], axis=0)
        assert np.max(np.abs(p1[:, inner] - (P @ X)[:, inner])) < 1e-5
        e = np.zeros(d)
        e[axis] = 1
        assert np.max(np.abs(T.reshape(d, -1)[:, inner] - (P @ e)[:, None])) < 1e-4


# ---- row 13: advection with the identity map reproduces affine fields ----
def test_advect_u_identity_affine(oracle):
    res = (12, 16, 10)
    d = 3
    dx = 1.0 / 16
    rng = np.random.default_rng(5)
    u0 = affine_mac_field(res, rng.normal(size=3), rng.normal(size=(3, 3)))
    for a in range(d):
        psi, T = oracle.reset_identity_grid(res, dx, a)
        u1 = np.empty_like(u0[a])
        oracle.advect_u(res, dx, a, u0, u1, T, psi)
        X = face_coords(res, a, dx)
        inner = np.all([(X[c] > 2.5 * dx) & (X[c] < (res[c] - 2.5) * dx) for c in range(d)], axis=0)
        assert np.max(np.abs(u1[inner] - u0[a][inner])) < 1e-5


# ---- row 17: APIC P2G reproduces affine momentum -------------------------
@pytest.mark.parametrize("res", [(16, 16), (12, 12, 12)])
def test_p2g_affine_reproduction(oracle, res):
    rng = np.random.default_rng(6)
    d = len(res)
    dx = 1.0 / res[1]
    a0 = rng.normal(size=d)
    B = rng.normal(size=(d, d))
    ppc_axis = 2
    n = int(np.prod(res)) * ppc_axis ** d
    pos = oracle.seed_particles(res, ppc_axis, dx / ppc_axis, n, seed=11)
    imp = (a0[:, None] + B @ pos.astype(np.float64)).astype(np.float32)
    C = np.repeat(B.astype(np.float32)[:, :, None], n, axis=2).copy()
    # identity flow maps: p2g_prep must return C and imp unchanged
    Tgm = np.tile(np.eye(d, dtype=np.float32)[:, :, None], (1, 1, n)).copy()
    C2, imp2 = oracle.p2g_prep(Tgm, Tgm.copy(), C, imp)
    np.testing.assert_array_equal(C2, C)
    np.testing.assert_array_equal(imp2, imp)
    u = oracle.p2g(res, dx, pos, imp, C)
    for a in range(d):
        X = face_coords(res, a, dx)
        want = a0[a] + np.tensordot(B[a], X, axes=1)
        inner = np.all([(X[c] > 1.6 * dx) & (X[c] < (res[c] - 1.6) * dx) for c in range(d)], axis=0)
        assert np.max(np.abs(u[a][inner] - want[inner])) < 2e-5


def test_p2g_prep_composition(oracle):
    """C_out = A^T C A and imp = A^T B imp0 with A = rows(T_gm), B = cols(T_init)."""
    rng = np.random.default_rng(7)
    d, n = 3, 20
    Tgm = rng.normal(size=(d, d, n)).astype(np.float32)
    Tin = rng.normal(size=(d, d, n)).astype(np.float32)
    C = rng.normal(size=(d, d, n)).astype(np.float32)
    m0 = rng.normal(size=(d, n)).astype(np.float32)
    Cout, imp = oracle.p2g_prep(Tgm, Tin, C, m0)
    for p in range(n):
        A = Tgm[:, :, p].astype(np.float64)          # row r = T_r
        Bm = Tin[:, :, p].astype(np.float64).T       # column m = T_m_init
        want_C = A.T @ C[:, :, p].astype(np.float64) @ A
        want_imp = A.T @ Bm @ m0[:, p].astype(np.float64)
        assert np.max(np.abs(Cout[:, :, p] - want_C)) < 1e-4 * (1 + np.abs(want_C).max())
        assert np.max(np.abs(imp[:, p] - want_imp)) < 1e-4 * (1 + np.abs(want_imp).max())
    # update_T: column m of the result = T_gm-matrix @ column m of T_in
    Tout = oracle.update_T(Tgm, Tin)
    for p in range(n):
        G = Tgm[:, :, p].astype(np.float64).T        # columns T_m_gm
        Bm = Tin[:, :, p].astype(np.float64).T
        assert np.max(np.abs(Tout[:, :, p].T - G @ Bm)) < 1e-4


def test_update_imp_columns_are_images_of_the_basis(oracle):
    """2D update_particles_imp (2D/run_2D.py:517-524): T = cols(T_x, T_y), so
    T_x and T_y are the images of e_x and e_y and imp = a T_x + b T_y for
    init_imp = (a, b).  A 90-degree rotation and a shear (both non-symmetric,
    small integers so every product is exact) catch a transposed T."""
    # particle 0: rotation by +90 degrees (e_x -> e_y, e_y -> -e_x)
    # particle 1: shear x += 3 y (e_x -> e_x, e_y -> (3, 1))
    # particle 2: rotation times shear, T_x = (0, 1), T_y = (-1, 3)
    T = np.zeros((2, 2, 3), np.float32)
    T[0, :, 0], T[1, :, 0] = (0, 1), (-1, 0)
    T[0, :, 1], T[1, :, 1] = (1, 0), (3, 1)
    T[0, :, 2], T[1, :, 2] = (0, 1), (-1, 3)
    m = np.array([[1, 2, 2], [0, 5, -1]], np.float32)  # init_imp columns per particle
    imp = oracle.update_imp(T, m)
    # rotated e_x is e_y; shear of (2, 5) is (17, 5); T @ (2, -1) = 2 (0,1) - (-1, 3) = (1, -1)
    want = np.array([[0, 17, 1], [1, 5, -1]], np.float32)
    assert np.array_equal(imp, want)
    # the transposed product differs on every particle, so the check has teeth
    wrong = np.einsum("mcp,cp->mp", T, m)
    assert all(not np.array_equal(wrong[:, p], want[:, p]) for p in range(3))


def test_cfl_dt_frame_clip_hand_computed(oracle):
    """dt logic of 3D/run_3D.py:725-743: dt = CFL dx / max_speed (max_speed a
    float32 read into a double), and the step that would reach visualize_dt is
    clipped to end the frame exactly, resetting sub_t."""
    dx = 1.0 / 128
    # max speed 0.5: dt = 0.5 / 128 / 0.5 = 1/128 exactly; 12 steps reach 12/128 =
    # 0.09375 and the 13th (0.09375 + 1/128 >= 0.1) is clipped to 0.1 - 0.09375
    sub_t, seen = 0.0, []
    for _ in range(14):
        dt, sub_t, end = oracle.cfl_dt(0.5, dx, 0.5, sub_t, 0.1)
        seen.append((dt, sub_t, end))
    for k in range(12):
        assert seen[k] == (1 / 128, (k + 1) / 128, False)
    assert seen[12][2] and seen[12][1] == 0.0 and abs(seen[12][0] - 0.00625) < 1e-15
    assert seen[13] == (1 / 128, 1 / 128, False)
    # ">=": with visualize_dt = 4/128 the 4th step lands exactly on it and ends the frame
    sub_t = 3 / 128
    dt, sub_t, end = oracle.cfl_dt(0.5, dx, 0.5, sub_t, 4 / 128)
    assert end and dt == 1 / 128 and sub_t == 0.0
    # the max speed is the float32 value of the field, not a double
    dt, _, _ = oracle.cfl_dt(np.float32(0.3), dx, 0.5, 0.0, 1.0)
    assert dt == 0.5 * dx / 0.30000001192092896 and dt != 0.5 * dx / 0.3


# ---- row 15: binning = stable argsort by truncated cell ------------------
@pytest.mark.parametrize("res", [(8, 8), (6, 8, 5)])
def test_binning_stable_argsort(oracle, res):
    rng = np.random.default_rng(8)
    d = len(res)
    dx = 1.0 / res[1]
    n = 3000
    ext = np.array(res) * dx
    pos = (rng.uniform(-0.3 * dx, 1.0, size=(d, n)) * np.r_[ext][:, None]).astype(np.float32)
    cell, counts, start, perm = oracle.bin_particles(res, dx, pos)
    inv = np.float32(1.0 / dx)
    ci = [np.clip((pos[c] * inv).astype(np.int32), 0, res[c] - 1) for c in range(d)]
    flat = np.ravel_multi_index(ci, res).astype(np.int32)
    np.testing.assert_array_equal(cell, flat)
    np.testing.assert_array_equal(perm, np.argsort(flat, kind="stable"))
    np.testing.assert_array_equal(counts, np.bincount(flat, minlength=int(np.prod(res))))
    np.testing.assert_array_equal(start[1:], np.cumsum(counts))
    ids = oracle.cell_lists(res, counts, perm, start, 12)
    assert ids.shape == tuple(res) + (12,)
    c0 = flat[0]
    lst = ids.reshape(-1, 12)[c0]
    assert lst[0] == 0 or perm[start[c0]] == lst[0]


# ---- row 6: seeding ------------------------------------------------------
@pytest.mark.parametrize("res,ppa", [((8, 6), 2), ((6, 4, 5), 2), ((5, 4), 3)])
def test_seeding_lattice(oracle, res, ppa):
    d = len(res)
    dx = 1.0 / res[1]
    dist = np.float32(dx / ppa)
    n = int(np.prod(res)) * ppa ** d
    pos = oracle.seed_particles(res, ppa, float(dist), n)
    lat = [np.arange(ppa * r) for r in res]
    grid = np.meshgrid(*lat, indexing="ij")
    # reference order: x fastest (3D/init_conditions.py:243-252)
    want = np.stack([((g.transpose(*reversed(range(d))).ravel()).astype(np.float32) + np.float32(0.5)) * dist
                     for g in grid])
    np.testing.assert_array_equal(pos, want)
    # cell-major order is a permutation of the same lattice, cell by cell
    pos1 = oracle.seed_particles(res, ppa, float(dist), n, order=1)
    key = lambda p: np.lexsort(p[::-1])
    np.testing.assert_array_equal(pos1[:, key(pos1)], pos[:, key(pos)])
    cells = np.ravel_multi_index([(pos1[c] / np.float32(dx)).astype(int) for c in range(d)], res)
    assert np.all(np.diff(cells) >= 0)
    # jitter: same sub-cell, deterministic, differs from the lattice
    pj = oracle.seed_particles(res, ppa, float(dist), n, seed=5)
    pj2 = oracle.seed_particles(res, ppa, float(dist), n, seed=5)
    np.testing.assert_array_equal(pj, pj2)
    assert np.all(np.floor(pj / dist) == np.floor(pos / dist))
    assert np.abs(pj - pos).max() > 0


# ---- row 30: reinit schedule ---------------------------------------------
def test_reinit_schedule(oracle):
    for re, rg, dim in [(20, 1, 3), (12, 8, 2), (12, 4, 3), (20, 8, 2)]:
        sch = oracle.reinit_schedule(100, re, rg, True, dim)
        for i, (lr, rs, sr) in enumerate(sch):
            assert lr == (i % re == 0)
            assert rs == lr
            assert sr == ((i % re) % rg == 0)
    # 3D without reseeding keeps ik = 0 (3D/run_3D.py:751-755)
    sch = oracle.reinit_schedule(50, 12, 5, False, 3)
    for i, (lr, rs, sr) in enumerate(sch):
        assert rs == (i == 0)
        assert sr == (i % 5 == 0)


# ---- row 31: centre/face averaging and the vortex ICs --------------------
def test_split_and_get_central(oracle):
    rng = np.random.default_rng(9)
    for res in [(6, 5), (4, 6, 5)]:
        d = len(res)
        vc = rng.normal(size=(d,) + res).astype(np.float32)
        u = oracle.split_central(res, vc)
        for a in range(d):
            pad = [(0, 0)] * d
            pad[a] = (1, 1)
            p = np.pad(vc[a], pad, mode="edge")
            sl_r = [slice(None)] * d
            sl_l = [slice(None)] * d
            sl_r[a] = slice(1, None)
            sl_l[a] = slice(0, -1)
            want = np.float32(0.5) * (p[tuple(sl_r)] + p[tuple(sl_l)])
            np.testing.assert_array_equal(u[a], want)
        back = oracle.get_central(res, u)
        for a in range(d):
            sl_r = [slice(None)] * d
            sl_l = [slice(None)] * d
            sl_r[a] = slice(1, None)
            sl_l[a] = slice(0, -1)
            np.testing.assert_array_equal(back[a], np.float32(0.5) * (u[a][tuple(sl_r)] + u[a][tuple(sl_l)]))


def test_point_vortices_formula(oracle):
    res = (32, 32)
    dx = 1.0 / 32
    cen = [(0.25, 0.45), (0.35, 0.55)]
    w = [1.0, -1.0]
    vc = oracle.point_vortices(res, dx, cen, w)
    X = (np.arange(32) + 0.5) * dx
    x, y = np.meshgrid(X, X, indexing="ij")
    ux = np.zeros_like(x)
    uy = np.zeros_like(x)
    for (cx, cy), wq in zip(cen, w):
        px, py = x - cx, y - cy
        r = np.sqrt(px * px + py * py) + 1e-6
        ang = (-0.01 / r * (1 - np.exp(-(r * r) / 0.02 ** 2))) / r * wq
        ux += ang * -py
        uy += ang * px
    assert rel_l2(vc[0], ux) < 1e-5 and rel_l2(vc[1], uy) < 1e-5


def test_vortex_ring_biot_savart(oracle):
    res = (8, 8, 8)
    dx = 1.0 / 8
    vc = np.zeros((3,) + res, np.float32)
    c = [0.5, 0.5, 0.4]
    ex, ey = [1.0, 0, 0], [0, 1.0, 0]
    ns = 64
    oracle.vortex_ring(res, dx, vc, 1.0, 0.2, 0.05, c, ex, ey, ns)
    g = (np.arange(8) + 0.5) * dx
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    want = np.zeros_like(P)
    cl = 2 * np.pi * 0.2 / ns
    for l in range(ns):
        th = l / ns * 2 * np.pi
        ps = 0.2 * (np.cos(th) * np.array(ex) + np.sin(th) * np.array(ey)) + c
        pd = P - ps
        r = np.linalg.norm(pd, axis=-1)[..., None]
        wv = -np.sin(th) * np.array(ex) + np.cos(th) * np.array(ey)
        want += cl * (-1 / (4 * np.pi * r ** 3) * (1 - np.exp(-(r / 0.05) ** 3))) * np.cross(pd, wv)
    assert rel_l2(np.moveaxis(vc, 0, -1), want) < 1e-4


def test_kinetic_energy_and_max_speed(oracle):
    res = (6, 5, 4)
    u = smooth_mac_field(res, seed=3)
    E = oracle.kinetic_energy(res, u)
    want = sum(float(np.sum(x.astype(np.float64) ** 2)) / x.size for x in u)
    assert abs(E - want) < 1e-12 * max(1, want)
    ms = oracle.max_speed(res, u)
    cx = 0.5 * (u[0][1:] + u[0][:-1])
    cy = 0.5 * (u[1][:, 1:] + u[1][:, :-1])
    cz = 0.5 * (u[2][:, :, 1:] + u[2][:, :, :-1])
    assert ms == max(np.float32(1e-3), np.sqrt(cx * cx + cy * cy + cz * cz).max())


# ---- rows 18-29: MGPCG ---------------------------------------------------
@pytest.mark.parametrize("N", [(16, 16), (8, 8, 8), (16, 8, 8)])
def test_mg_operator_matches_laplacian(oracle, N):
    mg = oracle.MG(N, base_level=2)
    A = neumann_laplacian(N)
    p = np.random.default_rng(10).normal(size=N).astype(np.float32)
    Ap = mg.apply_A(0, p)
    want = (A @ p.astype(np.float64).ravel()).reshape(N)
    assert np.max(np.abs(Ap - want)) < 1e-5


def test_mg_smooth_restrict_prolong_invariants(oracle):
    N = (8, 8, 8)
    mg = oracle.MG(N, base_level=2)
    rng = np.random.default_rng(11)
    A = neumann_laplacian(N)
    r = rng.normal(size=N).astype(np.float32)
    mg.field("r", 0)[...] = r
    mg.field("z", 0)[...] = rng.normal(size=N).astype(np.float32)
    I = np.indices(N).sum(0) & 1
    for phase in (0, 1):
        mg.smooth(0, phase)
        z = mg.field("z", 0).astype(np.float64)
        res = r - (A @ z.ravel()).reshape(N)
        assert np.max(np.abs(res[I == phase])) < 1e-5
    z = mg.field("z", 0).astype(np.float64)
    resid = r - (A @ z.ravel()).reshape(N)
    mg.restrict(0)
    rc = mg.field("r", 1)
    want = resid.reshape(4, 2, 4, 2, 4, 2).sum(axis=(1, 3, 5)) / 2.0
    assert np.max(np.abs(rc - want)) < 1e-5
    zc = rng.normal(size=(4, 4, 4)).astype(np.float32)
    mg.field("z", 1)[...] = zc
    z0 = mg.field("z", 0).copy()
    mg.prolongate(0)
    up = np.repeat(np.repeat(np.repeat(zc, 2, 0), 2, 1), 2, 2)
    np.testing.assert_array_equal(mg.field("z", 0), z0 + up)


def test_mg_vcycle_symmetric_positive(oracle):
    N = (16, 16, 16)
    mg = oracle.MG(N, base_level=3)
    rng = np.random.default_rng(12)
    a = rng.normal(size=N)
    b = rng.normal(size=N)
    a -= a.mean()
    b -= b.mean()
    Ma = mg.vcycle(a.astype(np.float32)).astype(np.float64)
    Mb = mg.vcycle(b.astype(np.float32)).astype(np.float64)
    lhs, rhs = np.sum(Ma * b), np.sum(a * Mb)
    assert abs(lhs - rhs) < 1e-3 * np.linalg.norm(Ma) * np.linalg.norm(b)
    assert np.sum(Ma * a) > 0


@pytest.mark.parametrize("N,fp64", [((32, 32), False), ((16, 16, 16), False), ((16, 16, 16), True)])
def test_mg_solve_matches_scipy(oracle, N, fp64):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    rng = np.random.default_rng(13)
    b = rng.normal(size=N)
    b -= b.mean()
    A = neumann_laplacian(N)
    n = A.shape[0]
    # zero-mean solution of the singular Neumann system
    Areg = (A + sp.csr_matrix(np.ones((n, n)) / n)).tocsc() if n <= 4096 else None
    x_ref = spl.spsolve(Areg, b.ravel()).reshape(N)
    mg = oracle.MG(N, base_level=3, fp64=fp64)
    mg.init(-b, -1.0)
    # the reference V-cycle is only approximately symmetric (50 bottom sweeps),
    # so even in FP64 PCG stalls near a 1e-9 relative residual
    it, rtr, conv = mg.solve(max_iters=200, tol=1e-8 if fp64 else 1e-6, tol_mode=1)
    assert conv
    x = mg.get_result().astype(np.float64)
    x -= x.mean()
    assert rel_l2(x, x_ref) < (1e-7 if fp64 else 1e-5)
    assert it < 40


def test_poisson_projection_divergence_free(oracle):
    res = (16, 16, 16)
    u = smooth_mac_field(res, seed=14)
    mg = oracle.MG(res, base_level=3)
    it, rtr, conv = mg.poisson(u, max_iters=400, tol=1e-6, tol_mode=1)
    assert conv
    # normal faces on the walls are zero and the discrete divergence vanishes
    assert np.all(u[0][0] == 0) and np.all(u[0][-1] == 0)
    div = (u[0][1:] - u[0][:-1]) + (u[1][:, 1:] - u[1][:, :-1]) + (u[2][:, :, 1:] - u[2][:, :, :-1])
    u0 = smooth_mac_field(res, seed=14)
    div0 = (u0[0][1:] - u0[0][:-1]) + (u0[1][:, 1:] - u0[1][:, :-1]) + (u0[2][:, :, 1:] - u0[2][:, :, :-1])
    assert np.abs(div).max() < 1e-4 * np.abs(div0).max()


# ---- diagnostics: curl ------------------------------------------------------
def test_curl_linear_field_and_walls(oracle):
    """Interior: the exact curl of a linear field.  Walls: sample() clamps the
    neighbour, so the difference is one-sided over 2 dx (numpy edge padding)."""
    dx = 0.125
    res = (6, 7, 8)
    x, y, z = [(np.arange(n) + 0.5) * dx for n in res]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    vc = np.stack([2 * Y + 3 * Z, -X + 0.5 * Z, 4 * X - Y]).astype(np.float32)
    c = oracle.curl(res, dx, vc)
    # curl (2y+3z, -x+z/2, 4x-y) = (-1 - 1/2, 3 - 4, -1 - 2)
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(c[inner][0], -1.5, atol=1e-5)
    np.testing.assert_allclose(c[inner][1], -1.0, atol=1e-5)
    np.testing.assert_allclose(c[inner][2], -3.0, atol=1e-5)
    # every cell against an edge-padded central difference in float64
    p = np.pad(vc.astype(np.float64), [(0, 0), (1, 1), (1, 1), (1, 1)], mode="edge")

    def dd(comp, axis):
        hi = [slice(1, -1)] * 3
        lo = [slice(1, -1)] * 3
        hi[axis] = slice(2, None)
        lo[axis] = slice(0, -2)
        return (p[comp][tuple(hi)] - p[comp][tuple(lo)]) / (2 * dx)

    want = np.stack([dd(2, 1) - dd(1, 2), dd(0, 2) - dd(2, 0), dd(1, 0) - dd(0, 1)])
    np.testing.assert_allclose(c, want, atol=1e-4)
    # 2D: scalar vorticity of (-y, x) is 2 inside
    res2 = (9, 6)
    x2, y2 = [(np.arange(n) + 0.5) * dx for n in res2]
    X2, Y2 = np.meshgrid(x2, y2, indexing="ij")
    v2 = np.stack([-Y2, X2]).astype(np.float32)
    c2 = oracle.curl(res2, dx, v2)
    np.testing.assert_allclose(c2[1:-1, 1:-1], 2.0, atol=1e-5)
    np.testing.assert_allclose(c2[0, 3], 1.5, atol=1e-5)  # one-sided in x at the wall


# ---- MGPCG with a solid obstacle (extension) -----------------------------
def _solid_case(N):
    from helpers import box_solid

    return box_solid(N, 0.3, 0.55)


@pytest.mark.parametrize("N", [(16, 16, 16), (24, 16)])
def test_mg_solid_operator_and_coarsening(oracle, N):
    from helpers import fluid_laplacian

    solid = _solid_case(N)
    assert solid.any() and not solid.all()
    mg = oracle.MG(N, base_level=2)
    mg.set_solid(solid)
    p = np.random.default_rng(40).normal(size=N).astype(np.float32)
    p[solid] = 0
    Ap = mg.apply_A(0, p)
    want = (fluid_laplacian(solid) @ p.astype(np.float64).ravel()).reshape(N)
    assert np.max(np.abs(Ap - want)) < 1e-5
    assert np.all(Ap[solid] == 0)
    # one colour of RB-GS zeroes the residual of that colour on fluid cells, solid z stays 0
    r = np.random.default_rng(41).normal(size=N).astype(np.float32)
    r[solid] = 0
    mg.field("r", 0)[...] = r
    mg.field("z", 0)[...] = 0
    mg.smooth(0, 0)
    z = mg.field("z", 0).astype(np.float64)
    res = r - (fluid_laplacian(solid) @ z.ravel()).reshape(N)
    col = np.indices(N).sum(0) & 1
    assert np.max(np.abs(res[(col == 0) & ~solid])) < 1e-5
    assert np.all(z[solid] == 0)
    # restriction: coarse value = sum of the fluid children's residuals / (d-1)
    mg.restrict(0)
    rc = mg.field("r", 1)
    d = len(N)
    children = res.reshape([n // 2 if i < d else n for i, n in enumerate(N)][:0] + sum(([n // 2, 2] for n in N), []))
    want_c = children.sum(axis=tuple(range(1, 2 * d, 2))) / (d - 1)
    assert np.max(np.abs(rc - want_c)) < 1e-5


def test_mg_solid_solve_vs_scipy(oracle):
    import scipy.sparse.linalg as spla
    from helpers import fluid_laplacian

    N = (16, 16, 16)
    solid = _solid_case(N)
    A = fluid_laplacian(solid)
    fl = ~solid.ravel()
    rng = np.random.default_rng(42)
    rhs = rng.normal(size=N)
    rhs[solid] = 0
    rhs[~solid] -= rhs[~solid].mean()
    # direct solve of the fluid block with the zero-mean constraint (bordered system)
    import scipy.sparse as sp

    Af = A[fl][:, fl]
    nf = Af.shape[0]
    K = sp.bmat([[Af, np.ones((nf, 1))], [np.ones((1, nf)), None]]).tocsc()
    sol = spla.spsolve(K, np.concatenate([-rhs.ravel()[fl], [0.0]]))[:nf]
    for fp64, tol in ((False, 1e-5), (True, 1e-7)):
        mg = oracle.MG(N, base_level=3, fp64=fp64)
        mg.set_solid(solid)
        mg.init(rhs.astype(np.float64 if fp64 else np.float32), -1.0)
        it, rtr, conv = mg.solve(max_iters=200, tol=1e-9 if fp64 else 1e-6, tol_mode=1)
        x = mg.get_result().astype(np.float64).ravel()
        assert np.all(x[~fl] == 0)
        xf = x[fl] - x[fl].mean()
        assert rel_l2(xf, sol - sol.mean()) < tol, (fp64, rel_l2(xf, sol - sol.mean()))


def test_mg_solid_projection(oracle):
    from helpers import box_solid

    res = (16, 16, 16)
    solid = box_solid(res, 0.3, 0.55)
    u = smooth_mac_field(res, seed=43)
    mg = oracle.MG(res, base_level=3)
    mg.set_solid(solid)
    it, rtr, conv = mg.poisson(u, max_iters=400, tol=1e-7, tol_mode=1)
    assert conv
    # faces touching the obstacle are zero; the divergence vanishes on fluid cells
    sp = np.pad(solid, 1)
    for a in range(3):
        lo = [slice(1, -1)] * 3
        hi = [slice(1, -1)] * 3
        lo[a] = slice(0, -1)
        hi[a] = slice(1, None)
        touch = sp[tuple(lo)] | sp[tuple(hi)]
        assert np.all(u[a][touch] == 0)
    div = (u[0][1:] - u[0][:-1]) + (u[1][:, 1:] - u[1][:, :-1]) + (u[2][:, :, 1:] - u[2][:, :, :-1])
    u0 = smooth_mac_field(res, seed=43)
    div0 = (u0[0][1:] - u0[0][:-1]) + (u0[1][:, 1:] - u0[1][:, :-1]) + (u0[2][:, :, 1:] - u0[2][:, :, :-1])
    assert np.abs(div[~solid]).max() < 1e-4 * np.abs(div0).max()


# ---- smoke transport on the grid (advect_smoke / clamp_smoke) -------------
def _identity_psi(res, dx, shift=0.0):
    return [(face_coords(res, a, dx) + shift).astype(np.float32) for a in range(len(res))]


def _cell_centres(res, dx):
    g = np.meshgrid(*[(np.arange(n) + 0.5) * dx for n in res], indexing="ij")
    return np.stack(g)


@pytest.mark.parametrize("res", [(9, 8, 7), (12, 10)])
def test_advect_smoke_identity_and_linear(oracle, res):
    """interp_1 at the mean of the face maps: the identity map returns the
    field (the 3D clamp s >= 1 makes the first cell of each axis read its
    neighbour), and a uniform shift of a linear field is exact inside."""
    d = len(res)
    dx = 1.0 / res[1]
    rng = np.random.default_rng(7)
    s0 = rng.uniform(0, 1, size=(3,) + res).astype(np.float32)
    s1 = oracle.advect_smoke(res, dx, s0, _identity_psi(res, dx))
    lo = 1 if d == 3 else 0
    inner = tuple(slice(lo, None) for _ in range(d))
    np.testing.assert_allclose(s1[(slice(None),) + inner], s0[(slice(None),) + inner], rtol=0, atol=2e-6)
    if d == 3:
        np.testing.assert_allclose(s1[:, 0, 2:, 2:], s0[:, 1, 2:, 2:], rtol=0, atol=2e-6)
    X = _cell_centres(res, dx)
    coef = np.array([0.3, -0.7, 0.45][:d])
    lin = (0.2 + np.tensordot(coef, X, axes=1)).astype(np.float32)[None]
    delta = 0.3 * dx
    s1 = oracle.advect_smoke(res, dx, lin, _identity_psi(res, dx, delta))
    want = 0.2 + np.tensordot(coef, X + delta, axes=1)
    core = tuple(slice(lo, n - 2) for n in res)
    np.testing.assert_allclose(s1[0][core], want[core], rtol=0, atol=5e-6)


def test_clamp_smoke_is_the_range_of_the_eight_samples(oracle):
    res = (8, 9, 10)
    dx = 1.0 / res[1]
    rng = np.random.default_rng(8)
    s0 = rng.uniform(0, 1, size=(2,) + res).astype(np.float32)
    psi = [(p + rng.uniform(-0.6, 0.6, size=p.shape[1:]).astype(np.float32) * dx).astype(np.float32)
           for p in _identity_psi(res, dx)]
    s1 = rng.uniform(-1, 2, size=s0.shape).astype(np.float32)
    out = oracle.advect_smoke(res, dx, s0, psi, smoke1=s1.copy(), mode=2)
    # numpy: the same psi_c and floor, then min / max over the 2^3 samples
    f32 = np.float32
    sx = psi[0][:, :-1] + psi[0][:, 1:]
    sy = sx + psi[1][:, :, :-1]
    sy = sy + psi[1][:, :, 1:]
    sz = sy + psi[2][:, :, :, :-1]
    sz = sz + psi[2][:, :, :, 1:]
    pc = f32(1.0 / 6.0) * sz
    u = pc * f32(1.0 / dx) - f32(0.5)
    hi = np.array([f32(n - 1 - 1e-6) for n in res], np.float32).reshape(3, 1, 1, 1)
    s = np.maximum(f32(1.0), np.minimum(u, hi))
    fl = np.floor(s).astype(np.int64)
    mn = np.full(s0.shape, np.inf, np.float32)
    mx = np.full(s0.shape, -np.inf, np.float32)
    for o in np.ndindex(2, 2, 2):
        idx = [np.clip(fl[a] + o[a], 0, res[a] - 1) for a in range(3)]
        v = s0[:, idx[0], idx[1], idx[2]]
        mn, mx = np.minimum(mn, v), np.maximum(mx, v)
    np.testing.assert_array_equal(out, np.minimum(np.maximum(s1, mn), mx))
    # advect then clamp (mode 3) equals the two passes
    both = oracle.advect_smoke(res, dx, s0, psi, mode=3)
    two = oracle.advect_smoke(res, dx, s0, psi)
    two = oracle.advect_smoke(res, dx, s0, psi, smoke1=two, mode=2)
    np.testing.assert_array_equal(both, two)


def test_vortex_ring_smoke_density(oracle):
    res = (10, 10, 10)
    dx = 1.0 / 10
    c, ex, ey, color = [0.5, 0.5, 0.4], [1.0, 0, 0], [0, 1.0, 0], [0.2, 0.5, 0.9]
    ns, rad, delta = 64, 0.2, 0.05
    sm = np.zeros((4,) + res, np.float32)
    oracle.vortex_ring_smoke(res, dx, sm, rad, delta, c, ex, ey, color, ns)
    P = np.moveaxis(_cell_centres(res, dx), 0, -1)
    dens = np.zeros(res)
    cl = 2 * np.pi * rad / ns
    for l in range(ns):
        th = l / ns * 2 * np.pi
        ps = rad * (np.cos(th) * np.array(ex) + np.sin(th) * np.array(ey)) + c
        r = np.linalg.norm(P - ps, axis=-1)
        dens += cl * np.exp(-(r / delta) ** 3)
    on = dens > 0.002
    near = np.abs(dens - 0.002) < 1e-6
    mismatch = (sm[3] == 1.0) != on
    assert np.all(~mismatch | near)
    assert 0 < on.sum() < on.size
    for k in range(3):
        np.testing.assert_array_equal(sm[k][sm[3] == 1.0], np.float32(color[k]))
        assert np.all(sm[k][sm[3] == 0.0] == 0)


# ---- particle smoke (init_particles_smoke / P2G_smoke / clamp_smoke_particle)
def _smoke_affine(res, dx):
    a = np.array([0.4, 0.5, 0.45, 0.55])
    b = np.array([[0.3, -0.2, 0.1], [-0.1, 0.25, 0.2], [0.05, 0.1, -0.3], [0.2, 0.15, 0.1]])
    X = _cell_centres(res, dx)
    field = (a[:, None, None, None] + np.tensordot(b, X, axes=1)).astype(np.float32)
    return a, b, field


def test_g2p_smoke_reproduces_affine_fields(oracle):
    """interp_grad_2_smoke: the quadratic B-spline value and gradient of an
    affine field are exact away from the clamp."""
    res = (12, 12, 12)
    dx = 1.0 / 12
    a, b, field = _smoke_affine(res, dx)
    rng = np.random.default_rng(9)
    pos = rng.uniform(2.0 * dx, 10.0 * dx, size=(3, 500)).astype(np.float32)
    ps, pg = oracle.g2p_smoke(res, dx, pos, field)
    want = a[:, None] + b @ pos.astype(np.float64)
    np.testing.assert_allclose(ps, want, rtol=0, atol=2e-6)
    for r in range(3):
        for c in range(4):
            np.testing.assert_allclose(pg[4 * r + c], b[c, r], rtol=0, atol=2e-4)


@pytest.mark.parametrize("apic", [True, False])
def test_p2g_smoke_affine_reproduction(oracle, apic):
    """P2G_smoke with APIC: particles carrying an affine smoke and its gradient
    (T = I) splat it back exactly on cells whose 4^3 neighbourhood is covered;
    without APIC a constant field stays constant."""
    res = (10, 10, 10)
    dx = 1.0 / 10
    n = int(np.prod(res)) * 8
    pos = oracle.seed_particles(res, 2, dx / 2, n, seed=4)
    a, b, _ = _smoke_affine(res, dx)
    if apic:
        ps = (a[:, None] + b @ pos.astype(np.float64)).astype(np.float32)
    else:
        ps = np.repeat(a[:, None], n, axis=1).astype(np.float32)
    grad = np.zeros((12, n), np.float32)
    for r in range(3):
        for c in range(4):
            grad[4 * r + c] = b[c, r]
    eye = np.zeros((3, 3, n), np.float32)
    for c in range(3):
        eye[c, c] = 1.0
    smoke, w = oracle.p2g_smoke(res, dx, pos, ps, grad, eye, eye, apic=apic)
    X = _cell_centres(res, dx)
    want = a[:, None, None, None] + np.tensordot(b, X, axes=1) * (1.0 if apic else 0.0)
    core = (slice(None),) + tuple(slice(2, n_ - 2) for n_ in res)
    np.testing.assert_allclose(smoke[core], want[core], rtol=0, atol=3e-6)
    assert np.all(w > 0)


def test_clamp_smoke_particle_brute_force(oracle):
    res = (5, 6, 4)
    dx = 1.0 / 6
    rng = np.random.default_rng(10)
    n = 300
    pos = rng.uniform(0, 1, size=(3, n)).astype(np.float32) * np.array(res, np.float32)[:, None] * np.float32(dx)
    ipos = np.clip(pos + rng.normal(0, 0.08, size=pos.shape).astype(np.float32), 0, None).astype(np.float32)
    init_smoke = rng.uniform(0, 1, size=(4,) + res).astype(np.float32)
    smoke0 = rng.uniform(-0.5, 1.5, size=(4,) + res).astype(np.float32)
    _cell, counts, start, perm = oracle.bin_particles(res, dx, pos)
    out = oracle.clamp_smoke_particle(res, dx, start, perm, 3, ipos, init_smoke, smoke0.copy())
    want = smoke0.copy()
    inv = np.float32(1.0 / dx)
    for q in range(int(np.prod(res))):
        I = np.unravel_index(q, res)
        ids = perm[start[q]:start[q + 1]][:3]
        vals = []
        for p in ids:
            base = (ipos[:, p] * inv).astype(np.int64)
            for o in np.ndindex(3, 3, 3):
                nb = base + np.array(o) - 1
                if np.all(nb >= 0) and np.all(nb < np.array(res)):
                    vals.append(init_smoke[(slice(None),) + tuple(nb)])
        if vals:
            v = np.stack(vals)
            want[(slice(None),) + I] = np.minimum(np.maximum(smoke0[(slice(None),) + I], v.min(0)), v.max(0))
    np.testing.assert_array_equal(out, want)
    assert np.any(counts > 3)  # the cap is exercised
