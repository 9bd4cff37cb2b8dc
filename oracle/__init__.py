"""CPU oracle for the particle-flow-map step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker the HIP library is
compared against, never part of the product path.

The arithmetic lives in ``pfm_oracle.c`` (an fp32 line-by-line restatement of
the reference, each function citing the reference file:line it follows); this
module loads it with ctypes and wraps it for numpy arrays.  It also holds the
host-side schedule restatement (``reinit_schedule``) and the full-step driver
(``OracleSim``) following ``3D/run_3D.py:716-831`` and ``2D/run_2D.py:772-909``.

Array conventions match the product: grid fields are numpy arrays shaped like
the reference's fields (face x of a 3D grid: ``(rx+1, ry, rz)``), particle
vector fields are SoA ``(d, n)`` float32, matrices per particle are ``(d, d, n)``
with ``[m, c, p]`` = component c of column vector m (``T_x, T_y, T_z``) or of
row m (``C_x, C_y, C_z``).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpfm_oracle.so")
_lib = None

F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        fl = ctypes.c_float
        ci = ctypes.c_int
        L.oracle_N2.restype = fl
        L.oracle_N2.argtypes = [fl]
        L.oracle_dN2.restype = fl
        L.oracle_dN2.argtypes = [fl]
        L.oracle_jitter_u.restype = fl
        L.oracle_jitter_u.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_rk4_particles.argtypes = [ci, I32P, fl, i64, vp, vp, vp, vp, vp, vp, vp, fl, ci]
        L.oracle_rk4_grid.argtypes = [ci, I32P, fl, ci, vp, vp, vp, vp, vp, fl]
        L.oracle_advect_u.argtypes = [ci, I32P, fl, ci, vp, vp, vp, vp, vp, vp]
        L.oracle_g2p.argtypes = [ci, I32P, fl, i64, vp, vp, vp, vp, vp, vp]
        L.oracle_seed_particles.argtypes = [ci, I32P, ci, fl, i64, vp, i64, ci]
        L.oracle_reset_identity_grid.argtypes = [ci, I32P, fl, ci, vp, vp]
        L.oracle_update_T.argtypes = [ci, i64, vp, vp, vp]
        L.oracle_p2g_prep.argtypes = [ci, i64, vp, vp, vp, vp, vp, vp]
        L.oracle_update_imp.argtypes = [ci, i64, vp, vp, vp]
        L.oracle_bin_particles.argtypes = [ci, I32P, fl, i64, vp, vp, vp, vp, vp]
        L.oracle_p2g.argtypes = [ci, I32P, fl, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_max_speed.restype = fl
        L.oracle_max_speed.argtypes = [ci, I32P, vp, vp, vp]
        L.oracle_split_central.argtypes = [ci, I32P, vp, vp, vp, vp]
        L.oracle_get_central.argtypes = [ci, I32P, vp, vp, vp, vp]
        L.oracle_curl.argtypes = [ci, I32P, fl, vp, vp]
        L.oracle_point_vortices.argtypes = [I32P, fl, ci, F32P, F32P, F32P, fl, fl, vp]
        L.oracle_vortex_ring.argtypes = [I32P, fl, fl, fl, fl, F32P, F32P, F32P, ci, vp]
        L.oracle_advect_smoke.argtypes = [ci, I32P, fl, ci, vp, vp, vp, vp, vp, ci]
        L.oracle_vortex_ring_smoke.argtypes = [I32P, fl, fl, fl, F32P, F32P, F32P, ci, F32P, vp]
        L.oracle_g2p_smoke.argtypes = [I32P, fl, i64, vp, vp, vp, vp]
        L.oracle_p2g_smoke.argtypes = [I32P, fl, i64, vp, vp, vp, vp, vp, ci, vp, vp]
        L.oracle_clamp_smoke_particle.argtypes = [I32P, fl, i64, vp, vp, ci, vp, vp, vp]
        L.oracle_kinetic_energy.restype = ctypes.c_double
        L.oracle_kinetic_energy.argtypes = [ci, I32P, vp, vp, vp]
        for pre in ("oracle_mg_", "oracle_mg64_"):
            getattr(L, pre + "create").restype = vp
            getattr(L, pre + "create").argtypes = [ci, I32P, ci, I32P, ci]
            getattr(L, pre + "destroy").argtypes = [vp]
            getattr(L, pre + "n_levels").restype = ci
            getattr(L, pre + "n_levels").argtypes = [vp]
            getattr(L, pre + "field").restype = vp
            getattr(L, pre + "field").argtypes = [vp, ci, ci]
            getattr(L, pre + "apply_A").argtypes = [vp, ci, vp, vp]
            getattr(L, pre + "smooth").argtypes = [vp, ci, ci]
            getattr(L, pre + "restrict").argtypes = [vp, ci]
            getattr(L, pre + "prolongate").argtypes = [vp, ci]
            getattr(L, pre + "vcycle").argtypes = [vp]
            getattr(L, pre + "init").argtypes = [vp, vp, ctypes.c_float if pre == "oracle_mg_" else ctypes.c_double]
            getattr(L, pre + "solve").restype = ci
            getattr(L, pre + "solve").argtypes = [vp, ci, ctypes.c_double, ctypes.c_double, ci,
                                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ci)]
            getattr(L, pre + "get_result").argtypes = [vp, vp]
            getattr(L, pre + "apply_bc").argtypes = [vp, vp, vp, vp]
            getattr(L, pre + "divergence").argtypes = [vp, vp, vp, vp]
            getattr(L, pre + "subtract_grad_p").argtypes = [vp, vp, vp, vp]
            getattr(L, pre + "set_solid").argtypes = [vp, vp]
            getattr(L, pre + "apply_solid_bc").argtypes = [vp, vp, vp, vp]
            getattr(L, pre + "poisson").restype = ci
            getattr(L, pre + "poisson").argtypes = [vp, vp, vp, vp, ci, ctypes.c_double, ci,
                                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ci)]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _res(res):
    r = list(res) + [1] * (3 - len(res))
    return np.array(r, dtype=np.int32)


def face_shape(res, axis):
    s = list(res)
    s[axis] += 1
    return tuple(s)


# --------------------------------------------------------------------------
# scalar primitives
# --------------------------------------------------------------------------
def N2(x: float) -> float:
    return lib().oracle_N2(float(x))


def dN2(x: float) -> float:
    return lib().oracle_dN2(float(x))


def jitter_u(seed: int, pid: int, comp: int) -> float:
    return lib().oracle_jitter_u(seed, pid, comp)


# --------------------------------------------------------------------------
# particle / grid kernels (in place on numpy arrays)
# --------------------------------------------------------------------------
def _u3(u):
    u = list(u) + [None] * (3 - len(u))
    return [_p(x) for x in u]


def rk4_particles(res, dx, pos, T, u, dt, advect_psi=True):
    """RK4_T_forward on pos (d,n) and T (d,d,n) (T[m] = column vector m)."""
    d, n = pos.shape
    Ts = [T[m] for m in range(d)] + [None] * (3 - d)
    lib().oracle_rk4_particles(d, _res(res), dx, n, _p(pos), *[_p(t) for t in Ts], *_u3(u),
                               float(dt), int(bool(advect_psi)))


def rk4_grid(res, dx, axis, psi, T, u, dt):
    d = len(res)
    lib().oracle_rk4_grid(d, _res(res), dx, axis, _p(psi), _p(T), *_u3(u), float(dt))


def advect_u(res, dx, axis, u0, u1_axis, T, psi):
    d = len(res)
    lib().oracle_advect_u(d, _res(res), dx, axis, *_u3(u0), _p(u1_axis), _p(T), _p(psi))


def g2p(res, dx, pos, u, imp, C=None):
    d, n = pos.shape
    lib().oracle_g2p(d, _res(res), dx, n, _p(pos), *_u3(u), _p(imp), _p(C))


def seed_particles(res, ppc_axis, dist, n, seed=-1, order=0):
    d = len(res)
    pos = np.empty((d, n), np.float32)
    lib().oracle_seed_particles(d, _res(res), ppc_axis, dist, n, _p(pos), int(seed), int(order))
    return pos


def reset_identity_grid(res, dx, axis):
    d = len(res)
    shp = face_shape(res, axis)
    psi = np.empty((d,) + shp, np.float32)
    T = np.empty((d,) + shp, np.float32)
    lib().oracle_reset_identity_grid(d, _res(res), dx, axis, _p(psi), _p(T))
    return psi, T


def update_T(Tgm, Tin):
    d, _, n = Tgm.shape
    out = np.empty_like(Tin)
    lib().oracle_update_T(d, n, _p(Tgm), _p(Tin), _p(out))
    return out


def p2g_prep(Tgm, Tin, C_in, init_imp):
    """Returns (C_out rows, imp).  Tin None -> 2D form (imp from Tgm as T)."""
    d, _, n = Tgm.shape
    C_out = np.empty_like(C_in)
    imp = np.empty((d, n), np.float32) if init_imp is not None else None
    lib().oracle_p2g_prep(d, n, _p(Tgm), _p(Tin), _p(C_in), _p(init_imp), _p(C_out), _p(imp))
    return C_out, imp


def update_imp(T, init_imp):
    d, _, n = T.shape
    imp = np.empty((d, n), np.float32)
    lib().oracle_update_imp(d, n, _p(T), _p(init_imp), _p(imp))
    return imp


def bin_particles(res, dx, pos):
    d, n = pos.shape
    C = int(np.prod(res))
    cell = np.empty(n, np.int32)
    counts = np.empty(C, np.int32)
    start = np.empty(C + 1, np.int32)
    perm = np.empty(n, np.int32)
    lib().oracle_bin_particles(d, _res(res), dx, n, _p(pos), _p(cell), _p(counts), _p(start), _p(perm))
    return cell, counts, start, perm


def cell_lists(res, counts, perm, start, cell_max):
    """Reference layout of get_particles_id_in_every_cell (3D/run_3D.py:342-351):
    ids[cell, slot] for slot < cell_max, -1 elsewhere; lowest ids first."""
    C = int(np.prod(res))
    ids = -np.ones((C, cell_max), np.int32)
    for c in np.nonzero(counts)[0]:
        m = min(int(counts[c]), cell_max)
        ids[c, :m] = perm[start[c]:start[c] + m]
    return ids.reshape(tuple(res) + (cell_max,))


def p2g(res, dx, pos, imp, C, with_weights=False):
    d, n = pos.shape
    u = [np.empty(face_shape(res, a), np.float32) for a in range(d)]
    W = [np.empty(face_shape(res, a), np.float32) for a in range(d)] if with_weights else [None] * d
    lib().oracle_p2g(d, _res(res), dx, n, _p(pos), _p(imp), _p(C), *_u3(u), *_u3(W))
    return (u, W) if with_weights else u


def max_speed(res, u):
    return float(lib().oracle_max_speed(len(res), _res(res), *_u3(u)))


def split_central(res, vc):
    d = len(res)
    u = [np.empty(face_shape(res, a), np.float32) for a in range(d)]
    lib().oracle_split_central(d, _res(res), _p(np.ascontiguousarray(vc, np.float32)), *_u3(u))
    return u


def get_central(res, u):
    d = len(res)
    vc = np.empty((d,) + tuple(res), np.float32)
    lib().oracle_get_central(d, _res(res), *_u3(u), _p(vc))
    return vc


def curl(res, dx, vc):
    """curl of the cell-centred vector vc (d, *res): (3, *res) in 3D, res in 2D."""
    d = len(res)
    out = np.empty((3,) + tuple(res) if d == 3 else tuple(res), np.float32)
    lib().oracle_curl(d, _res(res), dx, _p(np.ascontiguousarray(vc, np.float32)), _p(out))
    return out


def point_vortices(res, dx, centers, weights, rad=0.02, strength=-0.01):
    c = np.asarray(centers, np.float32)
    vc = np.empty((2,) + tuple(res), np.float32)
    lib().oracle_point_vortices(_res(res), dx, len(c), np.ascontiguousarray(c[:, 0]),
                                np.ascontiguousarray(c[:, 1]), np.asarray(weights, np.float32),
                                rad, strength, _p(vc))
    return vc


def vortex_ring(res, dx, vc, w, rad, delta, c, ex, ey, num_samples=2000):
    lib().oracle_vortex_ring(_res(res), dx, w, rad, delta, np.asarray(c, np.float32),
                             np.asarray(ex, np.float32), np.asarray(ey, np.float32),
                             num_samples, _p(vc))


def advect_smoke(res, dx, smoke0, psi, smoke1=None, mode=1):
    """advect_smoke / clamp_smoke (3D/run_3D.py:315-334, 2D/run_2D.py:444-450).
    smoke0: (nch, *res); psi: list of d arrays (d, *face_shape(a)); mode 1
    advects into a new array (or smoke1), 2 clamps smoke1 in place, 3 both."""
    d = len(res)
    s0 = np.ascontiguousarray(smoke0, np.float32)
    s1 = np.empty_like(s0) if smoke1 is None else smoke1
    ps = [np.ascontiguousarray(p, np.float32) for p in psi] + [None] * (3 - d)
    lib().oracle_advect_smoke(d, _res(res), dx, s0.shape[0], _p(s0), _p(s1), *[_p(p) for p in ps], mode)
    return s1


def vortex_ring_smoke(res, dx, smoke, rad, delta, c, ex, ey, color, num_samples=2000):
    lib().oracle_vortex_ring_smoke(_res(res), dx, rad, delta, np.asarray(c, np.float32),
                                   np.asarray(ex, np.float32), np.asarray(ey, np.float32), num_samples,
                                   np.asarray(color, np.float32), _p(smoke))


def g2p_smoke(res, dx, pos, smoke):
    """init_particles_smoke (3D/init_conditions.py:275-279): (p_smoke (4, n), p_grad (12, n))."""
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[1]
    ps = np.empty((4, n), np.float32)
    pg = np.empty((12, n), np.float32)
    lib().oracle_g2p_smoke(_res(res), dx, n, _p(pos), _p(np.ascontiguousarray(smoke, np.float32)), _p(ps), _p(pg))
    return ps, pg


def p2g_smoke(res, dx, pos, ps, grad=None, Tgm=None, Tin=None, apic=True):
    """P2G_smoke (3D/run_3D.py:554-593): (smoke (4, *res), weight res).  Tgm,
    Tin as (3, 3, n) column stacks."""
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[1]
    smoke = np.empty((4,) + tuple(res), np.float32)
    w = np.empty(tuple(res), np.float32)
    c = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
    lib().oracle_p2g_smoke(_res(res), dx, n, _p(pos), _p(c(ps)), _p(c(grad)), _p(c(Tgm)), _p(c(Tin)), int(apic),
                           _p(smoke), _p(w))
    return smoke, w


def clamp_smoke_particle(res, dx, start, ids, cell_max, init_pos, init_smoke, smoke):
    """clamp_smoke_particle (3D/run_3D.py:595-614) in place on smoke (4, *res)."""
    lib().oracle_clamp_smoke_particle(_res(res), dx, np.asarray(init_pos).shape[1], _p(np.ascontiguousarray(start)),
                                      _p(np.ascontiguousarray(ids)), int(cell_max),
                                      _p(np.ascontiguousarray(init_pos, np.float32)),
                                      _p(np.ascontiguousarray(init_smoke, np.float32)), _p(smoke))
    return smoke


def kinetic_energy(res, u):
    return float(lib().oracle_kinetic_energy(len(res), _res(res), *_u3(u)))


# --------------------------------------------------------------------------
# MGPCG (3D/mgpcg.py, 2D/mgpcg.py)
# --------------------------------------------------------------------------
class MG:
    """fp32 (the reference's) or fp64 (the Poisson oracle) MGPCG restatement."""

    def __init__(self, N, base_level=3, boundary_types=None, fp64=False, dirichlet_counts=None):
        self.dim = len(N)
        self.N = tuple(N)
        self.fp64 = fp64
        self.pre = "oracle_mg64_" if fp64 else "oracle_mg_"
        self.dt = np.float64 if fp64 else np.float32
        bt = np.array(boundary_types if boundary_types is not None else [[2, 2]] * self.dim,
                      np.int32).reshape(-1)
        if dirichlet_counts is None:
            dirichlet_counts = self.dim == 2
        self.h = self._f("create")(self.dim, _res(N), base_level, bt, int(dirichlet_counts))
        self.n_levels = self._f("n_levels")(self.h)

    def _f(self, name):
        return getattr(lib(), self.pre + name)

    def __del__(self):
        try:
            self._f("destroy")(self.h)
        except Exception:
            pass

    def level_shape(self, l):
        return tuple(n >> l for n in self.N)

    def field(self, which, level=0):
        """View of an internal field: r, z, x, p, Ap, u_div, pressure."""
        k = ["r", "z", "x", "p", "Ap", "u_div", "pressure"].index(which)
        shp = self.level_shape(level)
        ptr = self._f("field")(self.h, k, level)
        n = int(np.prod(shp))
        buf = (ctypes.c_double if self.fp64 else ctypes.c_float) * n
        return np.frombuffer(buf.from_address(ptr), dtype=self.dt).reshape(shp)

    def apply_A(self, level, p):
        p = np.ascontiguousarray(p, self.dt)
        out = np.empty_like(p)
        self._f("apply_A")(self.h, level, _p(p), _p(out))
        return out

    def smooth(self, l, phase):
        self._f("smooth")(self.h, l, phase)

    def restrict(self, l):
        self._f("restrict")(self.h, l)

    def prolongate(self, l):
        self._f("prolongate")(self.h, l)

    def vcycle(self, r):
        """z = M^-1 r with the reference V-cycle (3D/mgpcg.py:182-200)."""
        self.field("r", 0)[...] = r
        self._f("vcycle")(self.h)
        return self.field("z", 0).copy()

    def init(self, rhs, k=-1.0):
        rhs = np.ascontiguousarray(rhs, self.dt)
        self._f("init")(self.h, _p(rhs), k)

    def solve(self, max_iters=400, eps=1e-12, tol=1e-12, tol_mode=0):
        rtr = ctypes.c_double()
        conv = ctypes.c_int()
        it = self._f("solve")(self.h, max_iters, eps, tol, tol_mode, ctypes.byref(rtr), ctypes.byref(conv))
        return it, rtr.value, bool(conv.value)

    def get_result(self):
        out = np.empty(self.N, self.dt)
        self._f("get_result")(self.h, _p(out))
        return out

    def apply_bc(self, u):
        """MGPCG_3/MGPCG_2.apply_bc on float32 face arrays in place."""
        self._f("apply_bc")(self.h, *_u3(u))

    def set_solid(self, solid):
        """Solid obstacle cells (bool/uint8 array of the finest level's shape)."""
        m = np.ascontiguousarray(np.asarray(solid, bool).reshape(self.N), np.uint8)
        self._solid = m
        self._f("set_solid")(self.h, _p(m))

    def apply_solid_bc(self, u):
        """Zero the faces that touch a solid cell (free slip at the obstacle)."""
        self._f("apply_solid_bc")(self.h, *_u3(u))

    def divergence(self, u):
        """divergence -> the u_div field (3D/mgpcg.py:380-390)."""
        self._f("divergence")(self.h, *_u3(u))

    def poisson(self, u, max_iters=400, tol=1e-12, tol_mode=0):
        """Poisson on float32 face arrays in place; returns (iters, rTr, converged)."""
        rtr = ctypes.c_double()
        conv = ctypes.c_int()
        it = self._f("poisson")(self.h, *_u3(u), max_iters, tol, tol_mode, ctypes.byref(rtr),
                                ctypes.byref(conv))
        return it, rtr.value, bool(conv.value)


def poisson_fp64_solve(rhs, base_level=3, boundary_types=None, tol=1e-8, solid=None):
    """FP64 oracle: solve A x = -rhs (the reference's init(k=-1)) in double to a
    relative residual `tol` (PCG with the reference V-cycle stalls near 1e-9,
    giving solutions within ~1e-8 relative of a direct solve)."""
    N = rhs.shape
    mg = MG(N, base_level=base_level, boundary_types=boundary_types, fp64=True)
    if solid is not None:
        mg.set_solid(solid)
    mg.init(rhs.astype(np.float64), -1.0)
    it, rtr, conv = mg.solve(max_iters=200, tol=tol, tol_mode=1)
    return mg.get_result(), it, conv


# --------------------------------------------------------------------------
# schedule (row 30): 3D/run_3D.py:716-776, 2D/run_2D.py:772-856
# --------------------------------------------------------------------------
def reinit_schedule(n_steps, reinit_every, reinit_every_grad_m, reinit_particle_pos=True, dim=3):
    """Per step i: (long_reinit, reseed, short_reinit).  The 3D driver sets ik
    only when it reseeds (:751-755); the 2D driver sets ik at every long
    reinit (:819)."""
    out = []
    ik = 0
    for i in range(n_steps):
        j = i % reinit_every
        long_r = j == 0
        reseed = long_r and (reinit_particle_pos or i == 0)
        if long_r and (reseed or dim == 2):
            ik = i
        k = (i - ik) % reinit_every_grad_m
        out.append((long_r, reseed, k == 0))
    return out


def cfl_dt(max_speed_f32, dx, CFL, sub_t, visualize_dt):
    """dt logic of 3D/run_3D.py:725-743 in Python double; returns (dt, sub_t, frame_end)."""
    curr_dt = CFL * dx / float(np.float32(max_speed_f32))
    if sub_t + curr_dt >= visualize_dt:
        return visualize_dt - sub_t, 0.0, True
    return curr_dt, sub_t + curr_dt, False


# --------------------------------------------------------------------------
# full step (rows 11, 13-14, 17, 28, 30): 3D/run_3D.py:716-831,
# 2D/run_2D.py:731-909 on numpy arrays
# --------------------------------------------------------------------------
class OracleSim:
    """CPU restatement of the reference main loop (one scene)."""

    def __init__(self, res, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, visualize_dt=0.1,
                 particles_per_cell=8, base_level=3, max_iters=400, tol=1e-12, tol_mode=0, seed=-1,
                 particle_order=0, use_midpoint_vel=True, reinit_particle_pos=True):
        self.res = tuple(res)
        self.d = d = len(res)
        self.dx = 1.0 / res[1]
        self.ppc_axis = 2 if d == 3 else int(math.isqrt(particles_per_cell))
        self.dist = self.dx / self.ppc_axis
        self.n = int(np.prod(res)) * particles_per_cell
        self.reinit_every, self.reinit_every_grad_m = reinit_every, reinit_every_grad_m
        self.CFL, self.visualize_dt = CFL, visualize_dt
        self.max_iters, self.tol, self.tol_mode = max_iters, tol, tol_mode
        self.seed, self.order = seed, particle_order
        self.use_midpoint_vel, self.reinit_particle_pos = use_midpoint_vel, reinit_particle_pos
        self.mg = MG(res, base_level=base_level)
        n = self.n
        self.u = [np.zeros(face_shape(res, a), np.float32) for a in range(d)]
        self.pos = np.zeros((d, n), np.float32)
        self.init_imp = np.zeros((d, n), np.float32)
        self.imp_grad_m = np.zeros((d, n), np.float32)
        self.T_init = np.zeros((d, d, n), np.float32)
        self.T_grad_m = np.zeros((d, d, n), np.float32)
        self.init_C = np.zeros((d, d, n), np.float32)
        if d == 2:
            self.imp = np.zeros((d, n), np.float32)
            self.T = np.zeros((d, d, n), np.float32)
        self.i, self.ik, self.sub_t = -1, 0, 0.0
        self.iters = []

    @staticmethod
    def _eye(T):
        d = T.shape[0]
        T[...] = 0
        for m in range(d):
            T[m, m] = 1.0

    def project(self):
        it, rtr, conv = self.mg.poisson(self.u, self.max_iters, self.tol, self.tol_mode)
        self.iters.append(it)
        return it

    def set_velocity_from_central(self, vc, project=True):
        self.u = split_central(self.res, vc)
        if project:
            self.project()

    def seed_pos(self):
        self.pos = seed_particles(self.res, self.ppc_axis, self.dist, self.n, self.seed, self.order)

    def initialize_2d_particles(self):
        """2D/run_2D.py:731-738."""
        self.seed_pos()
        g2p(self.res, self.dx, self.pos, self.u, self.imp_grad_m, self.init_C)
        self._eye(self.T_grad_m)
        g2p(self.res, self.dx, self.pos, self.u, self.imp, None)
        self.init_imp[...] = self.imp

    def compute_dt(self):
        ms = max_speed(self.res, self.u)
        dt, self.sub_t, _ = cfl_dt(ms, self.dx, self.CFL, self.sub_t, self.visualize_dt)
        return dt

    def midpoint(self, dt):
        psi, T = [], []
        for a in range(self.d):
            p_, t_ = reset_identity_grid(self.res, self.dx, a)
            rk4_grid(self.res, self.dx, a, p_, t_, self.u, 0.5 * dt)
            psi.append(p_)
            T.append(t_)
        u0 = [x.copy() for x in self.u]
        for a in range(self.d):
            advect_u(self.res, self.dx, a, u0, self.u[a], T[a], psi[a])
        self.project()

    def step(self, dt=None):
        self.i += 1
        i = self.i
        dt = self.compute_dt() if dt is None else dt
        d = self.d
        if d == 3:
            if i % self.reinit_every == 0:
                if self.reinit_particle_pos or i == 0:
                    self.seed_pos()
                    self.ik = i
                g2p(self.res, self.dx, self.pos, self.u, self.init_imp, None)
                self._eye(self.T_init)
                self._eye(self.T_grad_m)
            if (i - self.ik) % self.reinit_every_grad_m == 0:
                g2p(self.res, self.dx, self.pos, self.u, self.imp_grad_m, self.init_C)
                self.T_init = update_T(self.T_grad_m, self.T_init)
                self._eye(self.T_grad_m)
        else:
            if i % self.reinit_every == 0:
                if self.reinit_particle_pos or i == 0:
                    self.seed_pos()
                self.ik = i
                g2p(self.res, self.dx, self.pos, self.u, self.imp, None)
                self.init_imp[...] = self.imp
                self._eye(self.T)
                self._eye(self.T_init)
            if (i - self.ik) % self.reinit_every_grad_m == 0:
                g2p(self.res, self.dx, self.pos, self.u, self.imp_grad_m, self.init_C)
                self._eye(self.T_grad_m)
                self.T_init[...] = self.T
        if self.use_midpoint_vel:
            self.midpoint(dt)
        rk4_particles(self.res, self.dx, self.pos, self.T_grad_m, self.u, dt, True)
        if d == 3:
            C, imp = p2g_prep(self.T_grad_m, self.T_init, self.init_C, self.init_imp)
        else:
            C, _ = p2g_prep(self.T_grad_m, None, self.init_C, None)
            self.T = update_T(self.T_grad_m, self.T_init)
            self.imp = update_imp(self.T, self.init_imp)
            imp = self.imp
        self.u = p2g(self.res, self.dx, self.pos, imp, C)
        self.project()
        return dt
