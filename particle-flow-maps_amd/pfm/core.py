"""Dimension-generic host wrappers over the libpfm C ABI.

Field conventions (all float32, contiguous, on the GPU):
  * face field of axis a: shape ``res`` with ``res[a] + 1`` (the reference's
    ``u_x`` is ``(rx+1, ry, rz)``);
  * particle vector field: ``(d, n)`` SoA; a per-particle matrix is held as d
    vector fields (``T_x, T_y, T_z`` columns or ``C_x, C_y, C_z`` rows);
  * grid-attached vector field (``psi_x_grid``, ``T_x_grid``): ``(d, *face_shape)``.

The reference-named modules (``pfm.d3.run_3D``, ``pfm.d2.run_2D`` ...) are thin
argument-order adapters over these functions.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Sequence

import torch

from . import _lib as L


def face_shape(res: Sequence[int], axis: int):
    s = list(res)
    s[axis] += 1
    return tuple(s)


def res_from_faces(u_x: torch.Tensor, dim: Optional[int] = None):
    """Cell resolution of a grid from its x-face field (rx+1, ry[, rz])."""
    s = list(u_x.shape)
    s[0] -= 1
    return tuple(s)


def default_dx(res) -> float:
    """dx = 1/res_y (3D/run_3D.py:16, 2D/run_2D.py:14)."""
    return 1.0 / res[1]


def _stream(stream):
    return L.stream_ptr(stream)


# Writes made by the library through raw pointers do not bump torch's version
# counter, so positions written here are tracked separately for the cached
# binning in ``Binning.bin``.
_WRITES = {}


def _mark_written(t):
    k = t.data_ptr()
    _WRITES[k] = _WRITES.get(k, 0) + 1


def write_count(t) -> int:
    return _WRITES.get(t.data_ptr(), 0)


def _u_ptrs(u, res):
    d = len(res)
    out = []
    for a in range(3):
        if a < d:
            t = u[a]
            if tuple(t.shape) != face_shape(res, a):
                raise L.PfmError(f"face field {a} has shape {tuple(t.shape)}, expected {face_shape(res, a)}")
            out.append(L.ptr(t, name=f"u[{a}]"))
        else:
            out.append(None)
    return out


def _pvec(t, d, n, name):
    if tuple(t.shape) != (d, n):
        raise L.PfmError(f"{name} must have shape ({d}, {n}), got {tuple(t.shape)}")
    return L.ptr(t, name=name)


def _cols(Ts, d, n, name):
    out = []
    for m in range(3):
        out.append(_pvec(Ts[m], d, n, f"{name}[{m}]") if m < d else None)
    return out


# --------------------------------------------------------------------------
# flow-map advection
# --------------------------------------------------------------------------
def rk4_particles(res, dx, pos, T, u, dt, advect_psi=True, stream=None):
    d, n = pos.shape
    g = L.make_grid(res, dx)
    L.call("pfm_rk4_particles", ctypes.byref(g), n, _pvec(pos, d, n, "pos"), *_cols(T, d, n, "T"),
           *_u_ptrs(u, res), float(dt), int(bool(advect_psi)), _stream(stream))
    if advect_psi:
        _mark_written(pos)


def face_axis(res, t, lead=1):
    """Which face axis a grid-attached field (lead dims + face shape) belongs to."""
    sp = tuple(t.shape[lead:])
    for a in range(len(res)):
        if sp == face_shape(res, a):
            return a
    raise L.PfmError(f"shape {tuple(t.shape)} is not a face-attached field of grid {tuple(res)}")


def rk4_grid(res, dx, psi, T, u, dt, stream=None):
    d = len(res)
    a = face_axis(res, psi)
    if tuple(T.shape) != tuple(psi.shape) or psi.shape[0] != d:
        raise L.PfmError("psi and T must both be (d, *face_shape)")
    g = L.make_grid(res, dx)
    L.call("pfm_rk4_grid", ctypes.byref(g), a, L.ptr(psi, name="psi"), L.ptr(T, name="T"),
           *_u_ptrs(u, res), float(dt), _stream(stream))


def advect_u(res, dx, u0, u1, T, psi, stream=None):
    d = len(res)
    g = L.make_grid(res, dx)
    Tp, Pp = [], []
    for a in range(3):
        if a < d:
            if face_axis(res, T[a]) != a or face_axis(res, psi[a]) != a:
                raise L.PfmError(f"T/psi of axis {a} have the wrong face shape")
            Tp.append(L.ptr(T[a], name=f"T[{a}]"))
            Pp.append(L.ptr(psi[a], name=f"psi[{a}]"))
        else:
            Tp.append(None)
            Pp.append(None)
    L.call("pfm_advect_u", ctypes.byref(g), *_u_ptrs(u0, res), *_u_ptrs(u1, res), *Tp, *Pp,
           _stream(stream))


def midpoint_advect(res, dx, u0, u1, psi, T, dt, stream=None):
    """reset_identity_grid + rk4_grid(dt) on every axis, then advect_u(u0 -> u1),
    in one launch (3D/run_3D.py:778-784).  u0 is the copy of the velocity that
    both sample; u1 must not alias it."""
    d = len(res)
    g = L.make_grid(res, dx)
    Pp, Tp = [], []
    for a in range(3):
        if a < d:
            if face_axis(res, psi[a]) != a or tuple(T[a].shape) != tuple(psi[a].shape) or psi[a].shape[0] != d:
                raise L.PfmError(f"psi/T of axis {a} must both be (d, *face_shape(axis {a}))")
            Pp.append(L.ptr(psi[a], name=f"psi[{a}]"))
            Tp.append(L.ptr(T[a], name=f"T[{a}]"))
        else:
            Pp.append(None)
            Tp.append(None)
    L.call("pfm_midpoint_advect", ctypes.byref(g), *_u_ptrs(u0, res), *_u_ptrs(u1, res), *Pp, *Tp, float(dt),
           _stream(stream))


def reset_identity_grid(res, dx, psi, T, stream=None):
    g = L.make_grid(res, dx)
    a = face_axis(res, psi)
    L.call("pfm_reset_identity_grid", ctypes.byref(g), a, L.ptr(psi, name="psi"), L.ptr(T, name="T"),
           _stream(stream))


def reset_T_identity(T, stream=None):
    d, n = T[0].shape
    L.call("pfm_reset_T_identity", d, n, *_cols(T, d, n, "T"), _stream(stream))


# --------------------------------------------------------------------------
# particle <-> grid
# --------------------------------------------------------------------------
def seed_particles(res, ppc_axis, dist, pos, seed=-1, order=0, stream=None):
    d, n = pos.shape
    g = L.make_grid(res, 1.0 / res[1])
    L.call("pfm_seed_particles", ctypes.byref(g), int(ppc_axis), float(dist), n, _pvec(pos, d, n, "pos"),
           int(seed), int(order), _stream(stream))
    _mark_written(pos)


def g2p(res, dx, pos, u, imp, C=None, stream=None):
    d, n = pos.shape
    g = L.make_grid(res, dx)
    Cp = _cols(C, d, n, "C") if C is not None else [None, None, None]
    ip = _pvec(imp, d, n, "imp") if imp is not None else None
    L.call("pfm_g2p", ctypes.byref(g), n, _pvec(pos, d, n, "pos"), *_u_ptrs(u, res), ip, *Cp,
           _stream(stream))


def update_T(Tgm, Tin, Tout, stream=None):
    d, n = Tgm[0].shape
    L.call("pfm_update_T", d, n, *_cols(Tgm, d, n, "T_grad_m"), *_cols(Tin, d, n, "T_in"),
           *_cols(Tout, d, n, "T_out"), _stream(stream))


def update_particles_imp(T, init_imp, imp, stream=None):
    d, n = init_imp.shape
    L.call("pfm_update_particles_imp", d, n, *_cols(T, d, n, "T"), _pvec(init_imp, d, n, "init_imp"),
           _pvec(imp, d, n, "imp"), _stream(stream))


def update_particles_grad_m(C, init_C, Tgm, stream=None):
    d, n = C[0].shape
    L.call("pfm_update_particles_grad_m", d, n, *_cols(C, d, n, "C"), *_cols(init_C, d, n, "init_C"),
           *_cols(Tgm, d, n, "T_grad_m"), _stream(stream))


class Binning:
    """Cell sort of one particle set (pfm_bin_particles) plus the P2G workspace.

    Keeps the last sort and re-bins only when the position tensor changed
    (tracked by data pointer and torch's in-place version counter)."""

    def __init__(self, res, n, device=None):
        self.res = tuple(res)
        self.n = int(n)
        dev = device or torch.device("cuda")
        C = math.prod(self.res)
        g = L.make_grid(self.res, 1.0 / self.res[1])
        ws = L.load().pfm_p2g_workspace_size(ctypes.byref(g), self.n)
        if ws == 0:
            raise L.PfmError(f"pfm_p2g_workspace_size failed: {L.last_error()}")
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=dev)
        self.cell_start = torch.empty(C + 1, dtype=torch.int32, device=dev)
        self.perm = torch.empty(max(self.n, 1), dtype=torch.int32, device=dev)
        self.cell_of = torch.empty(max(self.n, 1), dtype=torch.int32, device=dev)
        self._key = None

    def key(self, pos):
        return (pos.data_ptr(), pos._version, write_count(pos), tuple(pos.shape))

    def bin(self, pos, dx, stream=None, force=False):
        if not force and self._key == self.key(pos):
            return
        d, n = pos.shape
        if n != self.n:
            raise L.PfmError(f"Binning built for {self.n} particles, got {n}")
        g = L.make_grid(self.res, dx)
        L.call("pfm_bin_particles", ctypes.byref(g), n, _pvec(pos, d, n, "pos"),
               L.ptr(self.cell_of, dtype=torch.int32), L.ptr(self.cell_start, dtype=torch.int32),
               L.ptr(self.perm, dtype=torch.int32), self.workspace.data_ptr(), self.workspace.numel(),
               _stream(stream))
        self._key = self.key(pos)

    def cell_lists(self, dx, ids, counts, stream=None):
        g = L.make_grid(self.res, dx)
        cell_max = ids.shape[-1]
        L.call("pfm_cell_lists", ctypes.byref(g), L.ptr(self.cell_start, dtype=torch.int32),
               L.ptr(self.perm, dtype=torch.int32), int(cell_max), L.ptr(ids, dtype=torch.int32),
               L.ptr(counts, dtype=torch.int32), _stream(stream))

    def p2g(self, dx, pos, imp, C, u, W=None, Tgm=None, Tin=None, stream=None):
        d, n = pos.shape
        self.bin(pos, dx, stream)
        g = L.make_grid(self.res, dx)
        Wp = [L.ptr(W[a], name=f"W[{a}]") if (W is not None and a < d) else None for a in range(3)]
        if W is not None:
            for a in range(d):
                if tuple(W[a].shape) != face_shape(self.res, a):
                    raise L.PfmError(f"weight field {a} has the wrong shape")
        Gp = _cols(Tgm, d, n, "T_grad_m") if Tgm is not None else [None] * 3
        Ip = _cols(Tin, d, n, "T_init") if Tin is not None else [None] * 3
        L.call("pfm_p2g", ctypes.byref(g), n, _pvec(pos, d, n, "pos"), _pvec(imp, d, n, "imp"),
               *_cols(C, d, n, "C"), *Gp, *Ip, L.ptr(self.cell_start, dtype=torch.int32),
               L.ptr(self.perm, dtype=torch.int32), *_u_ptrs(u, self.res), *Wp,
               self.workspace.data_ptr(), self.workspace.numel(), _stream(stream))


    def rk4_record(self, dx, pos, Tgm, u_vel, dt, Tin, init_C, init_imp, stream=None, fold=False):
        """RK4_T_forward (3D/run_3D.py:795-796) with the P2G particle
        composition (3D/run_3D.py:460-479) fused in: advects pos/T_grad_m in
        place and writes one P2G record per particle into the workspace.
        fold: T_grad_m is the identity on entry and stays so in memory, and
        T_init becomes cols(T_grad_m) cols(T_init) (the next update_T)."""
        d, n = pos.shape
        g = L.make_grid(self.res, dx)
        lib = L.load()
        rec = lib.pfm_p2g_record_buffer(ctypes.byref(g), n, self.workspace.data_ptr())
        if not rec:
            raise L.PfmError("pfm_p2g_record_buffer failed")
        if fold:
            L.call("pfm_rk4_particles_p2g_record_fold", ctypes.byref(g), n, _pvec(pos, d, n, "pos"),
                   *_u_ptrs(u_vel, self.res), float(dt), *_cols(Tin, d, n, "T_init"), *_cols(init_C, d, n, "init_C"),
                   _pvec(init_imp, d, n, "init_imp"), rec, _stream(stream))
        else:
            L.call("pfm_rk4_particles_p2g_record", ctypes.byref(g), n, _pvec(pos, d, n, "pos"),
                   *_cols(Tgm, d, n, "T_grad_m"), *_u_ptrs(u_vel, self.res), float(dt),
                   *_cols(Tin, d, n, "T_init"), *_cols(init_C, d, n, "init_C"), _pvec(init_imp, d, n, "init_imp"),
                   rec, _stream(stream))
        _mark_written(pos)

    def rk4_record_count(self, dx, pos, u_vel, dt, Tin, init_C, init_imp, stream=None):
        """rk4_record(fold=True) with the count pass of the binning of the
        advected positions in the same kernel (finish with bin_finish)."""
        d, n = pos.shape
        g = L.make_grid(self.res, dx)
        L.call("pfm_rk4_particles_p2g_record_fold_count", ctypes.byref(g), n, _pvec(pos, d, n, "pos"),
               *_u_ptrs(u_vel, self.res), float(dt), *_cols(Tin, d, n, "T_init"), *_cols(init_C, d, n, "init_C"),
               _pvec(init_imp, d, n, "init_imp"), L.ptr(self.cell_of, dtype=torch.int32), self.workspace.data_ptr(),
               self.workspace.numel(), _stream(stream))
        _mark_written(pos)

    def bin_finish(self, pos, dx, stream=None):
        """The rest of bin() after rk4_record_count."""
        g = L.make_grid(self.res, dx)
        L.call("pfm_bin_finish", ctypes.byref(g), pos.shape[1], L.ptr(self.cell_of, dtype=torch.int32),
               L.ptr(self.cell_start, dtype=torch.int32), L.ptr(self.perm, dtype=torch.int32),
               self.workspace.data_ptr(), self.workspace.numel(), _stream(stream))
        self._key = self.key(pos)

    def p2g_from_records(self, dx, n, u_out, W=None, stream=None):
        """P2G (3D/run_3D.py:812, 3D/taichi_utils.py P2G) from the records of
        rk4_record, using the current binning of the advected positions."""
        d = len(self.res)
        g = L.make_grid(self.res, dx)
        Wp = [L.ptr(W[a], name=f"W[{a}]") if (W is not None and a < d) else None for a in range(3)]
        L.call("pfm_p2g_records", ctypes.byref(g), n, L.ptr(self.cell_start, dtype=torch.int32),
               L.ptr(self.perm, dtype=torch.int32), *_u_ptrs(u_out, self.res), *Wp, self.workspace.data_ptr(),
               self.workspace.numel(), _stream(stream))

    def rk4_then_p2g(self, dx, pos, Tgm, u_vel, dt, Tin, init_C, init_imp, u_out, W=None, stream=None, fold=False):
        """RK4_T_forward followed by P2G (3D/run_3D.py:795-812) with the P2G
        composition fused into the RK4 kernel: advect pos/T_grad_m in place,
        bin the new positions, then scatter into u_out (may alias u_vel only
        after the RK4 reads are done -- the launches are ordered on one stream)."""
        if fold:
            # the binning's count pass rides in the RK4 kernel
            self.rk4_record_count(dx, pos, u_vel, dt, Tin, init_C, init_imp, stream)
            self.bin_finish(pos, dx, stream)
        else:
            self.rk4_record(dx, pos, Tgm, u_vel, dt, Tin, init_C, init_imp, stream, fold=fold)
            self.bin(pos, dx, stream, force=True)
        self.p2g_from_records(dx, pos.shape[1], u_out, W, stream)


_BINNING_CACHE = {}


def binning_for(res, n, device=None) -> Binning:
    key = (tuple(res), int(n), str(device))
    b = _BINNING_CACHE.get(key)
    if b is None:
        b = Binning(res, n, device)
        _BINNING_CACHE[key] = b
    return b


# --------------------------------------------------------------------------
# grid utilities
# --------------------------------------------------------------------------
def max_speed(res, u, out=None, stream=None):
    """Device float32 tensor [1] with calc_max_speed's value."""
    if out is None:
        out = torch.empty(1, dtype=torch.float32, device=u[0].device)
    g = L.make_grid(res, default_dx(res))
    L.call("pfm_max_speed", ctypes.byref(g), *_u_ptrs(u, res), L.ptr(out, name="out"), _stream(stream))
    return out


def split_central(res, vc, u, stream=None):
    g = L.make_grid(res, default_dx(res))
    if tuple(vc.shape) != (len(res),) + tuple(res):
        raise L.PfmError(f"central field must be (d, *res), got {tuple(vc.shape)}")
    L.call("pfm_split_central", ctypes.byref(g), L.ptr(vc, name="vc"), *_u_ptrs(u, res), _stream(stream))


def get_central(res, u, vc, stream=None):
    g = L.make_grid(res, default_dx(res))
    if tuple(vc.shape) != (len(res),) + tuple(res):
        raise L.PfmError(f"central field must be (d, *res), got {tuple(vc.shape)}")
    L.call("pfm_get_central", ctypes.byref(g), *_u_ptrs(u, res), L.ptr(vc, name="vc"), _stream(stream))


def advect_smoke(res, dx, smoke0, smoke1, psi, mode=1, stream=None):
    """advect_smoke / clamp_smoke on the grid (3D/run_3D.py:315-334,
    2D/run_2D.py:444-450): smoke (nch, *res) channel-major, psi the d face maps
    (d, *face_shape(a)).  mode 1: smoke1 = interp_1(smoke0, psi_c); 2: clamp
    smoke1 in place to smoke0's range around psi_c (3D); 3: both."""
    d = len(res)
    nch = smoke0.shape[0]
    for name, t in (("smoke0", smoke0), ("smoke1", smoke1)):
        if tuple(t.shape) != (nch,) + tuple(res):
            raise L.PfmError(f"{name} must be ({nch}, *{tuple(res)}), got {tuple(t.shape)}")
    pp = []
    for a in range(3):
        if a < d:
            if tuple(psi[a].shape) != (d,) + face_shape(res, a):
                raise L.PfmError(f"psi[{a}] must be ({d}, *{face_shape(res, a)}), got {tuple(psi[a].shape)}")
            pp.append(L.ptr(psi[a], name=f"psi[{a}]"))
        else:
            pp.append(None)
    g = L.make_grid(res, dx)
    L.call("pfm_advect_smoke", ctypes.byref(g), int(nch), L.ptr(smoke0, name="smoke0"),
           L.ptr(smoke1, name="smoke1"), *pp, int(mode), _stream(stream))


def vortex_ring_smoke(res, dx, smoke, rad, delta, c, ex, ey, color, num_samples=2000, stream=None):
    """The smoke part of add_vortex_ring_and_smoke (3D/init_conditions.py:31-56)
    on smoke (4, *res)."""
    import numpy as np

    if tuple(smoke.shape) != (4,) + tuple(res):
        raise L.PfmError(f"smoke must be (4, *{tuple(res)}), got {tuple(smoke.shape)}")
    arr = [np.ascontiguousarray(np.asarray(v, np.float32)) for v in (c, ex, ey, color)]
    g = L.make_grid(res, dx)
    L.call("pfm_vortex_ring_smoke", ctypes.byref(g), float(rad), float(delta), *[v.ctypes.data for v in arr[:3]],
           int(num_samples), arr[3].ctypes.data, L.ptr(smoke, name="smoke"), _stream(stream))


def g2p_smoke(res, dx, pos, smoke, p_smoke, p_grad, stream=None):
    """init_particles_smoke (3D/init_conditions.py:275-279): p_smoke (4, n) and
    p_grad (12, n; row-major 3x4 per particle) from the cell smoke (4, *res)."""
    d, n = pos.shape
    if d != 3 or tuple(smoke.shape) != (4,) + tuple(res):
        raise L.PfmError("particle smoke is 3D with a (4, *res) smoke field")
    g = L.make_grid(res, dx)
    L.call("pfm_g2p_smoke", ctypes.byref(g), n, _pvec(pos, 3, n, "pos"), L.ptr(smoke, name="smoke"),
           _pvec(p_smoke, 4, n, "p_smoke"), _pvec(p_grad, 12, n, "p_grad"), _stream(stream))


def p2g_smoke(res, dx, pos, p_smoke, init_grad, Tgm, Tin, binning, smoke, weight, use_apic=True, stream=None):
    """P2G_smoke (3D/run_3D.py:554-593) over the cell sort of `binning`
    (a Binning already binned at these positions)."""
    d, n = pos.shape
    g = L.make_grid(res, dx)
    Gp = _cols(Tgm, 3, n, "T_grad_m") if use_apic else [None] * 3
    Ip = _cols(Tin, 3, n, "T_init") if use_apic else [None] * 3
    gp = _pvec(init_grad, 12, n, "init_grad") if use_apic else None
    L.call("pfm_p2g_smoke", ctypes.byref(g), n, _pvec(pos, 3, n, "pos"), _pvec(p_smoke, 4, n, "p_smoke"), gp,
           *Gp, *Ip, int(bool(use_apic)), L.ptr(binning.cell_start, dtype=torch.int32),
           L.ptr(binning.perm, dtype=torch.int32), L.ptr(smoke, name="smoke"), L.ptr(weight, name="weight"),
           _stream(stream))


def clamp_smoke_particle(res, dx, binning, cell_max, init_pos, init_smoke, smoke, stream=None):
    """clamp_smoke_particle (3D/run_3D.py:595-614) over the cell lists of `binning`."""
    n = init_pos.shape[1]
    g = L.make_grid(res, dx)
    L.call("pfm_clamp_smoke_particle", ctypes.byref(g), n, L.ptr(binning.cell_start, dtype=torch.int32),
           L.ptr(binning.perm, dtype=torch.int32), int(cell_max), _pvec(init_pos, 3, n, "init_pos"),
           L.ptr(init_smoke, name="init_smoke"), L.ptr(smoke, name="smoke"), _stream(stream))


def curl(res, dx, vc, out, stream=None):
    """curl (3D/taichi_utils.py:56-78; 2D/taichi_utils.py:97-103) of the
    cell-centred vector vc (d, *res): out is (3, *res) in 3D, res in 2D."""
    d = len(res)
    g = L.make_grid(res, dx)
    if tuple(vc.shape) != (d,) + tuple(res):
        raise L.PfmError(f"central field must be (d, *res), got {tuple(vc.shape)}")
    want = (3,) + tuple(res) if d == 3 else tuple(res)
    if tuple(out.shape) != want:
        raise L.PfmError(f"curl output must be {want}, got {tuple(out.shape)}")
    L.call("pfm_curl", ctypes.byref(g), L.ptr(vc, name="vc"), L.ptr(out, name="out"), _stream(stream))


def point_vortices(res, dx, centers, weights, vc, rad=0.02, strength=-0.01, stream=None):
    import numpy as np

    c = np.ascontiguousarray(np.asarray(centers, np.float32).reshape(-1, 2))
    cx = np.ascontiguousarray(c[:, 0])
    cy = np.ascontiguousarray(c[:, 1])
    w = np.ascontiguousarray(np.asarray(weights, np.float32))
    g = L.make_grid(res, dx)
    L.call("pfm_point_vortices", ctypes.byref(g), len(c), cx.ctypes.data, cy.ctypes.data, w.ctypes.data,
           float(rad), float(strength), L.ptr(vc, name="vc"), _stream(stream))


def vortex_ring(res, dx, vc, w, rad, delta, c, ex, ey, num_samples=2000, stream=None):
    import numpy as np

    cc = np.asarray(c, np.float32)
    xx = np.asarray(ex, np.float32)
    yy = np.asarray(ey, np.float32)
    g = L.make_grid(res, dx)
    L.call("pfm_vortex_ring", ctypes.byref(g), float(w), float(rad), float(delta), cc.ctypes.data,
           xx.ctypes.data, yy.ctypes.data, int(num_samples), L.ptr(vc, name="vc"), _stream(stream))


def add_fields(f1, f2, dest, multiplier, stream=None):
    """add_fields (3D/taichi_utils.py:18-21, the same kernel in 2D):
    dest = f1 + multiplier * f2 elementwise over same-shaped float32 fields."""
    if not (f1.shape == f2.shape == dest.shape):
        raise L.PfmError(f"add_fields shapes differ: {tuple(f1.shape)}, {tuple(f2.shape)}, {tuple(dest.shape)}")
    L.call("pfm_add_fields", int(dest.numel()), L.ptr(f1, name="f1"), L.ptr(f2, name="f2"),
           L.ptr(dest, name="dest"), float(multiplier), _stream(stream))


_KE_SCRATCH = {}


def kinetic_energy(res, u, stream=None) -> torch.Tensor:
    """Device float64 tensor [1]: sum_a sum u_a^2 / |F_a|."""
    dev = u[0].device
    # one scratch per (device, stream): calls on different streams may overlap
    key = (str(dev), _stream(stream))
    buf = _KE_SCRATCH.get(key)
    if buf is None:
        buf = torch.empty(1 + 3 * 1024, dtype=torch.float64, device=dev)
        _KE_SCRATCH[key] = buf
    g = L.make_grid(res, default_dx(res))
    L.call("pfm_kinetic_energy", ctypes.byref(g), *_u_ptrs(u, res), L.ptr(buf, dtype=torch.float64),
           _stream(stream))
    return buf[:1].clone()
