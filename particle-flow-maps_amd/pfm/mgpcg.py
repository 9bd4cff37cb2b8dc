"""MGPCG pressure solver with the reference's object interface.

``MGPCG`` / ``MGPCG_2`` / ``MGPCG_3`` follow ``3D/mgpcg.py:7-430`` and
``2D/mgpcg.py``: the constructor ``(boundary_types, N, base_level=3,
real=float)`` builds the hierarchy (on the GPU, through ``pfm_mg_create``),
``Poisson(u_x, u_y[, u_z])`` projects a MAC velocity in place, and
``init`` / ``solve`` / ``get_result`` / ``apply_preconditioner`` expose the
same building blocks.  Extensions (keyword-only): ``tol_mode`` ("abs", the
reference's ``rTr < tol``, or "rel": ``|r|/|r0| < tol``), ``max_iters`` for
``Poisson``, and the last solve's ``iterations`` / ``residual``.
"""
from __future__ import annotations

import ctypes
import math
import time

import torch

from . import _lib as L
from .core import face_shape


class _DevArray:
    """__cuda_array_interface__ view of a library-owned float32 buffer."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = {
            "shape": tuple(int(s) for s in shape),
            "typestr": "<f4",
            "data": (int(ptr), False),
            "version": 3,
            "strides": None,
        }


def _bt_list(boundary_types, dim):
    if boundary_types is None:
        return [[2, 2]] * dim
    if hasattr(boundary_types, "tolist"):
        boundary_types = boundary_types.tolist()
    bt = [list(map(int, row)) for row in boundary_types]
    if len(bt) != dim or any(len(r) != 2 for r in bt):
        raise ValueError(f"boundary_types must be {dim}x2")
    return bt


class MGPCG:
    """Grid MGPCG solver (3D/mgpcg.py:8-280) on the GPU."""

    def __init__(self, boundary_types, N, dim=2, base_level=3, real=float, *, precision=0):
        if real not in (float, torch.float32, "float32"):
            raise ValueError("only fp32 fields are supported (the reference's default)")
        self.use_multigrid = True
        self.N = [int(n) for n in N]
        self.dim = dim
        self.real = real
        self.base_level = base_level
        self.pre_and_post_smoothing = 2
        self.bottom_smoothing = 50
        self.boundary_types = _bt_list(boundary_types, dim)
        self.num_entries = math.prod(self.N)
        self.precision = precision
        lib = L.load()
        Narr = (ctypes.c_int32 * 3)(*(self.N + [1] * (3 - dim)))
        bc = (ctypes.c_int32 * 6)(*[v for row in self.boundary_types for v in row] + [0] * (6 - 2 * dim))
        h = ctypes.c_void_p()
        L.check(lib.pfm_mg_create(dim, ctypes.cast(Narr, ctypes.c_void_p), base_level,
                                  ctypes.cast(bc, ctypes.c_void_p), int(precision), ctypes.byref(h)),
                "pfm_mg_create")
        self._h = h
        self.n_mg_levels = lib.pfm_mg_n_levels(h)
        self.iterations = 0
        self.residual = float("nan")
        self.converged = False

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load(require_gpu=False).pfm_mg_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ---- solid obstacle (extension) ----
    def set_graphs(self, enable: bool):
        """Replay each PCG iteration from a HIP graph, or issue its launches
        directly (the default); the results are the same."""
        L.call("pfm_mg_set_graphs", self._h, int(bool(enable)))

    def set_solid(self, solid, stream=None):
        """Mark solid cells of the finest level (bool/uint8 CUDA tensor of shape N,
        nonzero = solid; None removes the obstacle).  Solid cells carry no
        pressure, their fluid neighbours treat them as walls, and apply_bc /
        Poisson zero the faces that touch them (free slip)."""
        lib = L.load()
        if solid is None:
            self._solid = None
            L.check(lib.pfm_mg_set_solid(self._h, None, L.stream_ptr(stream)), "pfm_mg_set_solid")
            return
        if tuple(solid.shape) != tuple(self.N):
            raise L.PfmError(f"solid mask must have shape {tuple(self.N)}, got {tuple(solid.shape)}")
        m = solid.to(torch.uint8).contiguous() if solid.dtype != torch.uint8 else solid.contiguous()
        self._solid = m  # keep alive until the kernels that read it have run
        L.check(lib.pfm_mg_set_solid(self._h, L.ptr(m, dtype=torch.uint8, name="solid"), L.stream_ptr(stream)),
                "pfm_mg_set_solid")

    # ---- field views (the reference's self.r[l], self.z[l], self.x, ...) ----
    def _level_shape(self, l):
        return tuple(max(1, n >> l) for n in self.N)

    def _field(self, which, level=0):
        p = L.load().pfm_mg_field(self._h, which, level)
        if not p:
            raise L.PfmError("field not available")
        return torch.as_tensor(_DevArray(p, self._level_shape(level)), device="cuda")

    @property
    def r(self):
        return [self._field(0, l) for l in range(self.n_mg_levels)]

    @property
    def z(self):
        return [self._field(1, l) for l in range(self.n_mg_levels)]

    @property
    def x(self):
        return self._field(2)

    @property
    def Ap(self):
        return self._field(4)

    @property
    def p(self):
        """Pressure (get_result target); the reference's dense ``self.p``."""
        return self._field(6)

    @property
    def u_div(self):
        return self._field(0)

    # ---- reference methods ----
    def _s(self, stream):
        return L.stream_ptr(stream)

    def init(self, r, k, stream=None):
        """r0 = k * r; z, Ap, p, x = 0 (3D/mgpcg.py:71-79)."""
        if tuple(r.shape) != tuple(self.N):
            raise L.PfmError(f"rhs shape {tuple(r.shape)} != {tuple(self.N)}")
        L.call("pfm_mg_init", self._h, L.ptr(r, name="r"), float(k), self._s(stream))

    def get_result(self, x, stream=None):
        """x <- solution (3D/mgpcg.py:82-90)."""
        x.copy_(self.x)

    def apply_preconditioner(self, stream=None):
        """z0 = V-cycle(r0) (3D/mgpcg.py:182-200)."""
        L.call("pfm_mg_vcycle", self._h, self._s(stream))

    def smooth(self, l, phase, stream=None):
        L.call("pfm_mg_smooth", self._h, int(l), int(phase), self._s(stream))

    def restrict(self, l, stream=None):
        L.call("pfm_mg_restrict", self._h, int(l), self._s(stream))

    def prolongate(self, l, stream=None):
        L.call("pfm_mg_prolongate", self._h, int(l), self._s(stream))

    def get_Ap(self, p, Ap, stream=None):
        """Ap = A p on the finest level (3D/mgpcg.py:124-129)."""
        L.call("pfm_mg_apply_A", self._h, L.ptr(p, name="p"), L.ptr(Ap, name="Ap"), self._s(stream))

    def solve(self, max_iters=-1, eps=1e-12, tol=1e-12, verbose=False, *, tol_mode="abs", stream=None):
        """PCG (3D/mgpcg.py:202-280); returns the iteration count."""
        it = ctypes.c_int()
        rtr = ctypes.c_double()
        t0 = time.time()
        rc = L.call("pfm_mg_solve", self._h, int(max_iters), float(eps), float(tol),
                    0 if tol_mode == "abs" else 1, ctypes.byref(it), ctypes.byref(rtr), self._s(stream))
        self.iterations, self.residual = it.value, math.sqrt(max(rtr.value, 0.0))
        self.converged = rc == 0
        if verbose:
            state = "Converged" if self.converged else "Return without converging"
            print(f"[MGPCG] {state} at iter: {it.value} with final error: {self.residual} "
                  f"using time: {time.time() - t0}")
        return it.value

    def _u(self, u):
        if len(u) != self.dim:
            raise L.PfmError(f"expected {self.dim} face fields")
        out = []
        for a, t in enumerate(u):
            if tuple(t.shape) != face_shape(self.N, a):
                raise L.PfmError(f"face field {a} has shape {tuple(t.shape)}, expected {face_shape(self.N, a)}")
            out.append(L.ptr(t, name=f"u[{a}]"))
        return out + [None] * (3 - self.dim)

    def apply_bc(self, *u, stream=None):
        L.call("pfm_mg_apply_bc", self._h, *self._u(u), self._s(stream))

    def divergence(self, *u, stream=None):
        """u_div = div u (3D/mgpcg.py:380-390); held in r0 until init."""
        L.call("pfm_mg_divergence_init", self._h, *self._u(u), 1.0, self._s(stream))

    def subtract_grad_p(self, *u, stream=None):
        L.call("pfm_mg_subtract_grad_p", self._h, *self._u(u), self._s(stream))

    def solve_pressure_MGPCG(self, verbose=False, *, max_iters=400, tol=1e-12, tol_mode="abs", stream=None):
        """init(u_div, -1); solve(400, tol=1e-12); get_result (3D/mgpcg.py:420-423)."""
        self.init(self.u_div, -1, stream=stream)
        return self.solve(max_iters=max_iters, tol=tol, verbose=verbose, tol_mode=tol_mode, stream=stream)

    def Poisson(self, *u, verbose=False, max_iters=400, tol=1e-12, tol_mode="abs", stream=None):
        """Pressure projection of u in place (3D/mgpcg.py:425-430); returns iterations."""
        it = ctypes.c_int()
        rtr = ctypes.c_double()
        t0 = time.time()
        rc = L.call("pfm_mg_poisson", self._h, *self._u(u), int(max_iters), float(tol),
                    0 if tol_mode == "abs" else 1, ctypes.byref(it), ctypes.byref(rtr), self._s(stream))
        self.iterations, self.residual = it.value, math.sqrt(max(rtr.value, 0.0))
        self.converged = rc == 0
        if verbose:
            state = "Converged" if self.converged else "Return without converging"
            print(f"[MGPCG] {state} at iter: {it.value} with final error: {self.residual} "
                  f"using time: {time.time() - t0}")
        return it.value


class MGPCG_2(MGPCG):
    """2D solver (2D/mgpcg.py:283-366)."""

    def __init__(self, boundary_types, N, base_level=3, real=float, **kw):
        super().__init__(boundary_types, N, dim=2, base_level=base_level, real=real, **kw)


class MGPCG_3(MGPCG):
    """3D solver (3D/mgpcg.py:350-430)."""

    def __init__(self, boundary_types, N, base_level=3, real=float, **kw):
        super().__init__(boundary_types, N, dim=3, base_level=base_level, real=real, **kw)
