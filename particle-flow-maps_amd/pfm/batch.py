"""Batched 2D scenes: many independent 2D scenes of different sizes advanced
together by the library (pfm_batch2d_* in include/pfm.h).

Each scene follows the reference 2D main loop (2D/run_2D.py:772-909) with its
own dt (its own max speed, CFL and frame clip) and its own PCG iteration
count; every phase of the step is one persistent launch over all scenes, and
the per-phase launch sequences are replayed from HIP graphs.  Fields live in
concatenated torch tensors (scene order); ``u_view`` / ``particle_view`` give
one scene's part in the shapes of the single-scene ``Simulation``.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from . import core
from .simulation import StepConfig

# kernel kinds of pfm_batch2d_tuning, in order (then the coarse kernel's block size); "rk4_grid" is the
# fused midpoint launch (identity, RK4_grid and advect_u per face) and "advect_u" is no longer launched
# (its slot keeps the index layout)
KINDS = ("rk4", "g2p", "seed", "rk4_grid", "advect_u", "faces", "p2g_tile", "bin", "cells", "mg_down", "mg_up")


class Batch2D:
    """Many 2D scenes; ``res`` lists (res_x, res_y) per scene, dx = 1/res_y."""

    def __init__(self, res: Sequence[Tuple[int, int]], cfg: Optional[StepConfig] = None,
                 seeds: Optional[Sequence[int]] = None, device="cuda", graphs: bool = True):
        self.lib = L.load()
        self.res = [(int(a), int(b)) for a, b in res]
        self.S = len(self.res)
        if self.S < 1:
            raise L.PfmError("a batch needs at least one scene")
        self.cfg = cfg or StepConfig()
        c = self.cfg
        self.ppc_axis = int(math.isqrt(c.particles_per_cell))
        if self.ppc_axis ** 2 != c.particles_per_cell:
            raise L.PfmError(f"particles_per_cell={c.particles_per_cell} is not a square")
        self.seeds = [int(s) for s in (seeds if seeds is not None else [c.seed] * self.S)]
        if len(self.seeds) != self.S:
            raise L.PfmError("one seed per scene")
        self._res_np = np.ascontiguousarray(np.asarray(self.res, np.int32).reshape(-1))
        self._seed_np = np.ascontiguousarray(np.asarray(self.seeds, np.int64))
        tot = (ctypes.c_int64 * 4)()
        L.check(self.lib.pfm_batch2d_sizes(self.S, self._res_np.ctypes.data, self.ppc_axis, tot),
                "pfm_batch2d_sizes")
        self.P, self.F = int(tot[0]), (int(tot[1]), int(tot[2]))
        self.C = int(tot[3])
        # per-scene offsets (the layout the library derives from the same sizes)
        self.n = [a * b * self.ppc_axis ** 2 for a, b in self.res]
        self.nf = [((a + 1) * b, a * (b + 1)) for a, b in self.res]
        self.p_off = np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64).tolist()
        self.f_off = [np.concatenate([[0], np.cumsum([f[a] for f in self.nf])[:-1]]).astype(np.int64).tolist()
                      for a in range(2)]
        self.c_off = np.concatenate([[0], np.cumsum([a * b for a, b in self.res])[:-1]]).astype(np.int64).tolist()
        dev = torch.device(device)
        self.device = dev
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        P, F = self.P, self.F

        def pv():
            return torch.zeros((2, P), **f32)

        self.u = [torch.zeros(F[a], **f32) for a in range(2)]
        self.p2g_weight = [torch.zeros(F[a], **f32) for a in range(2)]
        self.psi_grid = [torch.zeros((2, F[a]), **f32) for a in range(2)]
        self.T_grid = [torch.zeros((2, F[a]), **f32) for a in range(2)]
        self.pos = pv()
        self.T = [pv(), pv()]
        self.T_init = [pv(), pv()]
        self.T_grad_m = [pv(), pv()]
        self.Cm = [pv(), pv()]
        self.init_C = [pv(), pv()]
        self.imp = pv()
        self.init_imp = pv()
        self.imp_grad_m = pv() if c.write_imp_grad_m else None
        self.cell_of = torch.zeros(P, **i32)
        self.cell_start = torch.zeros(self.C + 1, **i32)
        self.perm = torch.zeros(P, **i32)
        D = L.PfmBatch2dDesc()
        D.n_scenes = self.S
        D.res = self._res_np.ctypes.data
        D.seed = self._seed_np.ctypes.data
        D.ppc_axis, D.particle_order = self.ppc_axis, int(c.particle_order)
        D.reinit_every, D.reinit_every_grad_m = int(c.reinit_every), int(c.reinit_every_grad_m)
        D.use_midpoint_vel, D.reinit_particle_pos = int(c.use_midpoint_vel), int(c.reinit_particle_pos)
        D.CFL, D.visualize_dt = float(c.CFL), float(c.visualize_dt)
        D.base_level, D.max_iters = int(c.base_level), int(c.max_iters)
        D.tol_mode = 0 if c.tol_mode == "abs" else 1
        D.tol = float(c.tol)
        for a in range(2):
            D.u[a], D.p2g_weight[a] = self.u[a].data_ptr(), self.p2g_weight[a].data_ptr()
            D.psi_grid[a], D.T_grid[a] = self.psi_grid[a].data_ptr(), self.T_grid[a].data_ptr()
            D.T[a], D.T_init[a], D.T_grad_m[a] = (self.T[a].data_ptr(), self.T_init[a].data_ptr(),
                                                 self.T_grad_m[a].data_ptr())
            D.C[a], D.init_C[a] = self.Cm[a].data_ptr(), self.init_C[a].data_ptr()
        D.pos, D.imp, D.init_imp = self.pos.data_ptr(), self.imp.data_ptr(), self.init_imp.data_ptr()
        D.imp_grad_m = self.imp_grad_m.data_ptr() if self.imp_grad_m is not None else None
        D.cell_of, D.cell_start, D.perm = self.cell_of.data_ptr(), self.cell_start.data_ptr(), self.perm.data_ptr()
        self._desc = D
        h = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        L.check(self.lib.pfm_batch2d_create(ctypes.byref(D), ctypes.byref(h), ctypes.byref(nbytes)),
                "pfm_batch2d_create")
        self._h = h
        self.library_bytes = int(nbytes.value)
        self.state = L.PfmBatch2dState()
        self.state.i = -1
        if not graphs:
            self.set_graphs(False)

    # ---------------------------------------------------------------- views
    def dx(self, s: int) -> float:
        return 1.0 / self.res[s][1]

    def u_view(self, s: int, a: int) -> torch.Tensor:
        """Scene s's face velocity of axis a, shape core.face_shape(res_s, a)."""
        o, n = self.f_off[a][s], self.nf[s][a]
        return self.u[a][o:o + n].view(core.face_shape(self.res[s], a))

    def particle_view(self, s: int, t: torch.Tensor) -> torch.Tensor:
        """Scene s's part of a (2, P) particle field (a strided view)."""
        return t[:, self.p_off[s]:self.p_off[s] + self.n[s]]

    # ------------------------------------------------------------------- IC
    def set_point_vortices(self, s: int, centers, weights, rad: float, strength: float = -0.01):
        """Point-vortex velocity of scene s (leapfrog_vel_func, 2D/init_conditions.py:29-60, for any
        vortex list) averaged to its faces; project() afterwards."""
        res = self.res[s]
        vc = torch.zeros((2,) + res, dtype=torch.float32, device=self.device)
        core.point_vortices(res, self.dx(s), centers, weights, vc, rad, strength)
        core.split_central(res, vc, [self.u_view(s, 0), self.u_view(s, 1)])

    def project(self):
        L.check(self.lib.pfm_batch2d_project(self._h, L.stream_ptr()), "pfm_batch2d_project")

    def init_particles(self):
        """initialize_2d_particles (2D/run_2D.py:731-738) for every scene."""
        L.check(self.lib.pfm_batch2d_init_particles(self._h, L.stream_ptr()), "pfm_batch2d_init_particles")
        core._mark_written(self.pos)

    # ----------------------------------------------------------------- step
    def step(self):
        """One step of every scene; returns the lockstep PCG iteration counts."""
        st = self.state
        L.check(self.lib.pfm_batch2d_step(self._h, ctypes.byref(st), L.stream_ptr()), "pfm_batch2d_step")
        core._mark_written(self.pos)
        return list(st.lockstep_iters[:st.n_solves])

    def scene_info(self):
        """Per-scene dt, sub_t, frame and the (iterations, converged) of the
        last step's projections."""
        S = self.S
        dt = np.zeros(S, np.float64)
        sub_t = np.zeros(S, np.float64)
        frame = np.zeros(S, np.int32)
        iters = np.zeros(2 * S, np.int32)
        conv = np.zeros(2 * S, np.int32)
        L.check(self.lib.pfm_batch2d_scene_info(self._h, dt.ctypes.data, sub_t.ctypes.data, frame.ctypes.data,
                                                iters.ctypes.data, conv.ctypes.data), "pfm_batch2d_scene_info")
        return {"dt": dt, "sub_t": sub_t, "frame": frame, "iters": iters.reshape(S, 2),
                "converged": conv.reshape(S, 2)}

    # ---------------------------------------------------------- launch shapes
    def tune(self):
        """Autotune the launch shapes on the current velocity, then restore it
        and the schedule (the next step is step 0 again)."""
        saved = [t.clone() for t in self.u]
        st = L.PfmBatch2dState()
        st.i = -1
        L.check(self.lib.pfm_batch2d_tune(self._h, ctypes.byref(st), L.stream_ptr()), "pfm_batch2d_tune")
        for a in range(2):
            self.u[a].copy_(saved[a])
        self.state = L.PfmBatch2dState()
        self.state.i = -1
        return self.tuning()

    def tuning(self):
        vals = (ctypes.c_int32 * (len(KINDS) + 1))()
        n = L.check(self.lib.pfm_batch2d_tuning(self._h, vals, len(KINDS) + 1), "pfm_batch2d_tuning")
        out = {k: int(vals[i]) for i, k in enumerate(KINDS[:n])}
        if n > len(KINDS):
            out["coarse_threads"] = int(vals[len(KINDS)])
        return out

    def set_enabled(self, s: int, enabled: bool):
        """Step scene s (True) or leave its fields as they are (False)."""
        L.check(self.lib.pfm_batch2d_set_enabled(self._h, int(s), int(bool(enabled))), "pfm_batch2d_set_enabled")

    def set_graphs(self, enable: bool):
        L.check(self.lib.pfm_batch2d_set_graphs(self._h, int(bool(enable))), "pfm_batch2d_set_graphs")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.pfm_batch2d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_vortex_scenes(n_scenes: int, seed: int = 2, lo: int = 64, hi: int = 512, vortices=(8, 32),
                         core_cells: float = 1.5):
    """Scenes of the batched-step workload: square grids with log-uniform
    sizes in [lo, hi], each with 8-32 Gaussian point vortices at uniform
    positions in [0.2, 0.8]^2, circulations N(0, 1) and core radius
    core_cells * dx.  Returns [(res, centers, weights, rad)] (numpy, seeded)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n_scenes))).astype(np.int64)
    out = []
    for N in sizes:
        k = int(rng.integers(vortices[0], vortices[1] + 1))
        centers = rng.uniform(0.2, 0.8, (k, 2))
        weights = rng.normal(0.0, 1.0, k)
        out.append(((int(N), int(N)), centers, weights, core_cells / float(N)))
    return out
