"""Step drivers that follow the reference main loops on the GPU.

``Simulation`` owns every field of one scene (particles, flow maps, face
velocities, the MGPCG hierarchy) resident in device memory and advances it
with the reference schedule:

  * 3D: 3D/run_3D.py:716-831 (long reinit every ``reinit_every`` steps with
    reseeding, short reinit every ``reinit_every_grad_m`` steps, midpoint
    velocity, particle RK4, binning, P2G with the in-kernel T/C composition,
    projection);
  * 2D: 2D/run_2D.py:731-909 (initial G2P before the loop, T/imp/C evolution
    in separate kernels, no binning call).

Host-side arithmetic (dt from the CFL condition, the frame clip) is done in
Python double exactly as the reference does it.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import List, Optional, Sequence

import torch

from . import _lib as L
from . import core
from .mgpcg import MGPCG


@dataclasses.dataclass
class StepConfig:
    """The hyperparameters.py knobs that the step reads (same names)."""

    reinit_every: int = 20
    reinit_every_grad_m: int = 1
    CFL: float = 0.5
    visualize_dt: float = 0.1
    use_midpoint_vel: bool = True
    reinit_particle_pos: bool = True
    particles_per_cell: int = 8
    # MGPCG (3D/mgpcg.py:29-30,420-423) and extensions
    base_level: int = 3
    max_iters: int = 400
    tol: float = 1e-12
    tol_mode: str = "abs"
    # extensions: seeded jitter (seed < 0: the reference lattice) and the
    # particle numbering (0: reference x-fastest lattice, 1: cell-major)
    seed: int = -1
    particle_order: int = 0
    write_imp_grad_m: bool = True
    # MGPCG level storage: 0 fp32 (the reference), 1 fp16, 2 bf16
    mg_precision: int = 0
    # 3D with reinit_every_grad_m == 1 and the fused RK4 + P2G record: the
    # RK4 kernel applies the next step's update_T itself (same arithmetic)
    fold_T: bool = True


def reinit_decision(i: int, ik: int, cfg: StepConfig, dim: int):
    """Reinitialisation decisions of step i: (long, reseed, short, new ik).

    Long reinit when i % reinit_every == 0 (3D/run_3D.py:748, 2D/run_2D.py:804);
    the 3D driver moves ik only when it reseeds (:751-755), the 2D driver at
    every long reinit (2D/run_2D.py:819); short reinit when (i - ik) %
    reinit_every_grad_m == 0 (3D/run_3D.py:767-768)."""
    long_r = i % cfg.reinit_every == 0
    reseed = long_r and (cfg.reinit_particle_pos or i == 0)
    if long_r and (reseed or dim == 2):
        ik = i
    short = (i - ik) % cfg.reinit_every_grad_m == 0
    return long_r, reseed, short, ik


class Simulation:
    """One scene of the particle-flow-map solver, all state in HBM."""

    def __init__(self, res: Sequence[int], cfg: Optional[StepConfig] = None, device="cuda",
                 boundary_types=None, native: bool = True):
        self.res = tuple(int(r) for r in res)
        self.dim = len(self.res)
        self.cfg = cfg or StepConfig()
        d = self.dim
        self.dx = 1.0 / self.res[1]
        ppc = self.cfg.particles_per_cell
        self.ppc_axis = 2 if d == 3 else int(math.isqrt(ppc))
        if self.ppc_axis ** d != ppc:
            raise ValueError(f"particles_per_cell={ppc} is not a {d}-th power")
        self.dist = self.dx / self.ppc_axis
        self.n = math.prod(self.res) * ppc
        dev = torch.device(device)
        self.device = dev
        f32 = dict(dtype=torch.float32, device=dev)
        n = self.n

        def pv():
            return torch.zeros((d, n), **f32)

        self.u = [torch.zeros(core.face_shape(self.res, a), **f32) for a in range(d)]
        self.p2g_weight = [torch.zeros(core.face_shape(self.res, a), **f32) for a in range(d)]
        self.psi_grid = [torch.zeros((d,) + core.face_shape(self.res, a), **f32) for a in range(d)]
        self.T_grid = [torch.zeros((d,) + core.face_shape(self.res, a), **f32) for a in range(d)]
        self.pos = pv()
        self.init_imp = pv()
        self.imp_grad_m = pv() if self.cfg.write_imp_grad_m else None
        self.T_init = [pv() for _ in range(d)]
        self.T_grad_m = [pv() for _ in range(d)]
        self.init_C = [pv() for _ in range(d)]
        if d == 2:
            self.imp = pv()
            self.T = [pv() for _ in range(d)]
            self.C = [pv() for _ in range(d)]
        bt = boundary_types if boundary_types is not None else [[2, 2]] * d
        self.solver = MGPCG(bt, list(self.res), dim=d, base_level=self.cfg.base_level,
                            precision=self.cfg.mg_precision)
        self.binning = core.Binning(self.res, n, dev)
        self.max_speed_buf = torch.empty(1, **f32)
        # 3D: fuse the P2G particle composition into the RK4 kernel
        self.fused_p2g = d == 3
        # schedule state (3D/run_3D.py:708-713)
        self.i = -1
        self.ik = 0
        self.sub_t = 0.0
        self.frame_idx = 0
        self.last_dt = 0.0
        self.poisson_iters: List[int] = []
        # step() runs the whole step in the library (pfm_scene_step) unless
        # native is False, when it issues the same entry points from Python
        self.native = native
        self._scene = None

    # ------------------------------------------------------------------ IC
    def set_velocity_from_central(self, vc: torch.Tensor, project: bool = True):
        """split_central_vector + Poisson (3D/run_3D.py:667-670)."""
        core.split_central(self.res, vc, self.u)
        if project:
            self.project()

    def init_leapfrog_2d(self, centers, weights, rad=0.02, strength=-0.01):
        vc = torch.zeros((2,) + self.res, dtype=torch.float32, device=self.device)
        core.point_vortices(self.res, self.dx, centers, weights, vc, rad, strength)
        self.set_velocity_from_central(vc)

    def init_rings_3d(self, rings, num_samples=2000):
        vc = torch.zeros((3,) + self.res, dtype=torch.float32, device=self.device)
        for r in rings:
            core.vortex_ring(self.res, self.dx, vc, r["w"], r["rad"], r["delta"], r["c"], r["unit_x"],
                             r["unit_y"], num_samples)
        self.set_velocity_from_central(vc)

    def project(self):
        it = self.solver.Poisson(*self.u, max_iters=self.cfg.max_iters, tol=self.cfg.tol,
                                 tol_mode=self.cfg.tol_mode)
        self.poisson_iters.append(it)
        return it

    # ---------------------------------------------------------------- step
    def _seed(self):
        core.seed_particles(self.res, self.ppc_axis, self.dist, self.pos, seed=self.cfg.seed,
                            order=self.cfg.particle_order)

    def compute_dt(self):
        """calc_max_speed + CFL + frame clip (3D/run_3D.py:725-743), host double."""
        core.max_speed(self.res, self.u, out=self.max_speed_buf)
        ms = float(self.max_speed_buf.item())
        curr_dt = self.cfg.CFL * self.dx / ms
        if self.sub_t + curr_dt >= self.cfg.visualize_dt:
            curr_dt = self.cfg.visualize_dt - self.sub_t
            self.sub_t = 0.0
            self.frame_idx += 1
        else:
            self.sub_t += curr_dt
        return curr_dt

    def _midpoint(self, dt):
        """3D/run_3D.py:778-789 (2D/run_2D.py:865-872)."""
        for a in range(self.dim):
            core.reset_identity_grid(self.res, self.dx, self.psi_grid[a], self.T_grid[a])
        for a in range(self.dim):
            core.rk4_grid(self.res, self.dx, self.psi_grid[a], self.T_grid[a], self.u, 0.5 * dt)
        for a in range(self.dim):
            self.p2g_weight[a].copy_(self.u[a])
        core.advect_u(self.res, self.dx, self.p2g_weight, self.u, self.T_grid, self.psi_grid)
        self.project()

    def step(self, dt: Optional[float] = None):
        """One reference step; returns the dt used."""
        if self.native:
            return self._native_step(dt)
        self.i += 1
        i = self.i
        curr_dt = self.compute_dt() if dt is None else dt
        self.last_dt = curr_dt
        if self.dim == 3:
            self._step3(i, curr_dt)
        else:
            self._step2(i, curr_dt)
        return curr_dt

    def _native_step(self, dt):
        if self._scene is None:
            self._scene = _NativeScene(self)
        return self._scene.step(self, dt)

    @property
    def folds_T(self) -> bool:
        """Whether the RK4 kernel applies update_T for the next step (3D, fused
        RK4 + P2G record, a short reinitialisation every step)."""
        return bool(self.dim == 3 and self.fused_p2g and self.cfg.fold_T and self.cfg.reinit_every_grad_m == 1)

    def _step3(self, i, dt):
        cfg = self.cfg
        fold = self.folds_T
        long_r, reseed, short, self.ik = reinit_decision(i, self.ik, cfg, 3)
        if long_r:
            if reseed:
                self._seed()
            core.g2p(self.res, self.dx, self.pos, self.u, self.init_imp, None)
            core.reset_T_identity(self.T_init)
            core.reset_T_identity(self.T_grad_m)
        if short:
            core.g2p(self.res, self.dx, self.pos, self.u,
                     self.imp_grad_m if self.imp_grad_m is not None else None, self.init_C)
            if not fold:  # else the previous RK4 applied it and left T_grad_m = I
                core.update_T(self.T_grad_m, self.T_init, self.T_init)
                core.reset_T_identity(self.T_grad_m)
        if cfg.use_midpoint_vel:
            self._midpoint(dt)
        if self.fused_p2g:
            self.binning.rk4_then_p2g(self.dx, self.pos, self.T_grad_m, self.u, dt, self.T_init, self.init_C,
                                      self.init_imp, self.u, W=self.p2g_weight, fold=fold)
        else:
            core.rk4_particles(self.res, self.dx, self.pos, self.T_grad_m, self.u, dt, True)
            self.binning.bin(self.pos, self.dx, force=True)
            self.binning.p2g(self.dx, self.pos, self.init_imp, self.init_C, self.u, W=self.p2g_weight,
                             Tgm=self.T_grad_m, Tin=self.T_init)
        self.project()

    def initialize_2d_particles(self):
        """Particle set-up before the 2D loop (2D/run_2D.py:731-738)."""
        self._seed()
        core.g2p(self.res, self.dx, self.pos, self.u, self.imp_grad_m, self.init_C)
        core.reset_T_identity(self.T_grad_m)
        core.g2p(self.res, self.dx, self.pos, self.u, self.imp, None)
        self.init_imp.copy_(self.imp)

    def _step2(self, i, dt):
        cfg = self.cfg
        long_r, reseed, short, self.ik = reinit_decision(i, self.ik, cfg, 2)
        if long_r:
            if reseed:
                self._seed()
            core.g2p(self.res, self.dx, self.pos, self.u, self.imp, None)
            self.init_imp.copy_(self.imp)
            core.reset_T_identity(self.T)
            core.reset_T_identity(self.T_init)
        if short:
            core.g2p(self.res, self.dx, self.pos, self.u, self.imp_grad_m, self.init_C)
            core.reset_T_identity(self.T_grad_m)
            for a in range(2):
                self.T_init[a].copy_(self.T[a])
        if cfg.use_midpoint_vel:
            self._midpoint(dt)
        core.rk4_particles(self.res, self.dx, self.pos, self.T_grad_m, self.u, dt, True)
        core.update_particles_grad_m(self.C, self.init_C, self.T_grad_m)
        core.update_T(self.T_grad_m, self.T_init, self.T)
        core.update_particles_imp(self.T, self.init_imp, self.imp)
        self.binning.bin(self.pos, self.dx, force=True)
        self.binning.p2g(self.dx, self.pos, self.imp, self.C, self.u, W=self.p2g_weight)
        self.project()

    # ----------------------------------------------------------- diagnostics
    def kinetic_energy(self) -> float:
        return float(core.kinetic_energy(self.res, self.u).item())


class _NativeScene:
    """pfm_scene_create over a Simulation's fields (captured once: every
    Simulation method updates its fields in place)."""

    def __init__(self, sim: "Simulation"):
        lib = L.load()
        cfg, d = sim.cfg, sim.dim
        D = L.PfmSceneDesc()
        D.dim = d
        for a in range(3):
            D.res[a] = sim.res[a] if a < d else 1
        D.reinit_every, D.reinit_every_grad_m = cfg.reinit_every, cfg.reinit_every_grad_m
        D.use_midpoint_vel, D.reinit_particle_pos = int(cfg.use_midpoint_vel), int(cfg.reinit_particle_pos)
        D.CFL, D.visualize_dt = float(cfg.CFL), float(cfg.visualize_dt)
        D.ppc_axis, D.particle_order, D.seed = sim.ppc_axis, int(cfg.particle_order), int(cfg.seed)
        D.max_iters, D.tol = int(cfg.max_iters), float(cfg.tol)
        D.tol_mode = 0 if cfg.tol_mode == "abs" else 1
        D.fused_p2g = int(sim.fused_p2g)
        D.fold_T = int(sim.folds_T)
        D.n = sim.n

        def p(t):
            return None if t is None else L.ptr(t)

        for a in range(d):
            D.u[a], D.p2g_weight[a] = p(sim.u[a]), p(sim.p2g_weight[a])
            D.psi_grid[a], D.T_grid[a] = p(sim.psi_grid[a]), p(sim.T_grid[a])
            D.T_init[a], D.T_grad_m[a], D.init_C[a] = p(sim.T_init[a]), p(sim.T_grad_m[a]), p(sim.init_C[a])
            if d == 2:
                D.T[a], D.C[a] = p(sim.T[a]), p(sim.C[a])
        D.pos, D.init_imp, D.imp_grad_m = p(sim.pos), p(sim.init_imp), p(sim.imp_grad_m)
        if d == 2:
            D.imp = p(sim.imp)
        b = sim.binning
        D.cell_of, D.cell_start, D.perm = (L.ptr(b.cell_of, dtype=torch.int32), L.ptr(b.cell_start, dtype=torch.int32),
                                           L.ptr(b.perm, dtype=torch.int32))
        D.workspace, D.ws_bytes = b.workspace.data_ptr(), b.workspace.numel()
        D.max_speed = p(sim.max_speed_buf)
        D.mg = sim.solver._h.value
        h = ctypes.c_void_p()
        L.check(lib.pfm_scene_create(ctypes.byref(D), ctypes.byref(h)), "pfm_scene_create")
        self._lib = lib
        self._h = h
        self.desc = D
        self.state = L.PfmSceneState()

    def step(self, sim: "Simulation", dt):
        st = self.state
        st.i, st.ik, st.frame_idx = sim.i, sim.ik, sim.frame_idx
        st.sub_t, st.last_dt = sim.sub_t, sim.last_dt
        rc = self._lib.pfm_scene_step(self._h, ctypes.byref(st), -1.0 if dt is None else float(dt),
                                      L.stream_ptr())
        L.check(rc, "pfm_scene_step")
        sim.i, sim.ik, sim.frame_idx = st.i, st.ik, st.frame_idx
        sim.sub_t, sim.last_dt = st.sub_t, st.last_dt
        k = min(int(st.n_solves), 2)
        sim.poisson_iters.extend(st.iters[:k])
        sv = sim.solver
        if k:
            sv.iterations, sv.converged = st.iters[k - 1], all(st.converged[:k])
        core._mark_written(sim.pos)
        return st.last_dt

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.pfm_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass
