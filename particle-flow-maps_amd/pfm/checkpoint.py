"""Checkpoint and resume.

* Velocity checkpoints use the reference's files: ``np.save`` of the fp32 face
  fields as ``vel_x_numpy_<frame>.npy`` (``vel_y``, ``vel_z``), written at
  3D/run_3D.py:699-701,846-849 and read back at :676-680 (2D/run_2D.py:726-728,
  753-755, 956-958).  A face tensor here has exactly the reference's numpy
  shape and order, so the files are interchangeable.
* Reference-style restart (``resume_from_frame``): velocity (and smoke) from a
  frame's files, flow maps restarted, as the reference does with
  ``from_frame > 0``.
* Full state (absent in the reference, which can only restart from the
  velocity): every array a step reads -- velocities, particle positions, the
  flow maps, impulses and C -- one ``.npy`` each, plus ``state.json`` with the
  grid, the step knobs and the schedule counters.  Resuming reproduces the
  uninterrupted run bit for bit.  The MGPCG search direction is re-zeroed by
  every solve, so the solver needs no state beyond its knobs.

Files are read with ``numpy.load(allow_pickle=False)`` only.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Sequence

import numpy as np
import torch

FORMAT = "pfm-state-1"
_AXES = "xyz"


def velocity_files(ckptdir: str, frame: int, dim: int):
    return [os.path.join(ckptdir, f"vel_{_AXES[a]}_numpy_{frame}.npy") for a in range(dim)]


def save_velocity(ckptdir: str, frame: int, u: Sequence[torch.Tensor]) -> None:
    """np.save(os.path.join(ckptdir, "vel_x_numpy_" + str(frame)), u_x.to_numpy()) for each axis."""
    os.makedirs(ckptdir, exist_ok=True)
    for path, t in zip(velocity_files(ckptdir, frame, len(u)), u):
        np.save(path, t.detach().to("cpu", torch.float32).numpy())


def load_velocity(ckptdir: str, frame: int, u: Sequence[torch.Tensor]) -> None:
    """u_x.from_numpy(np.load(... "vel_x_numpy_" + str(frame) + ".npy")) for each axis, in place."""
    for path, t in zip(velocity_files(ckptdir, frame, len(u)), u):
        a = np.load(path, allow_pickle=False)
        if a.dtype != np.float32 or tuple(a.shape) != tuple(t.shape):
            raise ValueError(f"{path}: expected float32 {tuple(t.shape)}, got {a.dtype} {a.shape}")
        t.copy_(torch.from_numpy(a))


def smoke_file(ckptdir: str, frame: int) -> str:
    return os.path.join(ckptdir, f"smoke_numpy_{frame}.npy")


def save_smoke(ckptdir: str, frame: int, smoke: torch.Tensor) -> None:
    """np.save of the cell smoke in the reference's layout: a 4-vector field's
    to_numpy() is (*res, 4); the device field here is (4, *res)."""
    os.makedirs(ckptdir, exist_ok=True)
    s = smoke.detach().to("cpu", torch.float32)
    np.save(smoke_file(ckptdir, frame), np.ascontiguousarray(np.moveaxis(s.numpy(), 0, -1)))


def load_smoke(ckptdir: str, frame: int, smoke: torch.Tensor) -> None:
    """smoke.from_numpy(np.load(... "smoke_numpy_" + str(frame) + ".npy"))
    (3D/run_3D.py:680, 2D/run_2D.py:728) into a (4, *res) field, in place."""
    path = smoke_file(ckptdir, frame)
    a = np.load(path, allow_pickle=False)
    want = tuple(smoke.shape[1:]) + (smoke.shape[0],)
    if a.dtype != np.float32 or tuple(a.shape) != want:
        raise ValueError(f"{path}: expected float32 {want}, got {a.dtype} {a.shape}")
    smoke.copy_(torch.from_numpy(np.ascontiguousarray(np.moveaxis(a, -1, 0))))


def resume_from_frame(sim, ckptdir: str, from_frame: int, smoke=None) -> None:
    """The reference's restart with ``from_frame > 0`` (3D/run_3D.py:676-680,
    708-713; 2D/run_2D.py:726-738): the face velocities (and, when given, the
    smoke) are read from the frame's files, not projected again, and the flow
    maps restart: the schedule counters start over (i = -1, ik = 0, sub_t =
    0, frame_idx = from_frame), so the first step reseeds the particles and
    reinitialises every map; a 2D scene also runs its particle set-up
    (initialize_2d_particles) as the 2D driver does before its loop."""
    if from_frame <= 0:
        raise ValueError("from_frame must be positive (frame 0 starts from the initial condition)")
    load_velocity(ckptdir, from_frame, sim.u)
    if smoke is not None:
        load_smoke(ckptdir, from_frame, smoke)
    sim.i, sim.ik, sim.sub_t, sim.frame_idx, sim.last_dt = -1, 0, 0.0, int(from_frame), 0.0
    if sim.dim == 2:
        sim.initialize_2d_particles()


def _state_arrays(sim):
    """name -> tensor for every array the step reads."""
    out = {}
    for a, t in enumerate(sim.u):
        out[f"u_{_AXES[a]}"] = t
    out["pos"] = sim.pos
    out["init_imp"] = sim.init_imp
    if sim.imp_grad_m is not None:
        out["imp_grad_m"] = sim.imp_grad_m
    for name in ("T_init", "T_grad_m", "init_C") + (("T", "C") if sim.dim == 2 else ()):
        for a, t in enumerate(getattr(sim, name)):
            out[f"{name}_{_AXES[a]}"] = t
    if sim.dim == 2:
        out["imp"] = sim.imp
    return out


def save_state(sim, path: str) -> None:
    """Write the full state of a Simulation into directory `path`."""
    os.makedirs(path, exist_ok=True)
    arrays = _state_arrays(sim)
    for name, t in arrays.items():
        np.save(os.path.join(path, name + ".npy"), t.detach().to("cpu").numpy())
    meta = {
        "format": FORMAT,
        "res": list(sim.res),
        "cfg": dataclasses.asdict(sim.cfg),
        "schedule": {"i": sim.i, "ik": sim.ik, "sub_t": sim.sub_t, "frame_idx": sim.frame_idx,
                     "last_dt": sim.last_dt},
        "arrays": sorted(arrays),
    }
    with open(os.path.join(path, "state.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


def load_state(path: str, device="cuda", boundary_types=None):
    """A Simulation restored from `path` (written by save_state)."""
    from .simulation import Simulation, StepConfig

    with open(os.path.join(path, "state.json")) as fh:
        meta = json.load(fh)
    if meta.get("format") != FORMAT:
        raise ValueError(f"{path}: unknown state format {meta.get('format')!r}")
    sim = Simulation(tuple(meta["res"]), StepConfig(**meta["cfg"]), device=device, boundary_types=boundary_types)
    arrays = _state_arrays(sim)
    if sorted(arrays) != sorted(meta["arrays"]):
        raise ValueError(f"{path}: arrays {meta['arrays']} do not match the configuration's {sorted(arrays)}")
    for name, t in arrays.items():
        a = np.load(os.path.join(path, name + ".npy"), allow_pickle=False)
        if a.dtype != np.float32 or tuple(a.shape) != tuple(t.shape):
            raise ValueError(f"{name}: expected float32 {tuple(t.shape)}, got {a.dtype} {a.shape}")
        t.copy_(torch.from_numpy(a))
    sch = meta["schedule"]
    sim.i, sim.ik = int(sch["i"]), int(sch["ik"])
    sim.sub_t, sim.frame_idx, sim.last_dt = float(sch["sub_t"]), int(sch["frame_idx"]), float(sch["last_dt"])
    return sim
