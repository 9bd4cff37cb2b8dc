"""ctypes binding of libpfm.so (the HIP C ABI declared in include/pfm.h).

The library is built in-tree by ``particle-flow-maps_amd/build.py``.  There is
no CPU fallback: if the library is missing or no GPU is visible, every entry
point raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfm.so")

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_f = ctypes.c_float
c_d = ctypes.c_double
c_vp = ctypes.c_void_p
c_sz = ctypes.c_size_t


class PfmGrid(ctypes.Structure):
    """pfm_grid of include/pfm.h."""

    _fields_ = [("dim", ctypes.c_int32), ("res", ctypes.c_int32 * 3), ("dx", ctypes.c_float)]


class PfmSceneDesc(ctypes.Structure):
    """pfm_scene_desc of include/pfm.h."""

    _fields_ = [
        ("dim", ctypes.c_int32), ("res", ctypes.c_int32 * 3),
        ("reinit_every", ctypes.c_int32), ("reinit_every_grad_m", ctypes.c_int32),
        ("use_midpoint_vel", ctypes.c_int32), ("reinit_particle_pos", ctypes.c_int32),
        ("CFL", ctypes.c_double), ("visualize_dt", ctypes.c_double),
        ("ppc_axis", ctypes.c_int32), ("particle_order", ctypes.c_int32), ("seed", ctypes.c_int64),
        ("max_iters", ctypes.c_int32), ("tol_mode", ctypes.c_int32), ("tol", ctypes.c_double),
        ("fused_p2g", ctypes.c_int32), ("fold_T", ctypes.c_int32), ("n", ctypes.c_int64),
        ("u", c_vp * 3), ("p2g_weight", c_vp * 3), ("psi_grid", c_vp * 3), ("T_grid", c_vp * 3),
        ("pos", c_vp), ("init_imp", c_vp), ("imp_grad_m", c_vp),
        ("T_init", c_vp * 3), ("T_grad_m", c_vp * 3), ("init_C", c_vp * 3),
        ("imp", c_vp), ("T", c_vp * 3), ("C", c_vp * 3),
        ("cell_of", c_vp), ("cell_start", c_vp), ("perm", c_vp), ("workspace", c_vp), ("ws_bytes", c_sz),
        ("max_speed", c_vp), ("mg", c_vp),
    ]


class PfmSceneState(ctypes.Structure):
    """pfm_scene_state of include/pfm.h."""

    _fields_ = [
        ("i", ctypes.c_int64), ("ik", ctypes.c_int64), ("frame_idx", ctypes.c_int64),
        ("sub_t", ctypes.c_double), ("last_dt", ctypes.c_double),
        ("n_solves", ctypes.c_int32), ("iters", ctypes.c_int32 * 2), ("converged", ctypes.c_int32 * 2),
    ]


class PfmBatch2dDesc(ctypes.Structure):
    """pfm_batch2d_desc of include/pfm.h."""

    _fields_ = [
        ("n_scenes", ctypes.c_int32), ("res", c_vp), ("seed", c_vp),
        ("ppc_axis", ctypes.c_int32), ("particle_order", ctypes.c_int32),
        ("reinit_every", ctypes.c_int32), ("reinit_every_grad_m", ctypes.c_int32),
        ("use_midpoint_vel", ctypes.c_int32), ("reinit_particle_pos", ctypes.c_int32),
        ("CFL", ctypes.c_double), ("visualize_dt", ctypes.c_double),
        ("base_level", ctypes.c_int32), ("max_iters", ctypes.c_int32), ("tol_mode", ctypes.c_int32),
        ("tol", ctypes.c_double),
        ("u", c_vp * 2), ("p2g_weight", c_vp * 2), ("psi_grid", c_vp * 2), ("T_grid", c_vp * 2),
        ("pos", c_vp), ("T", c_vp * 2), ("T_init", c_vp * 2), ("T_grad_m", c_vp * 2), ("C", c_vp * 2),
        ("init_C", c_vp * 2), ("imp", c_vp), ("init_imp", c_vp), ("imp_grad_m", c_vp),
        ("cell_of", c_vp), ("cell_start", c_vp), ("perm", c_vp),
    ]


class PfmBatch2dState(ctypes.Structure):
    """pfm_batch2d_state of include/pfm.h."""

    _fields_ = [("i", ctypes.c_int64), ("ik", ctypes.c_int64), ("n_solves", ctypes.c_int32),
                ("lockstep_iters", ctypes.c_int32 * 2)]


class PfmError(RuntimeError):
    pass


# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "pfm_abi_version": [],
    "pfm_last_error": [],
    "pfm_rk4_particles": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_f, c_int, c_vp],
    "pfm_rk4_grid": [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_f, c_vp],
    "pfm_advect_u": [c_vp] + [c_vp] * 12 + [c_vp],
    "pfm_midpoint_advect": [c_vp] + [c_vp] * 12 + [c_f, c_vp],
    "pfm_reset_identity_grid": [c_vp, c_int, c_vp, c_vp, c_vp],
    "pfm_reset_T_identity": [c_int, c_i64, c_vp, c_vp, c_vp, c_vp],
    "pfm_seed_particles": [c_vp, c_int, c_f, c_i64, c_vp, c_i64, c_int, c_vp],
    "pfm_g2p": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_update_T": [c_int, c_i64] + [c_vp] * 9 + [c_vp],
    "pfm_update_particles_imp": [c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_update_particles_grad_m": [c_int, c_i64] + [c_vp] * 9 + [c_vp],
    "pfm_p2g_workspace_size": [c_vp, c_i64],
    "pfm_bin_particles": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp],
    "pfm_cell_lists": [c_vp, c_vp, c_vp, ctypes.c_int32, c_vp, c_vp, c_vp],
    "pfm_p2g": [c_vp, c_i64] + [c_vp] * 11 + [c_vp, c_vp] + [c_vp] * 6 + [c_vp, c_sz, c_vp],
    "pfm_rk4_particles_p2g_record": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_f] + [c_vp] * 8
                                    + [c_vp],
    "pfm_p2g_records": [c_vp, c_i64, c_vp, c_vp] + [c_vp] * 6 + [c_vp, c_sz, c_vp],
    "pfm_rk4_particles_p2g_record_fold": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_f] + [c_vp] * 7 + [c_vp, c_vp],
    "pfm_rk4_particles_p2g_record_fold_count": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_f] + [c_vp] * 7
                                               + [c_vp, c_vp, c_sz, c_vp],
    "pfm_bin_finish": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp],
    "pfm_p2g_record_buffer": [c_vp, c_i64, c_vp],
    "pfm_max_speed": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_split_central": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_get_central": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_curl": [c_vp, c_vp, c_vp, c_vp],
    "pfm_point_vortices": [c_vp, c_int, c_vp, c_vp, c_vp, c_f, c_f, c_vp, c_vp],
    "pfm_vortex_ring": [c_vp, c_f, c_f, c_f, c_vp, c_vp, c_vp, c_int, c_vp, c_vp],
    "pfm_kinetic_energy": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_kinetic_energy_noalloc": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_add_fields": [c_i64, c_vp, c_vp, c_vp, c_f, c_vp],
    "pfm_advect_smoke": [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp],
    "pfm_vortex_ring_smoke": [c_vp, c_f, c_f, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp],
    "pfm_g2p_smoke": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_p2g_smoke": [c_vp, c_i64, c_vp, c_vp, c_vp] + [c_vp] * 6 + [c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_clamp_smoke_particle": [c_vp, c_i64, c_vp, c_vp, ctypes.c_int32, c_vp, c_vp, c_vp, c_vp],
    "pfm_mg_create": [c_int, c_vp, c_int, c_vp, c_int, ctypes.POINTER(c_vp)],
    "pfm_mg_destroy": [c_vp],
    "pfm_mg_set_graphs": [c_vp, c_int],
    "pfm_mg_n_levels": [c_vp],
    "pfm_mg_field": [c_vp, c_int, c_int],
    "pfm_mg_apply_bc": [c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_mg_set_solid": [c_vp, c_vp, c_vp],
    "pfm_mg_divergence_init": [c_vp, c_vp, c_vp, c_vp, c_f, c_vp],
    "pfm_mg_init": [c_vp, c_vp, c_f, c_vp],
    "pfm_mg_solve": [c_vp, c_int, c_d, c_d, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_d), c_vp],
    "pfm_mg_subtract_grad_p": [c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_mg_poisson": [c_vp, c_vp, c_vp, c_vp, c_int, c_d, c_int, ctypes.POINTER(c_int),
                       ctypes.POINTER(c_d), c_vp],
    "pfm_mg_vcycle": [c_vp, c_vp],
    "pfm_mg_smooth": [c_vp, c_int, c_int, c_vp],
    "pfm_mg_restrict": [c_vp, c_int, c_vp],
    "pfm_mg_prolongate": [c_vp, c_int, c_vp],
    "pfm_mg_apply_A": [c_vp, c_vp, c_vp, c_vp],
    "pfm_scene_create": [ctypes.POINTER(PfmSceneDesc), ctypes.POINTER(c_vp)],
    "pfm_scene_step": [c_vp, ctypes.POINTER(PfmSceneState), c_d, c_vp],
    "pfm_scene_destroy": [c_vp],
    "pfm_batch2d_sizes": [ctypes.c_int32, c_vp, ctypes.c_int32, c_vp],
    "pfm_batch2d_create": [ctypes.POINTER(PfmBatch2dDesc), ctypes.POINTER(c_vp), ctypes.POINTER(c_i64)],
    "pfm_batch2d_project": [c_vp, c_vp],
    "pfm_batch2d_init_particles": [c_vp, c_vp],
    "pfm_batch2d_step": [c_vp, ctypes.POINTER(PfmBatch2dState), c_vp],
    "pfm_batch2d_scene_info": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "pfm_batch2d_tune": [c_vp, ctypes.POINTER(PfmBatch2dState), c_vp],
    "pfm_batch2d_tuning": [c_vp, c_vp, ctypes.c_int32],
    "pfm_batch2d_set_enabled": [c_vp, ctypes.c_int32, ctypes.c_int32],
    "pfm_batch2d_set_graphs": [c_vp, ctypes.c_int32],
    "pfm_batch2d_destroy": [c_vp],
}
_RESTYPES = {
    "pfm_last_error": ctypes.c_char_p,
    "pfm_p2g_workspace_size": c_sz,
    "pfm_mg_field": c_vp,
    "pfm_p2g_record_buffer": c_vp,
}

_lib = None


def exported_symbols():
    return sorted(_SIGS)


def load(require_gpu: bool = True):
    """Load libpfm.so (raises if it was not built or, when required, no GPU)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PfmError(f"libpfm.so not built at {LIB_PATH}; run particle-flow-maps_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, c_int)
        _lib = L
    if require_gpu and not torch.cuda.is_available():
        raise PfmError("libpfm needs a ROCm GPU; torch.cuda.is_available() is False")
    return _lib


def last_error() -> str:
    msg = load(require_gpu=False).pfm_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise PfmError(f"{what} failed ({rc}): {last_error()}")
    return rc


def call(name: str, *args) -> int:
    return check(getattr(load(), name)(*args), name)


def stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def ptr(t, *, dtype=torch.float32, name="tensor", allow_none=False):
    """Device pointer of a contiguous CUDA tensor of the given dtype."""
    if t is None:
        if allow_none:
            return None
        raise PfmError(f"{name} is None")
    if not isinstance(t, torch.Tensor):
        raise PfmError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise PfmError(f"{name} must live on the GPU (got {t.device})")
    if t.dtype != dtype:
        raise PfmError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise PfmError(f"{name} must be contiguous")
    return t.data_ptr()


def make_grid(res, dx) -> PfmGrid:
    g = PfmGrid()
    g.dim = len(res)
    r = list(res) + [1] * (3 - len(res))
    for a in range(3):
        g.res[a] = int(r[a])
    g.dx = float(dx)
    return g
