"""Field utilities of 3D/taichi_utils.py used on the step path."""
from __future__ import annotations

from .. import core


def copy_to(source, dest):
    """3D/taichi_utils.py:8-11."""
    dest.copy_(source)


def split_central_vector(vc, vx, vy, vz):
    """3D/taichi_utils.py:87-100: cell-centred (3, rx, ry, rz) -> faces."""
    core.split_central(tuple(vc.shape[1:]), vc, (vx, vy, vz))


def get_central_vector(vx, vy, vz, vc):
    """3D/taichi_utils.py:80-85: faces -> cell-centred (3, rx, ry, rz)."""
    core.get_central(tuple(vc.shape[1:]), (vx, vy, vz), vc)


def curl(vf, cf, dx):
    """3D/taichi_utils.py:56-78: vf, cf are cell-centred (3, rx, ry, rz)."""
    core.curl(tuple(vf.shape[1:]), dx, vf, cf)


def add_fields(f1, f2, dest, multiplier):
    """3D/taichi_utils.py:18-21: dest = f1 + multiplier * f2 (e.g. the smoke fields of the
    initial conditions)."""
    core.add_fields(f1, f2, dest, multiplier)
