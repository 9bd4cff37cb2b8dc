"""Field utilities of 2D/taichi_utils.py used on the step path."""
from __future__ import annotations

from .. import core


def copy_to(source, dest):
    """2D/taichi_utils.py:18-21."""
    dest.copy_(source)


def split_central_vector(central, horizontal, vertical):
    """2D/taichi_utils.py:56-65: cell-centred (2, rx, ry) -> faces."""
    core.split_central(tuple(central.shape[1:]), central, (horizontal, vertical))


def get_central_vector(horizontal, vertical, central):
    """2D/taichi_utils.py:50-54."""
    core.get_central(tuple(central.shape[1:]), (horizontal, vertical), central)


def curl(vf, cf, dx):
    """2D/taichi_utils.py:97-103: vf is cell-centred (2, rx, ry), cf (rx, ry)."""
    core.curl(tuple(vf.shape[1:]), dx, vf, cf)


def add_fields(f1, f2, dest, multiplier):
    """2D/taichi_utils.py:13-16: dest = f1 + multiplier * f2 (e.g. the smoke fields of the
    initial conditions)."""
    core.add_fields(f1, f2, dest, multiplier)
