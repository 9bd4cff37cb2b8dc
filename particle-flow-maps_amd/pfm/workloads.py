"""Synthetic inputs of the benchmark configurations (set-up only, not the
step path): seeded Fourier velocity fields and the vortex initial conditions."""
from __future__ import annotations

import math

import torch

from . import core


def fourier_velocity(res, n_modes=64, kmin=1.0, kmax=8.0, seed=0, device="cuda"):
    """Face velocities u_a = sum_m A_{m,a}/|k_m|^2 sin(2 pi k_m . x + phi_m) with
    k_m uniform in [kmin, kmax]^d and A ~ N(0, 1) (the mixed-precision Poisson
    configuration's right-hand side is the divergence of this field)."""
    d = len(res)
    g = torch.Generator(device="cpu").manual_seed(seed)
    k = kmin + (kmax - kmin) * torch.rand((n_modes, d), generator=g, dtype=torch.float64)
    amp = torch.randn((n_modes, d), generator=g, dtype=torch.float64) / (k * k).sum(1, keepdim=True)
    phi = 2 * math.pi * torch.rand((n_modes,), generator=g, dtype=torch.float64)
    dx = 1.0 / res[1]
    out = []
    for a in range(d):
        shp = core.face_shape(res, a)
        axes = []
        for c in range(d):
            idx = torch.arange(shp[c], dtype=torch.float64, device=device)
            axes.append((idx if c == a else idx + 0.5) * dx)
        u = torch.zeros(shp, dtype=torch.float64, device=device)
        for m in range(n_modes):
            arg = phi[m].item()
            ph = None
            for c in range(d):
                v = 2 * math.pi * k[m, c].item() * axes[c]
                view = [1] * d
                view[c] = -1
                v = v.view(view)
                ph = v if ph is None else ph + v
            u += amp[m, a].item() * torch.sin(ph + arg)
        out.append(u.to(torch.float32).contiguous())
    return out


def box_mask(res, lo=0.4, hi=0.6, device="cuda"):
    """Solid cells of an axis-aligned box [lo, hi]^d in the unit-height domain
    (dx = 1/res[1]): a cell is solid when its centre lies in the box
    (uint8 tensor of shape res)."""
    dx = 1.0 / res[1]
    m = None
    for a, n in enumerate(res):
        c = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * dx
        inside = (c >= lo) & (c <= hi)
        shape = [1] * len(res)
        shape[a] = n
        inside = inside.reshape(shape)
        m = inside if m is None else m & inside
    return m.to(torch.uint8).contiguous()
