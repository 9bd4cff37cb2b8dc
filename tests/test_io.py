"""Checkpoint files and the .vti writer (CPU): the reference's velocity file
names and numpy layout, and a VTK ImageData round trip."""
import os

import numpy as np
import torch

from pfm import checkpoint, vtk


def test_velocity_checkpoint_names_and_roundtrip(tmp_path):
    res = (5, 4, 3)
    u = [torch.randn(tuple(r + (a == b) for b, r in enumerate(res))) for a in range(3)]
    checkpoint.save_velocity(str(tmp_path), 7, u)
    # 3D/run_3D.py:846-848: vel_x_numpy_<frame>.npy, the field's numpy shape
    for a, ax in enumerate("xyz"):
        arr = np.load(os.path.join(str(tmp_path), f"vel_{ax}_numpy_7.npy"), allow_pickle=False)
        assert arr.dtype == np.float32 and arr.shape == tuple(u[a].shape)
        np.testing.assert_array_equal(arr, u[a].numpy())
    v = [torch.zeros_like(t) for t in u]
    checkpoint.load_velocity(str(tmp_path), 7, v)
    for a in range(3):
        assert torch.equal(u[a], v[a])


def test_velocity_checkpoint_rejects_wrong_shape(tmp_path):
    u = [torch.zeros(5, 4), torch.zeros(4, 5)]
    checkpoint.save_velocity(str(tmp_path), 0, u)
    bad = [torch.zeros(6, 4), torch.zeros(4, 5)]
    try:
        checkpoint.load_velocity(str(tmp_path), 0, bad)
    except ValueError as e:
        assert "expected float32" in str(e)
    else:
        raise AssertionError("shape mismatch accepted")


def test_vti_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    w = rng.normal(size=(6, 5, 4)).astype(np.float32)
    s = rng.uniform(size=(6, 5, 4)).astype(np.float32)
    vtk.write_w_and_smoke(w, s, str(tmp_path), 12)
    path = os.path.join(str(tmp_path), "field_012.vti")
    text = open(path).read()
    assert 'WholeExtent="0 5 0 4 0 3"' in text and 'Scalars="vorticity"' in text
    back = vtk.read_vti(path)
    np.testing.assert_array_equal(back["vorticity"], w)
    np.testing.assert_array_equal(back["smoke"], s)
    # 2D arrays become one z-slab
    vtk.write_vti(os.path.join(str(tmp_path), "f2.vti"), {"w": w[:, :, 0]})
    np.testing.assert_array_equal(vtk.read_vti(os.path.join(str(tmp_path), "f2.vti"))["w"][:, :, 0], w[:, :, 0])
