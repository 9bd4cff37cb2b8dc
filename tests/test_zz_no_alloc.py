"""pfm_kinetic_energy_noalloc (the entry point that allocates nothing) called
through the C ABI against pfm_kinetic_energy on the same fields: equal bits,
on two calls in a row, the second on a smaller grid (the library's fixed
scratch keeps the first call's partials, which the second must not read)."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _fields(res, seed):
    from pfm import core

    g = torch.Generator(device="cpu").manual_seed(seed)
    return [torch.randn(core.face_shape(res, a), generator=g).to(DEV) for a in range(len(res))]


def _both(L, res, u):
    lib = L.load()
    g = L.make_grid(res, 1.0 / res[1])
    ptrs = [t.data_ptr() for t in u] + [None] * (3 - len(u))
    ref = torch.zeros(1 + 3 * 1024, dtype=torch.float64, device=DEV)
    out = torch.full((1,), float("nan"), dtype=torch.float64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    assert lib.pfm_kinetic_energy(ctypes.byref(g), *ptrs, ref.data_ptr(), s) == 0
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    rc = lib.pfm_kinetic_energy_noalloc(ctypes.byref(g), *ptrs, out.data_ptr(), s)
    assert rc == 0, L.last_error()
    torch.cuda.synchronize()
    assert torch.cuda.memory_allocated() == before
    return float(ref[0]), float(out[0])


@pytest.mark.parametrize("shapes", [((96, 80, 64), (12, 10, 8)), ((300, 200), (20, 16))])
def test_noalloc_kinetic_energy_equals_workspace_entry_point(pfm, shapes):
    from pfm import _lib as L

    for k, res in enumerate(shapes):
        u = _fields(res, 40 + k)
        want, got = _both(L, res, u)
        assert got == want, (res, got, want)
        assert want > 0


def test_noalloc_rejects_bad_arguments(pfm):
    from pfm import _lib as L

    lib = L.load()
    g = L.make_grid((8, 8, 8), 1.0 / 8)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    assert lib.pfm_kinetic_energy_noalloc(ctypes.byref(g), None, None, None, out.data_ptr(), None) != 0
    assert "null" in L.last_error()
