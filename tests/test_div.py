"""The correctly rounded division by the stencil diagonal used by the MGPCG
level kernels (particle-flow-maps_amd/csrc/pfm_div.h), compiled for the host
from the same header and checked against IEEE fp32 division (numpy)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR_DIR = os.path.join(ROOT, "particle-flow-maps_amd", "csrc")

SRC = r"""
#include "pfm_div.h"
void div_array(const float* x, float* q, long n, int d) {
    const float nf = (float)d, y = pfm_rcp_small(d);
    for (long i = 0; i < n; ++i) q[i] = pfm_div_small(x[i], nf, y);
}
float rcp(int d) { return pfm_rcp_small(d); }
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("div")
    src = d / "div.c"
    so = d / "libdiv.so"
    src.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-mfma", "-shared", "-fPIC", f"-I{HDR_DIR}", str(src),
                    "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.div_array.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int]
    L.rcp.argtypes = [ctypes.c_int]
    L.rcp.restype = ctypes.c_float
    return L


def _edge_values():
    f32 = np.float32
    vals = [0.0, -0.0, np.inf, -np.inf, np.nan, np.finfo(f32).max, -np.finfo(f32).max,
            np.finfo(f32).tiny, np.finfo(f32).smallest_subnormal, 2.0 ** -125, 2.0 ** -124, 1.0, 3.0, 6.0]
    out = np.array(vals, dtype=f32)
    # every exponent, mantissas at the ends and the middle
    e = np.arange(256, dtype=np.uint32) << 23
    mant = np.array([0, 1, 2, 0x2AAAAA, 0x555555, 0x7FFFFE, 0x7FFFFF], dtype=np.uint32)
    bits = (e[:, None] | mant[None, :]).ravel()
    bits = np.concatenate([bits, bits | np.uint32(0x80000000)])
    return np.concatenate([out, bits.view(f32)])


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6])
def test_div_small_matches_ieee(lib, d):
    rng = np.random.default_rng(d)
    x = np.concatenate([_edge_values(), rng.integers(0, 2 ** 32, size=1 << 21, dtype=np.uint64)
                        .astype(np.uint32).view(np.float32)])
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.empty_like(x)
    lib.div_array(x.ctypes.data, q.ctypes.data, x.size, d)
    with np.errstate(all="ignore"):
        ref = x / np.float32(d)
    same = (q.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(q) & np.isnan(ref))
    assert same.all(), x[~same][:8]


def test_rcp_small_is_rounded_reciprocal(lib):
    for d in range(1, 7):
        assert lib.rcp(d) == np.float32(1.0) / np.float32(d)
    assert lib.rcp(0) == 0.0 and lib.rcp(7) == 0.0
