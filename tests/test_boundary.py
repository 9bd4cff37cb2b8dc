"""The C-ABI boundary: libpfm.so loads, exports every entry point that
include/pfm.h declares, and the Python surface fails loudly without a GPU.
No compute calls (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfm.h")
LIB = os.path.join(ROOT, "particle-flow-maps_amd", "pfm", "libpfm.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfm_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "pfm_rk4_particles" in names and "pfm_mg_poisson" in names and "pfm_p2g" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libpfm.so first (particle-flow-maps_amd/build.py)"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    from pfm import _lib

    assert set(_lib.exported_symbols()) == set(declared())


def test_abi_version_and_error_text():
    lib = ctypes.CDLL(LIB)
    assert lib.pfm_abi_version() == 1
    lib.pfm_last_error.restype = ctypes.c_char_p
    # argument validation runs on the host and needs no device
    from pfm import _lib

    g = _lib.make_grid((8, 8, 2), 0.125)  # res 2 < 4 -> rejected
    rc = lib.pfm_reset_identity_grid(ctypes.byref(g), 0, None, None, None)
    assert rc == -1
    assert b"res[2]" in lib.pfm_last_error()


def test_workspace_size_query():
    from pfm import _lib

    L = _lib.load(require_gpu=False)
    g = _lib.make_grid((16, 16, 16), 1 / 16)
    ws = L.pfm_p2g_workspace_size(ctypes.byref(g), 16 ** 3 * 8)
    assert ws >= 16 ** 3 * 8 * (4 + 4 + 15 * 4)


def test_product_path_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import pfm

    with pytest.raises(pfm._lib.PfmError):
        pfm.load_library(require_gpu=True)
    with pytest.raises(Exception):
        pfm.MGPCG_3([[2, 2]] * 3, [16, 16, 16])


def test_product_does_not_import_oracle():
    """Nothing in the package may reach the CPU oracle."""
    pkg = os.path.join(ROOT, "particle-flow-maps_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "libpfm_oracle" not in text, f


def test_scene_structs_match_header(tmp_path):
    """The ctypes mirrors of pfm_scene_desc / pfm_scene_state have the C
    layout (sizes and every field offset, from a gcc build of the header)."""
    import ctypes
    import shutil
    import subprocess

    from pfm import _lib

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    structs = {"pfm_scene_desc": _lib.PfmSceneDesc, "pfm_scene_state": _lib.PfmSceneState,
               "pfm_batch2d_desc": _lib.PfmBatch2dDesc, "pfm_batch2d_state": _lib.PfmBatch2dState}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _t in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([gcc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py)
        for f, _t in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_batch_sizes_query():
    """pfm_batch2d_sizes: totals of the concatenated batch layout, host only."""
    import numpy as np

    from pfm import _lib

    lib = _lib.load(require_gpu=False)
    res = np.ascontiguousarray(np.array([64, 64, 97, 80, 4, 5], np.int32))
    out = (ctypes.c_int64 * 4)()
    assert lib.pfm_batch2d_sizes(3, res.ctypes.data, 2, out) == 0
    pairs = [(64, 64), (97, 80), (4, 5)]
    assert list(out) == [sum(a * b * 4 for a, b in pairs), sum((a + 1) * b for a, b in pairs),
                         sum(a * (b + 1) for a, b in pairs), sum(a * b for a, b in pairs)]
    bad = np.ascontiguousarray(np.array([64, 3], np.int32))
    assert lib.pfm_batch2d_sizes(1, bad.ctypes.data, 2, out) == -1
    assert "out of" in _lib.last_error()


def test_batch_create_rejects_bad_descriptors():
    """pfm_batch2d_create validates the descriptor before touching the GPU."""
    import numpy as np

    from pfm import _lib

    lib = _lib.load(require_gpu=False)
    res = np.ascontiguousarray(np.array([64, 64, 80, 72], np.int32))
    seeds = np.ascontiguousarray(np.array([1, 2], np.int64))
    d = _lib.PfmBatch2dDesc()
    d.n_scenes, d.res, d.seed = 2, res.ctypes.data, seeds.ctypes.data
    d.ppc_axis, d.reinit_every, d.reinit_every_grad_m = 2, 12, 8
    d.CFL, d.visualize_dt, d.base_level, d.max_iters, d.tol = 0.5, 0.1, 4, 400, 1e-6
    h = ctypes.c_void_p()
    nbytes = ctypes.c_int64()
    # every device field is still NULL
    assert lib.pfm_batch2d_create(ctypes.byref(d), ctypes.byref(h), ctypes.byref(nbytes)) == -1
    assert "null field pointer" in _lib.last_error()
    d.max_iters = -1
    assert lib.pfm_batch2d_create(ctypes.byref(d), ctypes.byref(h), ctypes.byref(nbytes)) == -1
    assert "max_iters" in _lib.last_error()
    d.max_iters, d.reinit_every = 400, 0
    assert lib.pfm_batch2d_create(ctypes.byref(d), ctypes.byref(h), ctypes.byref(nbytes)) == -1
    assert not h.value


def test_fused_entry_points_validate_arguments():
    """The fused step entry points check their arguments on the host (no GPU
    call is made before the checks)."""
    from pfm import _lib

    lib = _lib.load(require_gpu=False)
    g = _lib.make_grid((16, 16, 16), 1 / 16)
    f = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    u0 = [ctypes.c_void_p(0x1000 + 0x100 * k) for k in range(3)]
    u1 = [ctypes.c_void_p(0x2000 + 0x100 * k) for k in range(3)]
    # the midpoint launch needs every face pointer of every axis
    rc = lib.pfm_midpoint_advect(ctypes.byref(g), *u0, *u1, f, None, f, f, f, f, ctypes.c_float(0.1), None)
    assert rc == -1 and "null pointer" in _lib.last_error()
    # ... and u1 must not alias u0
    rc = lib.pfm_midpoint_advect(ctypes.byref(g), f, f, f, f, f, f, f, f, f, f, f, f, ctypes.c_float(0.1), None)
    assert rc == -1 and "alias" in _lib.last_error()
    # the fused RK4 + record + count pass is 3D and needs n >= 1
    g2 = _lib.make_grid((16, 16), 1 / 16)
    args = [f] * 3 + [ctypes.c_float(0.1)] + [f] * 7 + [None, f, ctypes.c_size_t(1 << 30), None]
    rc = lib.pfm_rk4_particles_p2g_record_fold_count(ctypes.byref(g2), 8, f, *args)
    assert rc == -1 and "3D" in _lib.last_error()
    rc = lib.pfm_rk4_particles_p2g_record_fold_count(ctypes.byref(g), 0, f, *args)
    assert rc == -1 and "n out of range" in _lib.last_error()
    # the rest of the binning rejects a workspace smaller than the layout
    need = lib.pfm_p2g_workspace_size(ctypes.byref(g), 4096)
    rc = lib.pfm_bin_finish(ctypes.byref(g), 4096, f, f, f, f, ctypes.c_size_t(need - 1), None)
    assert rc == -1 and "workspace too small" in _lib.last_error()
