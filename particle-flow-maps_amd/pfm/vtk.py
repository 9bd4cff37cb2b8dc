"""VTK XML ImageData (.vti) output without the vtk package.

The reference writes vorticity magnitude and smoke with vtk's
vtkXMLImageDataWriter (3D/io_utils.py:101-124): dimensions = the array shape,
origin 0, spacing 1, point data, x fastest (``ravel(order="F")``).  This
writer produces the same dataset as an inline base64 binary .vti that
ParaView and vtk read; ``read_vti`` parses it back (used by the tests).
"""
from __future__ import annotations

import base64
import os
import struct
import xml.etree.ElementTree as ET
from typing import Dict

import numpy as np


def _encode(a: np.ndarray) -> str:
    raw = np.ascontiguousarray(a.ravel(order="F"), dtype="<f4").tobytes()
    return base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")


def write_vti(path: str, arrays: Dict[str, np.ndarray], scalars: str = None) -> None:
    """Write point-data arrays of one shape (nx, ny[, nz]) as a .vti file."""
    if not arrays:
        raise ValueError("no arrays")
    shape = None
    for name, a in arrays.items():
        a = np.asarray(a).squeeze()
        if shape is None:
            shape = a.shape
        elif a.shape != shape:
            raise ValueError(f"array {name} has shape {a.shape}, expected {shape}")
    dims = list(shape) + [1] * (3 - len(shape))
    ext = " ".join(f"0 {n - 1}" for n in dims)
    scalars = scalars or next(iter(arrays))
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" header_type="UInt32">',
             f'  <ImageData WholeExtent="{ext}" Origin="0 0 0" Spacing="1 1 1">',
             f'    <Piece Extent="{ext}">',
             f'      <PointData Scalars="{scalars}">']
    for name, a in arrays.items():
        lines.append(f'        <DataArray type="Float32" Name="{name}" format="binary">')
        lines.append("          " + _encode(np.asarray(a, np.float32).squeeze()))
        lines.append("        </DataArray>")
    lines += ["      </PointData>", "    </Piece>", "  </ImageData>", "</VTKFile>", ""]
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def read_vti(path: str) -> Dict[str, np.ndarray]:
    """Point-data arrays of a .vti written by write_vti, shaped (nx, ny, nz)."""
    root = ET.parse(path).getroot()
    img = root.find("ImageData")
    e = [int(v) for v in img.get("WholeExtent").split()]
    dims = (e[1] - e[0] + 1, e[3] - e[2] + 1, e[5] - e[4] + 1)
    out = {}
    for da in img.iter("DataArray"):
        blob = base64.b64decode("".join(da.text.split()))
        (nbytes,) = struct.unpack("<I", blob[:4])
        data = np.frombuffer(blob[4:4 + nbytes], dtype="<f4")
        out[da.get("Name")] = data.reshape(dims, order="F")
    return out


def write_w_and_smoke(w_numpy, smoke_numpy, outdir, i):
    """3D/io_utils.py:101-124: field_<iii>.vti with "vorticity" and "smoke"."""
    os.makedirs(outdir, exist_ok=True)
    write_vti(os.path.join(outdir, "field_{:03d}.vti".format(i)),
              {"vorticity": np.asarray(w_numpy), "smoke": np.asarray(smoke_numpy)}, scalars="vorticity")
