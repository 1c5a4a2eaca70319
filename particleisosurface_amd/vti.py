"""VTK XML image data (.vti) with one Float32 point array, raw appended data (UInt64 header).

The dense-grid output of UniformGrid::Run (src/marching.cpp:155-185) in a form ParaView reads;
``read_vti`` reads back what ``write_vti`` (and the C++ writeVtiFile) produce."""
import re

import numpy as np


def write_vti(path, scalars, origin, spacing):
    """scalars: (dims2, dims1, dims0) float32 array, x fastest."""
    a = np.ascontiguousarray(scalars, dtype="<f4")
    d2, d1, d0 = a.shape
    head = ('<?xml version="1.0"?>\n<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64">\n  <ImageData WholeExtent="0 %d 0 %d 0 %d" Origin="%.9g %.9g %.9g" '
            'Spacing="%.9g %.9g %.9g" Direction="1 0 0 0 1 0 0 0 1">\n    <Piece Extent="0 %d 0 %d 0 %d">\n'
            '      <PointData Scalars="scalars">\n        <DataArray type="Float32" Name="scalars" '
            'format="appended" offset="0"/>\n      </PointData>\n      <CellData>\n      </CellData>\n'
            '    </Piece>\n  </ImageData>\n  <AppendedData encoding="raw">\n   _'
            % (d0 - 1, d1 - 1, d2 - 1, origin[0], origin[1], origin[2], spacing, spacing, spacing,
               d0 - 1, d1 - 1, d2 - 1))
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(np.uint64(a.nbytes).tobytes())
        f.write(a.tobytes())
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vti(path):
    """Returns (scalars (dims2, dims1, dims0) float32, origin (3,), spacing (3,))."""
    raw = open(path, "rb").read()
    cut = raw.index(b"<AppendedData")
    xml = raw[:cut].decode()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', xml).group(1).split()]
    origin = np.array([float(v) for v in re.search(r'Origin="([^"]+)"', xml).group(1).split()])
    spacing = np.array([float(v) for v in re.search(r'Spacing="([^"]+)"', xml).group(1).split()])
    start = raw.index(b"_", cut) + 1
    nb = int(np.frombuffer(raw[start:start + 8], "<u8")[0])
    data = np.frombuffer(raw[start + 8:start + 8 + nb], "<f4")
    dims = (ext[5] - ext[4] + 1, ext[3] - ext[2] + 1, ext[1] - ext[0] + 1)
    return data.reshape(dims).copy(), origin, spacing
