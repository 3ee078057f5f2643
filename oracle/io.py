"""Oracle of the VTI writer (test infrastructure only): VoxelDomain.write_to_vti of
``pymoto/common/domain.py:452-604`` restated for numpy arrays, whole arrays at once.

Pinning: tests/test_host.py decodes the written arrays back (base64 Float32) and checks values,
names, component counts and extents.
"""
import base64
import struct
import sys

import numpy as np


def write_vti(domain, vectors, filename, scale=1.0):
    len_enc = ("<" if sys.byteorder == "little" else ">") + "Q"
    point = {k: v for k, v in vectors.items() if v.size % domain.nel != 0 and v.size % domain.nnodes == 0}
    cell = {k: v for k, v in vectors.items() if v.size % domain.nel == 0}

    def arr(f, name, ncomp, v):
        f.write(f'<DataArray type="Float32" Name="{name}" NumberOfComponents="{ncomp}" format="binary">\n'.encode())
        enc = base64.b64encode(np.ascontiguousarray(v, dtype=np.float32))
        f.write(base64.b64encode(struct.pack(len_enc, len(enc))))
        f.write(enc)
        f.write(b"\n</DataArray>\n")

    def cols(vec, n):
        ax = next(i for i, s in enumerate(vec.shape) if s % n == 0)
        if vec.ndim == 1:
            return [vec], vec.shape[ax] // n
        other = (ax + 1) % 2
        return [np.take(vec, i, axis=other) for i in range(vec.shape[other])], vec.shape[ax] // n

    bo = "LittleEndian" if sys.byteorder == "little" else "BigEndian"
    dx, dy, dz = np.asarray(domain.element_size[0:3]) * scale
    o = domain.origin * scale
    ext = f"0 {domain.nelx} 0 {domain.nely} 0 {domain.nelz}"
    with open(filename, "wb") as f:
        f.write(b'<?xml version="1.0"?>\n')
        f.write(f'<VTKFile type="ImageData" version="0.1" header_type="UInt64" byte_order="{bo}">\n'.encode())
        f.write(f'<ImageData WholeExtent="{ext}" Origin="{o[0]} {o[1]} {o[2]}" Spacing="{dx} {dy} {dz}">\n'.encode())
        f.write(f'<Piece Extent="{ext}">\n'.encode())
        if point:
            f.write(b"<PointData>\n")
            for key, vec in point.items():
                cs, nc = cols(vec, domain.nnodes)
                pad = nc == 2 and domain.dim == 2
                nz = int(np.ceil(np.log10(len(cs)))) if len(cs) > 1 else 0
                for i, v in enumerate(cs):
                    v = v.ravel()
                    if pad:
                        w = np.zeros(3 * domain.nnodes)
                        w[0::3], w[1::3] = v[0::2], v[1::2]
                        v = w
                    arr(f, key if len(cs) == 1 else key + f"({i:0{nz}d})", 3 if pad else nc, v)
            f.write(b"</PointData>\n")
        if cell:
            f.write(b"<CellData>\n")
            for key, vec in cell.items():
                cs, nc = cols(vec, domain.nel)
                for i, v in enumerate(cs):
                    arr(f, key if len(cs) == 1 else key + f"({i})", nc, v.ravel())
            f.write(b"</CellData>\n")
        f.write(b"</Piece>\n</ImageData>\n</VTKFile>")


def read_vti_arrays(filename):
    """{name: (ncomponents, float32 array)} decoded from a file written in this format."""
    import re
    data = open(filename, "rb").read()
    out = {}
    for m in re.finditer(rb'<DataArray type="Float32" Name="([^"]+)" NumberOfComponents="(\d+)" format="binary">\n'
                         rb"([A-Za-z0-9+/=]+)\n</DataArray>", data):
        blob = m.group(3)
        body = base64.b64decode(blob[12:])  # 8-byte length field -> 12 base64 characters
        out[m.group(1).decode()] = (int(m.group(2)), np.frombuffer(body, dtype=np.float32))
    return out
