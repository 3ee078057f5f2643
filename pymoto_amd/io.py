"""VTI (VTK ImageData) writer for fields on a VoxelDomain (``pymoto/common/domain.py:452-604``).

Same file layout as the reference: one ImageData piece over the element grid, vectors whose
length is a multiple of ``nel`` as CellData and of ``nnodes`` as PointData, Float32 DataArrays in
``format="binary"`` (base64, preceded by the base64 of the UInt64 length field the reference
writes), 2D nodal vectors with 2 components padded to 3, 2D blocks ``(n, k)``/``(k, n)`` written as
numbered vectors.  Device tensors are streamed to the host in chunks of whole base64 groups, so a
12.8M-dof displacement never needs a full host copy in FP64.
"""
import base64
import os
import struct
import sys
import warnings

import numpy as np

_CHUNK = 3 * (1 << 20)  # bytes per encoded chunk, a multiple of 3 so base64 chunks concatenate


def _float32_chunks(v, pad_to_vector=False):
    """Yield the Float32 bytes of a 1D host or device vector in chunks (2-component -> 3 padding)."""
    try:
        import torch
        is_t = isinstance(v, torch.Tensor)
    except ImportError:
        is_t = False
    n = v.shape[0]
    step = _CHUNK // (6 if pad_to_vector else 4)
    if pad_to_vector:
        step -= step % 2
    for s in range(0, n, step):
        part = v[s:s + step]
        part = part.detach().to("cpu").numpy() if is_t else np.asarray(part)
        part = part.astype(np.float32)
        if pad_to_vector:
            pad = np.zeros((part.size // 2) * 3, dtype=np.float32)
            pad[0::3], pad[1::3] = part[0::2], part[1::2]
            part = pad
        yield part.tobytes()


def _write_array(file, name, ncomp, v, pad_to_vector=False):
    nbytes = (v.shape[0] // 2) * 3 * 4 if pad_to_vector else v.shape[0] * 4
    enc_len = 4 * ((nbytes + 2) // 3)
    len_enc = ("<" if sys.byteorder == "little" else ">") + "Q"
    file.write(f'<DataArray type="Float32" Name="{name}" NumberOfComponents="{ncomp}" format="binary">\n'.encode())
    file.write(base64.b64encode(struct.pack(len_enc, enc_len)))
    buf = b""
    for chunk in _float32_chunks(v, pad_to_vector):
        buf += chunk
        cut = len(buf) - len(buf) % 3
        file.write(base64.b64encode(buf[:cut]))
        buf = buf[cut:]
    file.write(base64.b64encode(buf))
    file.write(b"\n</DataArray>\n")


def _columns(vec, n):
    """(vectors, ncomponents) of a 1D or 2D array whose data axis has a multiple of n entries."""
    shape = tuple(vec.shape)
    ax = next((i for i, s in enumerate(shape) if s % n == 0), None)
    ncomp = shape[ax] // n
    if len(shape) == 1:
        return [vec], ncomp
    if len(shape) != 2:
        raise AssertionError("Only for 1D and 2D arrays")
    other = (ax + 1) % 2
    return [vec[:, i] if other == 1 else vec[i, :] for i in range(shape[other])], ncomp


def write_to_vti(domain, vectors: dict, filename="out.vti", scale=1.0):
    if ".vti" not in os.path.splitext(filename)[-1].lower():
        filename += ".vti"
    point, cell = {}, {}
    for key, vec in vectors.items():
        size = int(np.prod(tuple(vec.shape)))
        if size % domain.nel == 0:
            cell[key] = vec
        elif size % domain.nnodes == 0:
            point[key] = vec
        else:
            warnings.warn(f"Vector {key} is neither cell- nor point-data. Skipping vector...")
    if not point and not cell:
        warnings.warn(f"Nothing to write to {filename}. Skipping file...")
        return
    byte_order = "LittleEndian" if sys.byteorder == "little" else "BigEndian"
    dx, dy, dz = (np.asarray(domain.element_size[0:3]) * scale)
    o = domain.origin * scale
    ext = f"0 {domain.nelx} 0 {domain.nely} 0 {domain.nelz}"
    with open(filename, "wb") as file:
        file.write(b'<?xml version="1.0"?>\n')
        file.write(f'<VTKFile type="ImageData" version="0.1" header_type="UInt64" byte_order="{byte_order}">\n'.encode())
        file.write(f'<ImageData WholeExtent="{ext}"'.encode())
        file.write(f' Origin="{o[0]} {o[1]} {o[2]}"'.encode())
        file.write(f' Spacing="{dx} {dy} {dz}">\n'.encode())
        file.write(f'<Piece Extent="{ext}">\n'.encode())
        if point:
            file.write(b"<PointData>\n")
            for key, vec in point.items():
                cols, ncomp = _columns(vec, domain.nnodes)
                pad = ncomp == 2 and domain.dim == 2
                nz = int(np.ceil(np.log10(len(cols)))) if len(cols) > 1 else 0
                for i, v in enumerate(cols):
                    name = key if len(cols) == 1 else key + f"({i:0{nz}d})"
                    _write_array(file, name, 3 if pad else ncomp, v.reshape(-1), pad)
            file.write(b"</PointData>\n")
        if cell:
            file.write(b"<CellData>\n")
            for key, vec in cell.items():
                cols, ncomp = _columns(vec, domain.nel)
                for i, v in enumerate(cols):
                    name = key if len(cols) == 1 else key + f"({i})"
                    _write_array(file, name, ncomp, v.reshape(-1))
            file.write(b"</CellData>\n")
        file.write(b"</Piece>\n")
        file.write(b"</ImageData>\n")
        file.write(b"</VTKFile>")
