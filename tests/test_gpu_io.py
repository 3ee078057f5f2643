"""VTI output of device tensors: the writer streams GPU vectors to the file in chunks; the bytes must
equal those of the reference's own VoxelDomain.write_to_vti (domain.py:452-604) on the same values
(fixture tests/golden/vti_cases.npz, written by tests/golden/make_golden.py)."""
import numpy as np
import pytest
import torch

import pymoto_amd as pm
from conftest import golden

pytestmark = pytest.mark.gpu


def test_vti_device_tensors_equal_reference_bytes(gpu_lib, tmp_path, monkeypatch):
    from pymoto_amd import io
    g = golden("vti_cases.npz")
    for chunk in (io._CHUNK, 48):  # default chunking and a tiny one: many chunks per array, odd remainders
        monkeypatch.setattr(io, "_CHUNK", chunk)
        for i in range(int(g["ncases"])):
            nel = tuple(int(v) for v in g[f"nel_{i}"])
            unit = g[f"unit_{i}"]
            d = pm.VoxelDomain(*nel, unitx=unit[0], unity=unit[1], unitz=unit[2])
            vecs = {k: torch.as_tensor(g[f"{k}_{i}"], device="cuda") for k in ("x", "u", "modes")}
            out = tmp_path / f"dev{i}_{chunk}.vti"
            d.write_to_vti(vecs, str(out), scale=float(g[f"scale_{i}"]))
            assert out.read_bytes() == g[f"bytes_{i}"].tobytes(), (nel, chunk)


def test_vti_device_field_decodes_to_its_values(gpu_lib, tmp_path):
    """A nodal field of 128x64x64 hex8 (1.6M FP32 values, 6.5 MB: several 3 MiB chunks) written from the
    device decodes back to its FP32 values: the chunked base64 stream concatenates to one valid block."""
    from oracle.io import read_vti_arrays
    d = pm.VoxelDomain(128, 64, 64)
    u = torch.linspace(-1.0, 1.0, 3 * d.nnodes, dtype=torch.float64, device="cuda")
    x = torch.rand(d.nel, dtype=torch.float64, device="cuda")
    out = tmp_path / "big.vti"
    d.write_to_vti({"u": u, "x": x}, str(out))
    arrays = read_vti_arrays(str(out))
    assert arrays["u"][0] == 3 and arrays["x"][0] == 1
    np.testing.assert_array_equal(arrays["u"][1], u.cpu().numpy().astype(np.float32))
    np.testing.assert_array_equal(arrays["x"][1], x.cpu().numpy().astype(np.float32))
