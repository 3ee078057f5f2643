"""ctypes binding of the C-ABI in ``include/pymoto_amd.h`` (library ``pymoto_amd/lib/libpymoto_amd.so``).

The library is the only compute path of this package: if it is missing, or no GPU is visible,
every device module raises - there is no CPU fallback.
"""
import ctypes
import os

import numpy as np
import torch

K0_EXT_LEN = 684  # PMT_K0_EXT_LEN: K0 plus its reflection-basis blocks (FP64, FP32)
_LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpymoto_amd.so")
_lib = None

c_int, c_ll, c_dbl, c_vp = ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p
c_char_p = ctypes.c_char_p

# name -> (restype, argtypes)
_SIGS = {
    "pmt_last_error": (c_char_p, []),
    "pmt_abi_version": (c_int, []),
    "pmt_filter_fwd": (c_int, [c_int] * 6 + [c_vp] * 5 + [c_vp]),
    "pmt_filter_bwd": (c_int, [c_int] * 6 + [c_vp] * 5 + [c_vp]),
    "pmt_density_filter": (c_int, [c_int] * 3 + [c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pmt_powlaw_fwd": (c_int, [c_ll, c_dbl, c_dbl, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_powlaw_bwd": (c_int, [c_ll, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_int, c_vp]),
    "pmt_heaviside_fwd": (c_int, [c_ll, c_dbl, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_heaviside_bwd": (c_int, [c_ll, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_int, c_vp, c_vp]),
    "pmt_oc_update": (c_int, [c_ll, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp]),
    "pmt_axpby": (c_int, [c_ll, c_dbl, c_vp, c_dbl, c_vp, c_vp]),
    "pmt_dot": (c_int, [c_ll, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pmt_reduce_work_size": (c_int, []),
    "pmt_dot_noalloc": (c_int, [c_int, c_ll, c_vp, c_vp, c_vp, c_vp]),
    "pmt_jacobi_apply": (c_int, [c_ll, c_dbl, c_vp, c_vp, c_vp, c_vp]),
    "pmt_bc_split": (c_int, [c_int] * 3 + [c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp]),
    "pmt_axpy_free": (c_int, [c_int] * 3 + [c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_bc_split_ndof": (c_int, [c_int] * 4 + [c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp]),
    "pmt_axpy_free_ndof": (c_int, [c_int] * 4 + [c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_stiffness_apply": (c_int, [c_int] * 3 + [c_vp, c_int, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_stiffness_residual": (c_int, [c_int] * 3 + [c_vp, c_int, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp]),
    "pmt_stiffness_apply_f32": (c_int, [c_int] * 3 + [c_vp, c_int, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_stiffness_apply_bf16": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_k0_pattern": (c_int, [c_int, c_vp]),
    "pmt_k0_extend": (c_int, [c_int, c_vp, c_vp]),
    "pmt_set_tuning": (c_int, [c_int, c_int]),
    "pmt_get_tuning": (c_int, [c_int]),
    "pmt_stiffness_diagonal": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_dbl, c_vp, c_vp]),
    "pmt_sens_contract": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp, c_int, c_vp]),
    "pmt_sens_contract_k0": (c_int, [c_int] * 3 + [c_vp, c_int, c_vp, c_vp, c_vp, c_dbl, c_vp, c_int, c_vp]),
    "pmt_csr_nnz": (c_ll, [c_int] * 4),
    "pmt_csr_pattern": (c_int, [c_int] * 5 + [c_vp, c_vp, c_vp]),
    "pmt_csr_values": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_dbl, c_int, c_vp, c_vp, c_vp]),
    "pmt_csr_values_ndof": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp, c_dbl, c_int, c_vp, c_vp, c_vp]),
    "pmt_restrict": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp]),
    "pmt_prolong_add": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp]),
    "pmt_mg_create": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp, c_vp]),
    "pmt_mg_destroy": (c_int, [c_vp]),
    "pmt_mg_set_precision": (c_int, [c_vp, c_int]),
    "pmt_mg_set_smoother": (c_int, [c_vp, c_int, c_dbl, c_dbl]),
    "pmt_mg_smoother_interval": (c_int, [c_vp, c_int, c_vp, c_vp]),
    "pmt_mg_update": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_vp]),
    "pmt_mg_vcycle": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pmt_mg_level_apply": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp]),
    "pmt_mg_level_ndof": (c_ll, [c_vp, c_int]),
    "pmt_mg_device_bytes": (c_ll, [c_vp]),
    "pmt_pcg_solve": (c_int, [c_vp, c_vp, c_vp, c_int, c_dbl, c_int, c_int, c_vp, c_vp, c_vp, c_vp]),
    "pmt_mg_set_coarse_space": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp]),
    "pmt_block_pcg_solve": (c_int, [c_vp, c_int, c_vp, c_vp, c_int, c_dbl, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                    c_vp]),
    # scalar (one unknown per node) operator and multigrid PCG
    "pmt_scalar_apply": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pmt_scalar_residual": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp]),
    "pmt_scalar_diagonal": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_dbl, c_vp, c_vp]),
    "pmt_scalar_sens_contract": (c_int, [c_int] * 3 + [c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp, c_int, c_vp]),
    "pmt_smg_create": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp, c_vp]),
    "pmt_smg_destroy": (c_int, [c_vp]),
    "pmt_smg_level_nodes": (c_ll, [c_vp, c_int]),
    "pmt_smg_update": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_vp]),
    "pmt_smg_vcycle": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pmt_smg_level_apply": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp]),
    "pmt_smg_pcg_solve": (c_int, [c_vp, c_vp, c_vp, c_int, c_dbl, c_int, c_int, c_vp, c_vp, c_vp, c_vp]),
    # learned warm start
    "pmt_conv3d_fwd": (c_int, [c_int] * 5 + [c_vp, c_vp, c_vp, c_int, c_vp, c_vp]),
    "pmt_conv3d_bwd_data": (c_int, [c_int] * 5 + [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pmt_conv3d_wgrad_work_size": (c_int, [c_int, c_int]),
    "pmt_conv3d_bwd_weight": (c_int, [c_int] * 5 + [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pmt_relu_mask_bf16": (c_int, [c_ll, c_vp, c_vp, c_vp]),
    "pmt_pool2_bf16": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp]),
    "pmt_pool2_bwd_bf16": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp]),
    "pmt_up2_add_bf16": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp, c_vp]),
    "pmt_up2_bwd_bf16": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp]),
    "pmt_elem_to_node": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp]),
    "pmt_elem_to_node_bwd": (c_int, [c_int] * 4 + [c_vp, c_vp, c_vp]),
    "pmt_adam_step": (c_int, [c_ll, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, c_int, c_vp, c_vp,
                              c_vp, c_vp, c_vp, c_vp]),
    "pmt_add_bf16": (c_int, [c_ll, c_vp, c_vp, c_vp, c_vp]),
    "pmt_f32_to_bf16": (c_int, [c_ll, c_vp, c_vp, c_vp]),
    "pmt_bf16_to_f32": (c_int, [c_ll, c_vp, c_vp, c_vp]),
}


class HIPError(RuntimeError):
    pass


class NumericalError(HIPError, FloatingPointError):
    """A solve met a non-finite right-hand side, start vector or residual (status PMT_ERR_NUMERIC)."""


PMT_ERR_NUMERIC = -4


def load(path: str = _LIBPATH):
    """Load the shared library and declare all signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HIPError(f"HIP library not built: {path} is missing (run `python -m pymoto_amd.build`)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    """The library, after checking that a GPU is visible (device modules call this)."""
    if not torch.cuda.is_available():
        raise HIPError("pymoto_amd needs a ROCm GPU (gfx950); none is visible and there is no CPU fallback")
    return load()


def exported_symbols():
    return list(_SIGS)


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().pmt_last_error().decode(errors="replace")
        cls = NumericalError if rc == PMT_ERR_NUMERIC else HIPError
        raise cls(f"{what or 'HIP call'} failed with status {rc}: {msg}")


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return ctypes.c_void_p(t.data_ptr())


def device():
    return torch.device("cuda", torch.cuda.current_device())


def as_device(x, dtype=torch.float64):
    """numpy / scalar / tensor -> contiguous device tensor of `dtype` (host data is copied once)."""
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.to(device())
        if t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()
    arr = np.ascontiguousarray(np.asarray(x), dtype=torch.empty(0, dtype=dtype).numpy().dtype)
    return torch.from_numpy(arr).to(device())


_work = {}


def reduce_work():
    dev = torch.cuda.current_device()
    if dev not in _work:
        _work[dev] = torch.empty(load().pmt_reduce_work_size(), dtype=torch.float64, device=device())
    return _work[dev]
