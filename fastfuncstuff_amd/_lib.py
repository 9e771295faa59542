"""ctypes binding of ``libffsreml.so`` (the C ABI declared in include/ffs_reml.h).

The library is required: every compute entry point raises when it is missing
or when no HIP device is present.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libffsreml.so"

_c_int = ctypes.c_int32
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (all return int32 status unless listed in _RESTYPES)
SIGNATURES: dict[str, list] = {
    "ffs_reml_version": [],
    "ffs_reml_last_error": [],
    "ffs_reml_plan_create": [_pp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _c_int, _c_int, _vp, _c_int,
                             _c_int, _c_int, _c_int, _vp, _vp],
    "ffs_reml_plan_create_ridge": [_pp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _c_int, _c_int, _vp, _c_int,
                                   _c_int, _c_int, ctypes.c_double, _vp],
    "ffs_reml_plan_destroy": [_vp],
    "ffs_reml_plan_rank": [_vp],
    "ffs_reml_dsort_workspace": [_vp, _i64, _c_int, _c_int, _vp],
    "ffs_reml_dsort_gls": [_vp, _vp, _i64, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                           _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_plan_create_many": [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _c_int, _c_int, _vp,
                                  _c_int, _c_int, _c_int, _c_int, _vp, _vp],
    "ffs_reml_group_create": [_pp, _c_int, _vp, _i64, _vp, _vp],
    "ffs_reml_group_destroy": [_vp],
    "ffs_reml_group_fit": [_vp, _vp, _i64, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_group_tune": [_vp, _vp, _i64, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _i64, _vp],
    "ffs_reml_dense_likelihood": [_vp, _vp, _c_int, _vp, _c_int, _c_int, _vp, _i64, _vp, _vp],
    "ffs_reml_plan_logdets": [_vp, _vp, _vp],
    "ffs_reml_plan_ainv": [_vp, _vp],
    "ffs_reml_plan_dims": [_vp, _vp, _vp, _vp],
    "ffs_reml_plan_set_projection": [_vp, _c_int],
    "ffs_reml_prep": [_vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _c_int, _vp, _vp],
    "ffs_reml_prep_features": [_vp, _vp, _i64, _i64, _vp, _c_int, _vp, _vp, _vp],
    "ffs_reml_whiten_cols": [_vp, _vp, _i64, _c_int, _vp, _i64, _vp, _vp],
    "ffs_reml_plan_coef_stride": [_vp, _vp],
    "ffs_reml_energy": [_vp, _vp, _i64, _i64, _vp, _i64, _vp],
    "ffs_reml_project": [_vp, _vp, _i64, _i64, _vp, _i64, _vp],
    "ffs_reml_select": [_vp, _vp, _i64, _i64, _c_int, _vp, _vp, _vp],
    "ffs_reml_surface_f64": [_vp, _vp, _c_int, _i64, _i64, _vp, _vp, _i64, _vp],
    "ffs_reml_select_f64": [_vp, _vp, _i64, _i64, _c_int, _vp, _vp, _vp],
    "ffs_reml_gls": [_vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "ffs_reml_gls_noalloc": [_vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "ffs_reml_ljung_box": [_vp, _i64, _i64, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp],
    "ffs_reml_local_workspace": [_vp, _i64, _vp],
    "ffs_reml_local_recenter": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_gls_workspace": [_vp, _i64, _vp],
    "ffs_reml_plan_yt_rows": [_vp, _vp],
    "ffs_reml_surface": [_vp, _vp, _i64, _i64, _vp, _i64, _c_int, _vp],
    "ffs_reml_plan_build_hier": [_vp, _vp],
    "ffs_reml_hier_workspace": [_vp, _i64, _vp],
    "ffs_reml_hier_search": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_local_prepare": [_vp, _vp, _i64, _i64, _vp, ctypes.c_float, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_local_energy": [_vp, _i64, _vp, _i64, _vp],
    "ffs_reml_local_project": [_vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_local_search": [_vp, _vp, _i64, _i64, _vp, ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_compact": [_vp, _i64, _vp, _vp, _vp],
    "ffs_reml_subset_workspace": [_vp, _i64, _vp],
    "ffs_reml_subset_search": [_vp, _vp, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_plan_set_bands": [_vp, ctypes.c_double, ctypes.c_double, _vp],
    "ffs_reml_plan_band_info": [_vp, _vp, _vp, _vp, _vp, _vp],
    "ffs_reml_plan_window_info": [_vp, _vp, _vp, _vp, _vp, _vp],
    "ffs_reml_band_workspace": [_vp, _i64, _vp],
    "ffs_reml_band_prepare": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_band_energy": [_vp, _i64, _vp, _i64, _vp],
    "ffs_reml_band_project": [_vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_band_search": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ffs_reml_scatter_best": [_vp, _i64, _vp, _vp, _vp, _vp, _vp],
    "ffs_mlp_param_count": [],
    "ffs_mlp_workspace_floats": [_i64],
    "ffs_mlp_last_error": [],
    "ffs_mlp_forward": [_vp, _vp, _vp, _i64, _vp, _vp],
    "ffs_mlp_grad": [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "ffs_mlp_train_steps": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _c_int, _c_int, _c_int, ctypes.c_float,
                            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp],
    "ffs_mlp_adam": [_vp, _vp, _vp, _vp, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                     _c_int, _vp],
}
_RESTYPES = {"ffs_reml_version": ctypes.c_char_p, "ffs_reml_last_error": ctypes.c_char_p,
             "ffs_mlp_last_error": ctypes.c_char_p}

_lib = None


class RemlLibraryError(RuntimeError):
    """A call into libffsreml.so failed."""


def load(required_symbols: bool = True):
    """Load the shared library (once) and bind the declared signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RemlLibraryError(
            f"{LIB_PATH} is missing: build it with `python -m fastfuncstuff_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            if required_symbols:
                raise RemlLibraryError(f"{LIB_PATH} does not export {name}")
            continue
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, ctypes.c_int32)
    _lib = lib
    return lib


def last_error() -> str:
    return load().ffs_reml_last_error().decode()


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = last_error()
        if rc == -2:
            raise torch.linalg.LinAlgError(f"{what}: {msg}")
        if rc == -3:
            raise ValueError(f"{what}: {msg}")
        raise RemlLibraryError(f"{what} failed ({rc}): {msg}")


def require_device(device: torch.device | None = None) -> torch.device:
    """The HIP device to run on; raises when there is none (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RemlLibraryError("fastfuncstuff_amd needs a HIP device (MI355X); none is visible")
    if device is None or device.type != "cuda":
        return torch.device("cuda", torch.cuda.current_device())
    return device


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(device: torch.device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


# Host -> device upload of large host arrays through two pinned staging
# buffers: the host copies chunk i + 1 into one buffer while the DMA of chunk i
# from the other runs, instead of the runtime's own staging of a pageable
# source (about 2x the rate for a few hundred MB).  The buffers are allocated
# pinned once per device and reused; a lock serialises their use.
_STAGE_BYTES = 32 << 20
_STAGE_MIN_BYTES = 16 << 20
_stage: dict = {}
_stage_lock = threading.Lock()


def upload_f32(src, device: torch.device) -> torch.Tensor:
    """``src`` (array-like, any real dtype) as a contiguous float32 tensor on
    ``device``, in stream order on the current stream (the values are those
    of ``torch.as_tensor(src).to(device, torch.float32)``)."""
    src = torch.as_tensor(src)
    if (src.device.type != "cpu" or src.numel() * 4 < _STAGE_MIN_BYTES or src.ndim == 0 or src.is_pinned()
            or not src.is_contiguous()):
        return src.to(device=device, dtype=torch.float32).contiguous()
    out = torch.empty(src.shape, dtype=torch.float32, device=device)
    flat_src = src.reshape(-1)
    flat_out = out.reshape(-1)
    n = flat_src.numel()
    chunk = _STAGE_BYTES // 4
    stream = torch.cuda.current_stream(device)
    with _stage_lock:
        key = device.index if device.index is not None else torch.cuda.current_device()
        st = _stage.get(key)
        if st is None:
            st = _stage[key] = {"buf": [torch.empty(chunk, dtype=torch.float32).pin_memory() for _ in range(2)],
                                "ev": [None, None]}
        for i, o in enumerate(range(0, n, chunk)):
            m = min(chunk, n - o)
            slot = i & 1
            if st["ev"][slot] is not None:
                st["ev"][slot].synchronize()  # the DMA that last read this buffer is done
            buf = st["buf"][slot][:m]
            buf.copy_(flat_src[o:o + m])
            flat_out[o:o + m].copy_(buf, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            st["ev"][slot] = ev
    return out
