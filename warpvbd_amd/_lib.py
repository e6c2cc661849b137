"""ctypes binding of libwarpvbd.so (declared in include/warpvbd.h).

There is no CPU implementation behind this module: if the shared library is
missing, or no HIP device is present, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# WARPVBD_LIB selects another build of the same library (e.g. one compiled
# with different layout defines); the default is the in-tree build.
LIB_PATH = os.environ.get("WARPVBD_LIB", os.path.join(_HERE, "libwarpvbd.so"))

VBD_MAX_SNAPSHOTS = 16
MATERIALS = {"snh": 0, "stvk": 1}

_ERRNAMES = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EMESH", -5: "ECOLORS", -6: "EVALENCE", -7: "EDEGENERATE"}

# name -> (restype, argtypes)
_P = ctypes.POINTER
_d, _i32, _i64, _u8, _vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint8, ctypes.c_void_p
SIGNATURES = {
    "vbd_create": (ctypes.c_int, [_P(_d), _i32, _P(_i32), _i32, _P(_d), _P(_d), _P(_d), _P(_u8), _P(_d), _d, _i32,
                                  _P(_vp)]),
    "vbd_destroy": (None, [_vp]),
    "vbd_reset": (ctypes.c_int, [_vp]),
    "vbd_step": (ctypes.c_int, [_vp, _P(_d), _d, _d, _i32, _P(_d), _P(_i32), _P(_d)]),
    "vbd_step_dev": (ctypes.c_int, [_vp, _vp, _d, _d, _i32, _vp, _vp, _P(_i32), _P(_d)]),
    "vbd_get_state": (ctypes.c_int, [_vp, _P(_d), _P(_d)]),
    "vbd_set_state": (ctypes.c_int, [_vp, _P(_d), _P(_d)]),
    "vbd_snapshot": (ctypes.c_int, [_vp, _i32]),
    "vbd_read_snapshot": (ctypes.c_int, [_vp, _i32, _P(_d)]),
    "vbd_load_snapshot": (ctypes.c_int, [_vp, _i32]),
    "vbd_info": (ctypes.c_int, [_vp, _P(_i32), _P(_i32), _P(_i32), _P(_i32), _P(_i32), _P(_i64)]),
    "vbd_get_colors": (ctypes.c_int, [_vp, _P(_i32)]),
    "vbd_get_color_sizes": (ctypes.c_int, [_vp, _P(_i32)]),
    "vbd_get_incidence": (ctypes.c_int, [_vp, _P(_i32)]),
    "vbd_color": (ctypes.c_int, [_P(_i32), _i32, _i32, _P(_i32), _P(_i32), _P(_i32)]),
    "vbd_color_graph": (ctypes.c_int, [_P(_i32), _P(_i32), _i32, _P(_i32), _P(_i32)]),
    "vbd_color_noalloc": (ctypes.c_int, [_vp, _i32, _i32, _vp, _P(_i32), _P(_i32), _vp]),
    "vbd_last_dx_max": (ctypes.c_int, [_vp, _P(_d)]),
    "vbd_profile_enable": (ctypes.c_int, [_vp, _i32, _i32]),
    "vbd_profile_read": (ctypes.c_int, [_vp, _P(_d), _P(_i64), _P(_d)]),
    "vbd_color_bytes": (ctypes.c_int, [_vp, _P(_d)]),
    "vbd_set_launch": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32]),
    "vbd_use_device": (ctypes.c_int, [_i32]),
    "vbd_set_graph": (ctypes.c_int, [_vp, _i32]),
    "vbd_batch_create": (ctypes.c_int, [_P(_d), _i32, _P(_i32), _i32, _P(_i32), _i32, _P(_d), _P(_d), _P(_d), _P(_u8),
                                        _P(_d), _d, _i32, _P(_vp)]),
    "vbd_batch_step": (ctypes.c_int, [_vp, _d, _d, _i32, _P(_i32), _P(_d)]),
    "vbd_batch_set_mode": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32]),
    "vbd_batch_results": (ctypes.c_int, [_vp, _P(_i32), _P(_d)]),
    "vbd_batch_get_mode": (ctypes.c_int, [_vp, _P(_i32), _P(_i32), _P(_i32), _P(_i32)]),
    "vbd_batch_autotune": (ctypes.c_int, [_vp, _d, _d, _i32, _i32, _P(_d)]),
    "vbd_set_uniform_material": (ctypes.c_int, [_vp, _d, _d]),
    "vbd_set_gravity": (ctypes.c_int, [_vp, _P(_d)]),
    "vbd_mpnn_create": (ctypes.c_int, [_vp, _i32, _P(_vp)]),
    "vbd_mpnn_destroy": (None, [_vp]),
    "vbd_mpnn_num_params": (ctypes.c_int, []),
    "vbd_mpnn_set_policy": (ctypes.c_int, [_vp, _i32]),
    "vbd_use_record_dictionary": (None, [_i32]),
    "vbd_use_neighbour_staging": (None, [_i32]),
    "vbd_color_layout_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_d)]),
    "vbd_records": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "vbd_staged_colors": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "vbd_tail_sweep": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "vbd_use_tail_sweep": (None, [_i32]),
    "vbd_edge_form": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "vbd_use_edge_form": (None, [_i32]),
    "vbd_mpnn_guess_counts": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "vbd_mpnn_set_params": (ctypes.c_int, [_vp, _P(ctypes.c_float)]),
    "vbd_mpnn_get_params": (ctypes.c_int, [_vp, _P(ctypes.c_float)]),
    "vbd_mpnn_set_material": (ctypes.c_int, [_vp, _d, _d]),
    "vbd_mpnn_set_body_materials": (ctypes.c_int, [_vp, _P(_d), _P(_d)]),
    "vbd_mpnn_collect": (ctypes.c_int, [_vp, _i32, _d, _d, _i32, _P(_i32)]),
    "vbd_mpnn_loss_grad": (ctypes.c_int, [_vp, _i32, _P(ctypes.c_float), _P(ctypes.c_float)]),
    "vbd_mpnn_train": (ctypes.c_int, [_vp, _i32, _i32, _i32, ctypes.c_float, ctypes.c_uint64, _P(ctypes.c_float)]),
    "vbd_mpnn_predict": (ctypes.c_int, [_vp, _i32, _d, _P(ctypes.c_float)]),
    "vbd_mpnn_read_sample": (ctypes.c_int, [_vp, _i32, _P(ctypes.c_float), _P(ctypes.c_float), _P(_i32), _P(_i32),
                                            _P(_i32)]),
    "vbd_step_warm": (ctypes.c_int, [_vp, _vp, _d, _d, _i32, _P(_i32), _P(_d), _P(_i32)]),
    "vbd_last_error": (ctypes.c_char_p, []),
}

_lib = None


class VBDError(RuntimeError):
    pass


def _share_torch_runtime():
    """Load PyTorch first when it is installed.  PyTorch ships its own HIP
    runtime with the same soname as the system one this library links; the
    first one loaded serves the whole process, and PyTorch cannot initialise
    its GPU on the system runtime.  With PyTorch loaded first, both use its
    runtime and tensors can be passed to the solver.  WARPVBD_NO_TORCH=1 skips
    this."""
    if os.environ.get("WARPVBD_NO_TORCH", "0") not in ("", "0"):
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _share_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise VBDError(f"{LIB_PATH} is missing: build it with `python -m warpvbd_amd.build` "
                           "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().vbd_last_error().decode(errors="replace")
        raise VBDError(f"{what} failed ({_ERRNAMES.get(rc, rc)}): {msg}")


def ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(to_numpy(a), dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(to_numpy(a), dtype=np.int32)


def to_numpy(a):
    """numpy view of array-likes: numpy arrays, torch tensors (any device),
    objects with a .numpy() method (e.g. the solver's position arrays)."""
    if isinstance(a, np.ndarray):
        return a
    mod = type(a).__module__
    if mod.startswith("torch"):
        return a.detach().cpu().numpy()
    if hasattr(a, "numpy") and callable(a.numpy):
        return a.numpy()
    return np.asarray(a)


def use_record_dictionary(enable: bool = True) -> None:
    """Share record dictionary entries between slots with equal constants
    (default on; off: one entry per slot); applies to solvers created
    afterwards."""
    load().vbd_use_record_dictionary(int(bool(enable)))


def use_neighbour_staging(enable: bool = True) -> None:
    """Stage each sweep block's neighbours and the dictionary in LDS
    (default on; same results either way); applies to solvers created
    afterwards."""
    load().vbd_use_neighbour_staging(int(bool(enable)))


def use_tail_sweep(enable: bool = True) -> None:
    """Sweep a single body's trailing small colours in one launch by clusters
    of vertices (default on; same results either way); applies to solvers
    created, and launch settings changed, afterwards."""
    load().vbd_use_tail_sweep(int(bool(enable)))


def use_edge_form(enable: bool = True) -> None:
    """Form the SNH terms cof(F) b and J - 1 from each dictionary entry's rest
    edges where the dictionary has at most 64 entries or belongs to a batch
    (default on; off: from F's rows); applies to solvers created
    afterwards."""
    load().vbd_use_edge_form(int(bool(enable)))


def use_device(device) -> int:
    """Make `device` ("cuda", "cuda:1", 1, ...) current for the calling thread."""
    if device is None or device == "cuda":
        idx = 0
    elif isinstance(device, int):
        idx = device
    elif isinstance(device, str) and device.startswith("cuda:"):
        idx = int(device.split(":")[1])
    elif isinstance(device, str) and device == "cpu":
        raise VBDError("device='cpu' is not supported: this solver runs only on the GPU")
    else:
        idx = int(getattr(device, "index", 0) or 0)
    check(load().vbd_use_device(idx), "vbd_use_device")
    return idx
