"""ctypes binding of ``include/pgeof_hip.h`` (the device-pointer entry points).

This is the binding a ctypes caller of ``libpgeof_hip.so`` writes; the Python
package's own functions go through the pybind11 module ``pgeof_ext`` instead.
Device pointers are plain integers (e.g. ``torch.Tensor.data_ptr()``) and
streams are ``hipStream_t`` handles as integers (``torch.cuda.Stream.cuda_stream``).
Loading fails loudly when the library is not built; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import json
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgeof_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "pgeof_hip.h")

PGEOF_OK, PGEOF_EINVAL, PGEOF_ERUNTIME, PGEOF_ENOMEM = 0, 1, 2, 3

_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_F = ctypes.c_float
_D = ctypes.c_double
_SZ = ctypes.c_size_t

_SIGS = {
    "pgeof_last_error": (ctypes.c_char_p, []),
    "pgeof_abi_version": (ctypes.c_int, []),
    "pgeof_device_count": (ctypes.c_int, []),
    "pgeof_knn_search": (ctypes.c_int, [_P, _I64, _P, _I64, _U32, _P, _P]),
    "pgeof_radius_search": (ctypes.c_int, [_P, _I64, _P, _I64, _F, _U32, _P, _P]),
    "pgeof_compute_features": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _U64, _P]),
    "pgeof_compute_features_multiscale": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _P, _I64, _P]),
    "pgeof_compute_features_optimal": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _U32, _U32, _U32, _P]),
    "pgeof_compute_features_selected_f32": (ctypes.c_int, [_P, _I64, _F, _U32, _P, _I64, _P]),
    "pgeof_compute_features_selected_f64": (ctypes.c_int, [_P, _I64, _D, _U32, _P, _I64, _P]),
    "pgeof_knn_workspace_size": (_SZ, [_I64, _I64, _U32]),
    "pgeof_radius_workspace_size": (_SZ, [_I64, _I64, _U32]),
    "pgeof_radius_csr_workspace_size": (_SZ, [_I64, _U32]),
    "pgeof_knn_search_device": (ctypes.c_int, [_P, _I64, _P, _I64, _U32, _P, _P, _P, _SZ, _P]),
    "pgeof_radius_search_device": (ctypes.c_int, [_P, _I64, _P, _I64, _F, _U32, _P, _P, _P, _SZ, _P]),
    "pgeof_knn_features_device": (ctypes.c_int, [_P, _I64, _U32, _P, _P, _P, _P, _P, _SZ, _P]),
    "pgeof_radius_csr_device": (ctypes.c_int, [_P, _I64, _F, _U32, _P, _P, ctypes.POINTER(_I64), _P, _SZ, _P]),
    "pgeof_compute_features_device": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _U64, _P, _P, _P]),
    "pgeof_radius_multiscale_workspace_size": (_SZ, [_I64, _U32, _I64]),
    "pgeof_radius_multiscale_device": (ctypes.c_int, [_P, _I64, _F, _U32, _P, _I64, _P, _P, _P, _P, _SZ, _P]),
    "pgeof_knn_optimal_device": (ctypes.c_int, [_P, _I64, _U32, _U32, _U32, _U32, _P, _P, _P, _P, _SZ, _P]),
    "pgeof_compute_features_device_noalloc": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _U64, _P, _P, _P]),
    "pgeof_compute_features_multiscale_device": (ctypes.c_int,
                                                 [_P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P]),
    "pgeof_compute_features_optimal_device": (ctypes.c_int,
                                              [_P, _I64, _P, _I64, _P, _I64, _U32, _U32, _U32, _P, _P, _P]),
    "pgeof_autotune": (ctypes.c_int, [ctypes.c_char_p, _P, _I64, _U32, _F, _P, _I64, ctypes.c_int, _P,
                                      ctypes.c_char_p, _SZ]),
    "pgeof_autotune_clear": (None, []),
    "pgeof_set_tuning": (ctypes.c_int, [ctypes.c_char_p, _D]),
    "pgeof_get_tuning": (_D, [ctypes.c_char_p]),
    "pgeof_profile_enable": (None, [ctypes.c_int]),
    "pgeof_profile_report": (ctypes.c_int, [ctypes.c_char_p, _SZ]),
    "pgeof_profile_reset": (None, []),
    "pgeof_server_create": (_P, [ctypes.c_int, _SZ, _U32]),
    "pgeof_server_submit": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _I64, _P, _P]),
    "pgeof_server_wait": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _D]),
    "pgeof_server_footprint": (_SZ, [_I64, _U32]),
    "pgeof_server_destroy": (None, [_P]),
}


def header_symbols() -> list[str]:
    """Function names declared in include/pgeof_hip.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(pgeof_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m point_geometric_features_amd._build`")
        from ._runtime import prefer_torch_hip_runtime

        prefer_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class PgeofError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != PGEOF_OK:
        msg = lib().pgeof_last_error().decode()
        if rc == PGEOF_EINVAL:
            raise ValueError(msg)
        raise PgeofError(f"pgeof error {rc}: {msg}")


def set_tuning(key: str, value: float) -> None:
    check(lib().pgeof_set_tuning(key.encode(), float(value)))


def get_tuning(key: str) -> float:
    return float(lib().pgeof_get_tuning(key.encode()))


def autotune(workload: str, xyz, k: int, radius: float = 0.0, k_scales=(), reps: int = 3, stream=None) -> dict:
    """pgeof_autotune on a device cloud `xyz` (torch tensor (n, 3) float32):
    times the workload under a sweep of knob sets, keeps the fastest for its
    family, k and size class, and returns the JSON report as a dict."""
    import numpy as np

    sc = np.ascontiguousarray(k_scales, dtype=np.uint32)
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().pgeof_autotune(workload.encode(), _ptr(xyz), xyz.shape[0], int(k), float(radius),
                               sc.ctypes.data if len(sc) else None, len(sc), int(reps), _stream(stream), buf,
                               len(buf)))
    return json.loads(buf.value.decode())


def autotune_clear() -> None:
    lib().pgeof_autotune_clear()


def profile_enable(on: bool = True) -> None:
    lib().pgeof_profile_enable(1 if on else 0)


def profile_reset() -> None:
    lib().pgeof_profile_reset()


def profile_report() -> dict:
    L = lib()
    n = L.pgeof_profile_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.pgeof_profile_report(buf, n + 1)
    return {k: (int(v[0]), float(v[1])) for k, v in json.loads(buf.value.decode() or "{}").items()}


# ---------------------------------------------------------------- torch helpers
def _ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


def _stream(stream) -> int | None:
    if stream is None:
        import torch

        return int(torch.cuda.current_stream().cuda_stream)
    return int(getattr(stream, "cuda_stream", stream))


def knn_search_device(data, query, k, indices, sqr_dist=None, workspace=None, stream=None):
    ws, wsb = (_ptr(workspace), workspace.numel() * workspace.element_size()) if workspace is not None else (None, 0)
    check(lib().pgeof_knn_search_device(_ptr(data), data.shape[0], _ptr(query), query.shape[0], int(k),
                                        _ptr(indices), _ptr(sqr_dist), ws, wsb, _stream(stream)))


def radius_search_device(data, query, r, max_knn, indices, sqr_dist, workspace=None, stream=None):
    ws, wsb = (_ptr(workspace), workspace.numel() * workspace.element_size()) if workspace is not None else (None, 0)
    check(lib().pgeof_radius_search_device(_ptr(data), data.shape[0], _ptr(query), query.shape[0], float(r),
                                           int(max_knn), _ptr(indices), _ptr(sqr_dist), ws, wsb,
                                           _stream(stream)))


def knn_features_device(xyz, k, nn, nn_ptr, features, sqr_dist=None, workspace=None, stream=None):
    ws, wsb = (_ptr(workspace), workspace.numel() * workspace.element_size()) if workspace is not None else (None, 0)
    check(lib().pgeof_knn_features_device(_ptr(xyz), xyz.shape[0], int(k), _ptr(nn), _ptr(nn_ptr), _ptr(sqr_dist),
                                          _ptr(features), ws, wsb, _stream(stream)))


def radius_csr_device(xyz, r, max_knn, nn, nn_ptr, workspace=None, stream=None, sync_nnz=True):
    """Returns nnz (synchronising the stream), or None with sync_nnz=False
    (asynchronous: nnz = nn_ptr[n] on the device)."""
    ws, wsb = (_ptr(workspace), workspace.numel() * workspace.element_size()) if workspace is not None else (None, 0)
    nnz = _I64(0)
    check(lib().pgeof_radius_csr_device(_ptr(xyz), xyz.shape[0], float(r), int(max_knn), _ptr(nn), _ptr(nn_ptr),
                                        ctypes.byref(nnz) if sync_nnz else None, ws, wsb, _stream(stream)))
    return int(nnz.value) if sync_nnz else None


def radius_multiscale_device(xyz, r, max_knn, k_scales, nn, nn_ptr, features, workspace=None, stream=None):
    """pgeof_radius_multiscale_device: radius CSR + multiscale features, the
    features computed in the search kernel."""
    import numpy as np

    sc = np.ascontiguousarray(k_scales, dtype=np.uint32)
    ws, wsb = (_ptr(workspace), workspace.numel() * workspace.element_size()) if workspace is not None else (None, 0)
    check(lib().pgeof_radius_multiscale_device(_ptr(xyz), xyz.shape[0], float(r), int(max_knn), sc.ctypes.data,
                                               len(sc), _ptr(nn), _ptr(nn_ptr), _ptr(features), ws, wsb,
                                               _stream(stream)))


def radius_multiscale_workspace_size(n: int, max_knn: int, n_scales: int) -> int:
    return int(lib().pgeof_radius_multiscale_workspace_size(n, max_knn, n_scales))


def knn_optimal_device(xyz, k, nn, nn_ptr, features, k_min=1, k_step=1, k_min_search=1, workspace=None, stream=None):
    """pgeof_knn_optimal_device: kNN CSR + optimal-k features in one call."""
    ws, wsb = (_ptr(workspace), workspace.numel() * workspace.element_size()) if workspace is not None else (None, 0)
    check(lib().pgeof_knn_optimal_device(_ptr(xyz), xyz.shape[0], int(k), int(k_min), int(k_step), int(k_min_search),
                                         _ptr(nn), _ptr(nn_ptr), _ptr(features), ws, wsb, _stream(stream)))


def knn_optimal_workspace_size(n: int, k: int) -> int:
    return knn_workspace_size(n, n, k) + 256


def compute_features_device(xyz, nn, nn_ptr, features, k_min=1, err=None, stream=None):
    check(lib().pgeof_compute_features_device(_ptr(xyz), xyz.shape[0], _ptr(nn), nn.shape[0], _ptr(nn_ptr),
                                              nn_ptr.shape[0], int(k_min), _ptr(features), _ptr(err),
                                              _stream(stream)))


def compute_features_device_noalloc(xyz, nn, nn_ptr, features, err, k_min=1, stream=None):
    """pgeof_compute_features_device_noalloc: no allocation, no synchronisation;
    `err` (one int32/uint32 device word, zeroed by the caller) is required."""
    check(lib().pgeof_compute_features_device_noalloc(_ptr(xyz), xyz.shape[0], _ptr(nn), nn.shape[0], _ptr(nn_ptr),
                                                      nn_ptr.shape[0], int(k_min), _ptr(features), _ptr(err),
                                                      _stream(stream)))


NOALLOC_MAX_ROWS = 4194304  # PGEOF_NOALLOC_MAX_ROWS


def compute_features_multiscale_device(xyz, nn, nn_ptr, k_scales, features, err=None, stream=None):
    import numpy as np

    sc = np.ascontiguousarray(k_scales, dtype=np.uint32)
    check(lib().pgeof_compute_features_multiscale_device(_ptr(xyz), xyz.shape[0], _ptr(nn), nn.shape[0],
                                                         _ptr(nn_ptr), nn_ptr.shape[0], sc.ctypes.data, len(sc),
                                                         _ptr(features), _ptr(err), _stream(stream)))


def compute_features_optimal_device(xyz, nn, nn_ptr, features, k_min=1, k_step=1, k_min_search=1, err=None,
                                    stream=None):
    check(lib().pgeof_compute_features_optimal_device(_ptr(xyz), xyz.shape[0], _ptr(nn), nn.shape[0], _ptr(nn_ptr),
                                                      nn_ptr.shape[0], int(k_min), int(k_step), int(k_min_search),
                                                      _ptr(features), _ptr(err), _stream(stream)))


class Server:
    """ctypes wrapper of the serving loop (pgeof_server_*)."""

    def __init__(self, n_slots: int, hbm_budget_bytes: int, k: int):
        self._h = lib().pgeof_server_create(int(n_slots), int(hbm_budget_bytes), int(k))
        if not self._h:
            raise PgeofError("pgeof_server_create failed")
        self.k = int(k)
        self._keep = {}

    def submit(self, rid: int, xyz, feat_out, nn_out=None) -> None:
        """xyz, feat_out (and nn_out) are C-contiguous numpy arrays kept alive
        until the request completes."""
        self._keep[int(rid)] = (xyz, feat_out, nn_out)
        check(lib().pgeof_server_submit(self._h, int(rid), xyz.ctypes.data, len(xyz),
                                        None if nn_out is None else nn_out.ctypes.data, feat_out.ctypes.data))

    def wait(self, max_items: int = 64, timeout_ms: float = 1000.0):
        import numpy as np

        ids = np.zeros(max_items, np.uint64)
        st = np.zeros(max_items, np.int32)
        lat = np.zeros(max_items, np.float64)
        c = lib().pgeof_server_wait(self._h, max_items, ids.ctypes.data, st.ctypes.data, lat.ctypes.data,
                                    float(timeout_ms))
        out = []
        for i in range(c):
            self._keep.pop(int(ids[i]), None)
            out.append((int(ids[i]), int(st[i]), float(lat[i])))
        return out

    @staticmethod
    def footprint(n: int, k: int) -> int:
        return int(lib().pgeof_server_footprint(int(n), int(k)))

    def close(self) -> None:
        if self._h:
            lib().pgeof_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def knn_workspace_size(n_data: int, n_query: int, k: int) -> int:
    return int(lib().pgeof_knn_workspace_size(n_data, n_query, k))


def radius_workspace_size(n_data: int, n_query: int, max_knn: int) -> int:
    return int(lib().pgeof_radius_workspace_size(n_data, n_query, max_knn))


def radius_csr_workspace_size(n: int, max_knn: int) -> int:
    return int(lib().pgeof_radius_csr_workspace_size(n, max_knn))
