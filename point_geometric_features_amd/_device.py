"""Device-array forms of the pgeof functions (DLPack).

The reference binds its arrays with ``nb::ndarray`` (include/pgeof.hpp:78-79),
which admits foreign arrays; this build accepts arrays that live on the GPU
(any object with ``__dlpack__`` / ``__dlpack_device__`` on a ROCm or CUDA
device: torch tensors, CuPy arrays, ...) and runs the device entry points of
``libpgeof_hip.so`` on them, with no host round trip.  Results are torch
tensors on the same device, computed on torch's current stream (the call
returns once the work is queued).  Dtype rules are those of the host
functions (float32 xyz of shape (n, 3), no implicit conversion -> TypeError),
except that the uint32 CSR arrays may also be int32 (the same bits; torch has
few uint32 operations).  compute_features_selected on device arrays goes
through host memory.
"""
from __future__ import annotations

import numpy as np

_DEVICE_TYPES = (2, 10)  # kDLCUDA, kDLROCM


def is_device_array(a) -> bool:
    if a is None or isinstance(a, np.ndarray):
        return False
    dev = getattr(a, "__dlpack_device__", None)
    if dev is None:
        return False
    try:
        return int(dev()[0]) in _DEVICE_TYPES
    except Exception:  # noqa: BLE001 (an object whose device cannot be read is not a device array)
        return False


def _tensor(a, name):
    import torch

    t = a if isinstance(a, torch.Tensor) else torch.from_dlpack(a)
    if t.device.type != "cuda":
        raise TypeError(f"incompatible function arguments: '{name}' must be a device array")
    return t


def _xyz(a, name="xyz"):
    import torch

    t = _tensor(a, name)
    if t.dtype != torch.float32:
        raise TypeError(f"incompatible function arguments: '{name}' must have dtype float32, got {t.dtype}")
    if t.dim() != 2 or t.shape[1] != 3:
        raise TypeError(f"incompatible function arguments: '{name}' must have shape (n, 3)")
    return t.contiguous()


def _u32(a, name):
    import torch

    t = _tensor(a, name)
    u32 = getattr(torch, "uint32", None)
    if t.dtype not in (torch.int32,) + ((u32,) if u32 is not None else ()):
        raise TypeError(f"incompatible function arguments: '{name}' must have dtype uint32 (or int32), got {t.dtype}")
    if t.dim() != 1:
        raise TypeError(f"incompatible function arguments: '{name}' must be 1-dimensional")
    return t.contiguous()


def _same_device(*ts):
    devs = {t.device for t in ts}
    if len(devs) != 1:
        raise TypeError(f"incompatible function arguments: the arrays live on different devices {sorted(map(str, devs))}")
    return ts[0].device


def _as_u32(t):
    import torch

    u32 = getattr(torch, "uint32", None)
    return t.view(u32) if u32 is not None else t


def _done(verbose):
    if verbose:
        print("100% done")


def compute_features(xyz, nn, nn_ptr, k_min=1, verbose=False):
    import torch

    from . import capi

    x, a, p = _xyz(xyz), _u32(nn, "nn"), _u32(nn_ptr, "nn_ptr")
    dev = _same_device(x, a, p)
    if p.shape[0] < 1:
        raise ValueError("nn_ptr must hold at least one entry")
    with torch.cuda.device(dev):
        out = torch.empty((p.shape[0] - 1, 11), dtype=torch.float32, device=dev)
        capi.compute_features_device(x, a, p, out, k_min=int(k_min))
    _done(verbose)
    return out


def compute_features_multiscale(xyz, nn, nn_ptr, k_scales, verbose=False):
    import torch

    from . import capi

    x, a, p = _xyz(xyz), _u32(nn, "nn"), _u32(nn_ptr, "nn_ptr")
    dev = _same_device(x, a, p)
    sc = [int(s) for s in k_scales]
    if p.shape[0] < 1:
        raise ValueError("nn_ptr must hold at least one entry")
    with torch.cuda.device(dev):
        out = torch.empty((p.shape[0] - 1, len(sc), 11), dtype=torch.float32, device=dev)
        capi.compute_features_multiscale_device(x, a, p, sc, out)
    _done(verbose)
    return out


def compute_features_optimal(xyz, nn, nn_ptr, k_min=1, k_step=1, k_min_search=1, verbose=False):
    import torch

    from . import capi

    x, a, p = _xyz(xyz), _u32(nn, "nn"), _u32(nn_ptr, "nn_ptr")
    dev = _same_device(x, a, p)
    if p.shape[0] < 1:
        raise ValueError("nn_ptr must hold at least one entry")
    with torch.cuda.device(dev):
        out = torch.empty((p.shape[0] - 1, 12), dtype=torch.float32, device=dev)
        capi.compute_features_optimal_device(x, a, p, out, int(k_min), int(k_step), int(k_min_search))
    _done(verbose)
    return out


def knn_search(data, query, knn):
    import torch

    from . import capi

    d, q = _xyz(data, "data"), _xyz(query, "query")
    dev = _same_device(d, q)
    with torch.cuda.device(dev):
        idx = torch.empty((q.shape[0], int(knn)), dtype=torch.int32, device=dev)
        d2 = torch.empty((q.shape[0], int(knn)), dtype=torch.float32, device=dev)
        # the same storage for data and query selects the self-search path
        qq = d if (q.data_ptr() == d.data_ptr() and q.shape == d.shape) else q
        capi.knn_search_device(d, qq, int(knn), idx, d2)
    return _as_u32(idx), d2


def radius_search(data, query, search_radius, max_knn):
    import torch

    from . import capi

    d, q = _xyz(data, "data"), _xyz(query, "query")
    dev = _same_device(d, q)
    with torch.cuda.device(dev):
        idx = torch.empty((q.shape[0], int(max_knn)), dtype=torch.int32, device=dev)
        d2 = torch.empty((q.shape[0], int(max_knn)), dtype=torch.float32, device=dev)
        qq = d if (q.data_ptr() == d.data_ptr() and q.shape == d.shape) else q
        capi.radius_search_device(d, qq, float(search_radius), int(max_knn), idx, d2)
    return idx, d2


def compute_features_selected(host_fn, xyz, search_radius, max_knn, selected_features):
    import torch

    t = _tensor(xyz, "xyz")
    out = host_fn(t.detach().cpu().numpy(), search_radius, max_knn, selected_features)
    return torch.from_numpy(out).to(t.device)
