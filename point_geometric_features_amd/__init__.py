"""MI355X (gfx950) implementation of pgeof's hot path.

Same import surface as the reference package (``src/pgeof/__init__.py:1-9``):
``EFeatureID, compute_features, compute_features_multiscale,
compute_features_optimal, knn_search, radius_search, compute_features_selected``.
The functions run hand-written HIP kernels through ``libpgeof_hip.so``; there is
no CPU fallback.  Without a build, importing the package still works (so that
its build recipe, ``point_geometric_features_amd._build``, can run from a
clean checkout), but every use of a pgeof name raises ImportError.
"""
from __future__ import annotations

import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

from ._runtime import prefer_torch_hip_runtime as _prefer_runtime

_prefer_runtime()

__all__ = [
    "EFeatureID",
    "compute_features",
    "compute_features_multiscale",
    "compute_features_optimal",
    "knn_search",
    "radius_search",
    "compute_features_selected",
]

try:
    from . import pgeof_ext as _ext
    from .pgeof_ext import EFeatureID  # noqa: F401
except ImportError as _e:  # no build yet: fail on first use, not on import
    _missing = ImportError(
        "point_geometric_features_amd: the gfx950 extension is not built "
        f"({_e}); run `python -m point_geometric_features_amd._build`"
    )

    def __getattr__(name: str):
        if name.startswith("__"):
            raise AttributeError(name)
        raise _missing
else:
    from . import _device

    # EFeatureID values at module level, as the reference's export_values()
    globals().update({_k: _v for _k, _v in EFeatureID.__members__.items()})

    # Host numpy arrays go to the extension module (the reference's binding
    # surface); arrays on the GPU (DLPack: torch tensors, CuPy arrays, ...)
    # go to the device entry points and come back as torch tensors.
    def compute_features(xyz, nn, nn_ptr, k_min=1, verbose=False):
        if _device.is_device_array(xyz) or _device.is_device_array(nn) or _device.is_device_array(nn_ptr):
            return _device.compute_features(xyz, nn, nn_ptr, k_min, verbose)
        return _ext.compute_features(xyz, nn, nn_ptr, k_min, verbose)

    def compute_features_multiscale(xyz, nn, nn_ptr, k_scales, verbose=False):
        if _device.is_device_array(xyz) or _device.is_device_array(nn) or _device.is_device_array(nn_ptr):
            return _device.compute_features_multiscale(xyz, nn, nn_ptr, k_scales, verbose)
        return _ext.compute_features_multiscale(xyz, nn, nn_ptr, k_scales, verbose)

    def compute_features_optimal(xyz, nn, nn_ptr, k_min=1, k_step=1, k_min_search=1, verbose=False):
        if _device.is_device_array(xyz) or _device.is_device_array(nn) or _device.is_device_array(nn_ptr):
            return _device.compute_features_optimal(xyz, nn, nn_ptr, k_min, k_step, k_min_search, verbose)
        return _ext.compute_features_optimal(xyz, nn, nn_ptr, k_min, k_step, k_min_search, verbose)

    def knn_search(data, query, knn):
        if _device.is_device_array(data) or _device.is_device_array(query):
            return _device.knn_search(data, query, knn)
        return _ext.knn_search(data, query, knn)

    def radius_search(data, query, search_radius, max_knn):
        if _device.is_device_array(data) or _device.is_device_array(query):
            return _device.radius_search(data, query, search_radius, max_knn)
        return _ext.radius_search(data, query, search_radius, max_knn)

    def compute_features_selected(xyz, search_radius, max_knn, selected_features):
        if _device.is_device_array(xyz):
            return _device.compute_features_selected(_ext.compute_features_selected, xyz, search_radius, max_knn,
                                                     selected_features)
        return _ext.compute_features_selected(xyz, search_radius, max_knn, selected_features)

    for _f, _src in ((compute_features, _ext.compute_features),
                     (compute_features_multiscale, _ext.compute_features_multiscale),
                     (compute_features_optimal, _ext.compute_features_optimal), (knn_search, _ext.knn_search),
                     (radius_search, _ext.radius_search), (compute_features_selected, _ext.compute_features_selected)):
        _f.__doc__ = (_src.__doc__ or "") + "\n\nDevice arrays (DLPack) run on the GPU and return torch tensors."

