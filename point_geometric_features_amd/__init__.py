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
    from .pgeof_ext import (  # noqa: F401
        EFeatureID,
        compute_features,
        compute_features_multiscale,
        compute_features_optimal,
        compute_features_selected,
        knn_search,
        radius_search,
    )
    from .pgeof_ext import *  # noqa: F401,F403  (EFeatureID values, as export_values())
except ImportError as _e:  # no build yet: fail on first use, not on import
    _missing = ImportError(
        "point_geometric_features_amd: the gfx950 extension is not built "
        f"({_e}); run `python -m point_geometric_features_amd._build`"
    )

    def __getattr__(name: str):
        if name.startswith("__"):
            raise AttributeError(name)
        raise _missing
