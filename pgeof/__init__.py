"""Drop-in import name of the reference package.

The reference is imported as ``pgeof`` (``src/pgeof/__init__.py:1-9``), which
re-exports its extension module's names.  This package does the same over the
gfx950 implementation, so ``import pgeof`` works unchanged for code written
against the reference.
"""
import point_geometric_features_amd as _impl
from point_geometric_features_amd import (  # noqa: F401
    EFeatureID,
    compute_features,
    compute_features_multiscale,
    compute_features_optimal,
    compute_features_selected,
    knn_search,
    radius_search,
)

globals().update({_k: _v for _k, _v in EFeatureID.__members__.items()})  # export_values()

__all__ = [
    "EFeatureID",
    "compute_features",
    "compute_features_multiscale",
    "compute_features_optimal",
    "knn_search",
    "radius_search",
    "compute_features_selected",
]
