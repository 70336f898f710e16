"""Shared parity checks between the GPU path and the CPU oracle.

Feature tolerance (vs the FP64 oracle): |gpu - oracle| <= 1e-4 + 1e-5 * |oracle|
(1e-4 absolute, plus a few float32 ulps for the large values that
1 / (sqrt(lambda_min) + 1e-3) produces on flat neighbourhoods).

Entries the reference itself does not determine are excluded and counted:
  * VerticalityPGEOF (col 3) on numerically rank-deficient rows
    (lambda_min <= 1e-9 lambda_max): whether it is computed at all depends on
    the sign of a rounding-level lambda_min (include/pca.hpp:189);
    columns 0-2 there are compared with rtol 1e-3;
  * the normal (cols 4-6) and VerticalityPGEOF when the two largest
    eigenvalues are within 1e-6 lambda_max (eigenvector not unique), and
    VerticalityPGEOF when the two smallest are (v0/v1 basis not unique).
"""
from __future__ import annotations

import numpy as np

ATOL = 1e-4
RTOL = 1e-5


def unpinned_masks(val: np.ndarray):
    l0, l1, l2 = val[:, 0], val[:, 1], val[:, 2]
    scale = np.maximum(l2, 1e-300)
    rankdef = l0 <= 1e-9 * scale
    top = (l2 - l1) <= 1e-6 * scale
    bottom = (l1 - l0) <= 1e-6 * scale
    return rankdef, top, bottom


def compare_features(gpu: np.ndarray, ref: np.ndarray, val: np.ndarray, cols: int = 11):
    """Returns (n_bad, max_err_per_col, n_excluded) for (n, >=11) feature arrays."""
    g = gpu[:, :11].astype(np.float64)
    r = ref[:, :11].astype(np.float64)
    err = np.abs(g - r)
    tol = ATOL + RTOL * np.abs(r)
    rankdef, top, bottom = unpinned_masks(val)
    skip = np.zeros_like(err, dtype=bool)
    skip[rankdef, 3] = True
    skip[top, 3:7] = True
    skip[bottom & ~rankdef, 3] = True
    tol[rankdef, 0:3] = ATOL + 1e-3 * np.abs(r[rankdef, 0:3])
    bad = (err > tol) & ~skip
    masked = np.where(skip, 0.0, err)
    return int(bad.sum()), masked.max(axis=0) if len(masked) else np.zeros(11), int(skip.sum())


def assert_features_close(gpu, ref, val, what=""):
    n_bad, max_err, n_skip = compare_features(gpu, ref, val)
    if n_bad:
        g = gpu[:, :11].astype(np.float64)
        r = ref[:, :11].astype(np.float64)
        rankdef, top, bottom = unpinned_masks(val)
        err = np.abs(g - r) > (ATOL + RTOL * np.abs(r))
        rows = np.nonzero(err.any(axis=1))[0][:5]
        detail = "\n".join(f"row {i}: gpu {gpu[i, :11]} ref {ref[i, :11]} val {val[i]}" for i in rows)
        raise AssertionError(f"{what}: {n_bad} feature entries beyond tolerance; max err {max_err}\n{detail}")
    return max_err, n_skip
