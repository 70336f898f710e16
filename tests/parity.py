"""Shared parity checks between the GPU path and the CPU oracle.

Feature tolerance (vs the FP64 oracle): |gpu - oracle| <= max(1e-4, 1e-5 |oracle|),
i.e. 1e-4 absolute wherever float32 can hold it; above |oracle| = 10 (flat
neighbourhoods, where 1 / (sqrt(lambda_min) + 1e-3) is large) the bound is
1e-5 relative, since a float32 ulp of 2^13 already exceeds 1e-4.

Entries the reference itself does not determine are excluded and counted:
  * VerticalityPGEOF (col 3) on numerically rank-deficient rows
    (lambda_min <= 1e-9 lambda_max): whether it is computed at all depends on
    the sign of a rounding-level lambda_min (include/pca.hpp:189);
    columns 0-2 there are compared with a relative bound of RTOL_RANKDEF,
    since sqrt(lambda_min) is itself a rounding-level number;
  * the normal (cols 4-6) and VerticalityPGEOF when the two largest
    eigenvalues are within 1e-6 lambda_max (eigenvector not unique), and
    VerticalityPGEOF when the two smallest are (v0/v1 basis not unique).
Callers bound the number of excluded entries (``max_skip_frac``).

Set PGEOF_PARITY_LOG to a file path to append one JSON line of error
statistics per comparison.
"""
from __future__ import annotations

import json
import os

import numpy as np

ATOL = 1e-4
RTOL = 1e-5
RTOL_RANKDEF = 1e-4


def unpinned_masks(val: np.ndarray):
    l0, l1, l2 = val[:, 0], val[:, 1], val[:, 2]
    scale = np.maximum(l2, 1e-300)
    rankdef = l0 <= 1e-9 * scale
    top = (l2 - l1) <= 1e-6 * scale
    bottom = (l1 - l0) <= 1e-6 * scale
    return rankdef, top, bottom


def tolerance(ref: np.ndarray, val: np.ndarray) -> np.ndarray:
    r = np.abs(ref[:, :11].astype(np.float64))
    tol = np.maximum(ATOL, RTOL * r)
    rankdef, _, _ = unpinned_masks(val)
    tol[rankdef, 0:3] = np.maximum(ATOL, RTOL_RANKDEF * r[rankdef, 0:3])
    return tol


def skip_mask(val: np.ndarray, shape) -> np.ndarray:
    rankdef, top, bottom = unpinned_masks(val)
    skip = np.zeros(shape, dtype=bool)
    skip[rankdef, 3] = True
    skip[top, 3:7] = True
    skip[bottom & ~rankdef, 3] = True
    return skip


def compare_features(gpu: np.ndarray, ref: np.ndarray, val: np.ndarray):
    """Returns (n_bad, max_err_per_col, n_excluded, worst tolerance ratio per
    col) for (n, >=11) feature arrays."""
    g = gpu[:, :11].astype(np.float64)
    r = ref[:, :11].astype(np.float64)
    err = np.abs(g - r)
    tol = tolerance(ref, val)
    skip = skip_mask(val, err.shape)
    bad = (err > tol) & ~skip
    masked = np.where(skip, 0.0, err)
    ratio = np.where(skip, 0.0, err / tol)
    if len(masked) == 0:
        return 0, np.zeros(11), 0, np.zeros(11)
    return int(bad.sum()), masked.max(axis=0), int(skip.sum()), ratio.max(axis=0)


def assert_features_close(gpu, ref, val, what="", max_skip_frac=0.05):
    n_bad, max_err, n_skip, ratio = compare_features(gpu, ref, val)
    log = os.environ.get("PGEOF_PARITY_LOG")
    if log:
        with open(log, "a") as fh:
            fh.write(json.dumps({"what": what, "rows": int(len(gpu)), "n_bad": n_bad, "n_skip": n_skip,
                                 "max_err": [float(x) for x in max_err],
                                 "tol_ratio": [float(x) for x in ratio]}) + "\n")
    if n_bad:
        g = gpu[:, :11].astype(np.float64)
        r = ref[:, :11].astype(np.float64)
        err = np.abs(g - r) > tolerance(ref, val)
        err &= ~skip_mask(val, err.shape)
        rows = np.nonzero(err.any(axis=1))[0][:5]
        detail = "\n".join(f"row {i}: gpu {gpu[i, :11]} ref {ref[i, :11]} val {val[i]}" for i in rows)
        raise AssertionError(f"{what}: {n_bad} feature entries beyond tolerance; max err {max_err}\n{detail}")
    n_entries = max(1, gpu.shape[0] * 11)
    assert n_skip <= max_skip_frac * n_entries, f"{what}: {n_skip} of {n_entries} entries excluded"
    return max_err, n_skip


def optimal_disagreement_is_tie(xyz, nn, k_row, rows, k_gpu, k_ref, tol=2e-6) -> bool:
    """Rows where the GPU's optimal k differs from the oracle's: the FP32
    eigentropies (include/pca.hpp:140-150) of the FP64 eigenvalues at the two
    k must differ by no more than `tol` (a few float32 ulps of values ~1),
    i.e. the argmin is a near-tie that float rounding may resolve either
    way.  `nn` holds rows of `k_row` neighbours."""
    from oracle import oracle as O

    if len(rows) == 0:
        return True
    nn2 = nn.reshape(-1, k_row)
    h = []
    for ks in (np.asarray(k_gpu, np.int64), np.asarray(k_ref, np.int64)):
        hv = np.zeros(len(rows))
        for kk in np.unique(ks):
            sel = ks == kk
            r = np.asarray(rows)[sel]
            sub_ptr = np.r_[0, np.full(len(r), kk).cumsum()].astype(np.uint32)
            sub_nn = nn2[r, :kk].ravel().copy()
            hv[sel] = O.eigentropy32_np(O.eigvals_csr_np(xyz, sub_nn, sub_ptr))
        h.append(hv)
    return bool((np.abs(h[0] - h[1]) <= tol).all())
