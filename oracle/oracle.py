"""CPU oracle for pgeof's hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker.  The product
(``point_geometric_features_amd``) never imports it.

Two independent restatements of the reference live here:

* ``lib()`` -- ctypes handle on ``oracle/libpgeof_oracle.so`` (built from
  ``oracle/pgeof_oracle.c`` by ``oracle/Makefile``): FP64 two-pass covariance +
  cyclic Jacobi eigensolver, KD-tree / brute-force neighbour search with the
  reference's float32 distance expression.
* the ``*_np`` functions -- numpy restatements of ``include/pca.hpp:71-200``
  and ``include/pgeof.hpp:75-310`` using ``numpy.linalg.eigh`` (LAPACK
  ``syevd``), which also returns eigenvalues in ascending order like Eigen's
  ``SelfAdjointEigenSolver``.

The tests pin each against the other, against closed forms and against
scipy's KD-tree (the reference's own comparator, ``tests/test_pgeof.py:8-27``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

EPS = 1e-3


def build() -> str:
    """Compile the C oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libpgeof_oracle.so")


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpgeof_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64, u32, i32 = ctypes.c_int64, ctypes.c_uint32, ctypes.c_int
        L.oracle_pca.argtypes = [P, P, i64, P, P]
        L.oracle_eig3.argtypes = [P, P, P]
        L.oracle_compute_features.argtypes = [P, i64, P, i64, P, i64, u32, P, i32]
        L.oracle_compute_features_multiscale.argtypes = [P, i64, P, i64, P, i64, P, i64, P, i32]
        L.oracle_compute_features_optimal.argtypes = [P, i64, P, i64, P, i64, u32, u32, u32, i32, P, i32]
        L.oracle_knn.argtypes = [P, i64, P, i64, u32, i32, P, P, i32]
        L.oracle_radius.argtypes = [P, i64, P, i64, ctypes.c_float, u32, i32, P, P, P, i32]
        L.oracle_compute_features_selected.argtypes = [P, i32, i64, ctypes.c_double, u32, P, i64, P]
        L.oracle_compute_features_selected_rows.argtypes = [P, i32, i64, ctypes.c_double, u32, P, i64, P, i64, P]
        L.oracle_knn_features_timed.argtypes = [P, i64, P, i64, u32, i32, P, P,
                                                ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_double)]
        for f in ("oracle_compute_features", "oracle_compute_features_multiscale",
                  "oracle_compute_features_optimal", "oracle_knn", "oracle_radius",
                  "oracle_compute_features_selected", "oracle_compute_features_selected_rows",
                  "oracle_knn_features_timed"):
            getattr(L, f).restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _threads(n_threads):
    if n_threads is None:
        n_threads = min(16, len(os.sched_getaffinity(0)))
    return int(n_threads)


class OracleError(ValueError):
    pass


def _check(rc: int):
    if rc == 1:
        raise OracleError("reference ValueError condition")
    if rc == 2:
        raise OracleError("invalid CSR / argument rejected")


# --------------------------------------------------------------------------
# C oracle wrappers
# --------------------------------------------------------------------------

def eig3(cov6) -> tuple[np.ndarray, np.ndarray]:
    """Ascending eigenvalues (clamped >= 0) and eigenvectors (rows) of a packed
    3x3 covariance, v2 oriented to z >= 0 (pca.hpp:79-96)."""
    c = np.ascontiguousarray(cov6, dtype=np.float64)
    val = np.zeros(3)
    vec = np.zeros(9)
    lib().oracle_eig3(_p(c), _p(val), _p(vec))
    return val, vec.reshape(3, 3)


def pca(points: np.ndarray):
    pts = np.ascontiguousarray(points, dtype=np.float32)
    idx = np.arange(len(pts), dtype=np.uint32)
    val = np.zeros(3)
    vec = np.zeros(9)
    lib().oracle_pca(_p(pts), _p(idx), len(pts), _p(val), _p(vec))
    return val, vec.reshape(3, 3)


def compute_features(xyz, nn, nn_ptr, k_min=1, n_threads=None) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    nn = np.ascontiguousarray(nn, dtype=np.uint32)
    nn_ptr = np.ascontiguousarray(nn_ptr, dtype=np.uint32)
    out = np.zeros((max(len(nn_ptr) - 1, 0), 11))
    _check(lib().oracle_compute_features(_p(xyz), len(xyz), _p(nn), len(nn), _p(nn_ptr), len(nn_ptr),
                                         int(k_min), _p(out), _threads(n_threads)))
    return out


def compute_features_multiscale(xyz, nn, nn_ptr, k_scales, n_threads=None) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    nn = np.ascontiguousarray(nn, dtype=np.uint32)
    nn_ptr = np.ascontiguousarray(nn_ptr, dtype=np.uint32)
    sc = np.ascontiguousarray(k_scales, dtype=np.uint32)
    out = np.zeros((max(len(nn_ptr) - 1, 0), len(sc), 11))
    _check(lib().oracle_compute_features_multiscale(_p(xyz), len(xyz), _p(nn), len(nn), _p(nn_ptr),
                                                    len(nn_ptr), _p(sc), len(sc), _p(out),
                                                    _threads(n_threads)))
    return out


def compute_features_optimal(xyz, nn, nn_ptr, k_min=1, k_step=1, k_min_search=1, eigentropy_fp32=True,
                             n_threads=None) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    nn = np.ascontiguousarray(nn, dtype=np.uint32)
    nn_ptr = np.ascontiguousarray(nn_ptr, dtype=np.uint32)
    out = np.zeros((max(len(nn_ptr) - 1, 0), 12))
    _check(lib().oracle_compute_features_optimal(_p(xyz), len(xyz), _p(nn), len(nn), _p(nn_ptr), len(nn_ptr),
                                                 int(k_min), int(k_step), int(k_min_search),
                                                 int(bool(eigentropy_fp32)), _p(out), _threads(n_threads)))
    return out


def knn(data, query, k, kdtree=True, n_threads=None):
    data = np.ascontiguousarray(data, dtype=np.float32)
    query = np.ascontiguousarray(query, dtype=np.float32)
    idx = np.zeros((len(query), k), dtype=np.uint32)
    d = np.zeros((len(query), k), dtype=np.float32)
    _check(lib().oracle_knn(_p(data), len(data), _p(query), len(query), int(k), int(bool(kdtree)),
                            _p(idx), _p(d), _threads(n_threads)))
    return idx, d


def radius(data, query, r, max_knn, kdtree=True, n_threads=None):
    data = np.ascontiguousarray(data, dtype=np.float32)
    query = np.ascontiguousarray(query, dtype=np.float32)
    idx = np.zeros((len(query), max_knn), dtype=np.int32)
    d = np.zeros((len(query), max_knn), dtype=np.float32)
    cnt = np.zeros(len(query), dtype=np.int32)
    _check(lib().oracle_radius(_p(data), len(data), _p(query), len(query), float(r), int(max_knn),
                               int(bool(kdtree)), _p(idx), _p(d), _p(cnt), _threads(n_threads)))
    return idx, d, cnt


def compute_features_selected(xyz, r, max_knn, fids, rows=None) -> np.ndarray:
    """include/pgeof.hpp:325-375 (brute-force radius search in the input
    precision); for the given rows only when `rows` is set."""
    is64 = np.asarray(xyz).dtype == np.float64
    x = np.ascontiguousarray(xyz, dtype=np.float64 if is64 else np.float32)
    f = np.ascontiguousarray([int(v) for v in fids], dtype=np.int32)
    if rows is None:
        out = np.zeros((len(x), len(f)))
        lib().oracle_compute_features_selected(_p(x), int(is64), len(x), float(r), int(max_knn), _p(f), len(f),
                                               _p(out))
        return out
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    assert rows.size == 0 or int(rows.max()) < len(x)
    out = np.zeros((len(rows), len(f)))
    lib().oracle_compute_features_selected_rows(_p(x), int(is64), len(x), float(r), int(max_knn), _p(f), len(f),
                                                _p(rows), len(rows), _p(out))
    return out


def knn_features_timed(xyz, sample, k, n_threads=None):
    """KD-tree over all of xyz + kNN and FP64 features for xyz[sample]."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    feats = np.zeros((len(sample), 11))
    idx = np.zeros((len(sample), k), dtype=np.uint32)
    tb, tq = ctypes.c_double(), ctypes.c_double()
    _check(lib().oracle_knn_features_timed(_p(xyz), len(xyz), _p(sample), len(sample), int(k),
                                           _threads(n_threads), _p(feats), _p(idx), ctypes.byref(tb),
                                           ctypes.byref(tq)))
    return feats, idx, tb.value, tq.value


# --------------------------------------------------------------------------
# numpy restatement (independent eigen-solver: LAPACK via numpy.linalg.eigh)
# --------------------------------------------------------------------------

def orient_v2_np(v2: np.ndarray) -> np.ndarray:
    """pca.hpp:96 plus this build's convention for v2.z == 0 (unpinned)."""
    v2 = v2.copy()
    flip = (v2[:, 2] < 0) | ((v2[:, 2] == 0) & ((v2[:, 1] < 0) | ((v2[:, 1] == 0) & (v2[:, 0] < 0))))
    v2[flip] *= -1
    return v2


def features_from_eig_np(val: np.ndarray, vec: np.ndarray) -> np.ndarray:
    """pca.hpp:160-200 vectorised.  val (n,3) ascending clamped; vec (n,3,3)
    with vec[:, i] the i-th eigenvector."""
    n = len(val)
    f = np.zeros((n, 11))
    s = np.sqrt(val)
    fact = 1.0 / (s[:, 0] + EPS)
    f[:, 4:7] = vec[:, 2]
    f[:, 0] = (s[:, 0] - s[:, 1]) * fact
    f[:, 1] = (s[:, 1] - s[:, 2]) * fact
    f[:, 2] = s[:, 2] * fact
    f[:, 7] = s[:, 0]
    f[:, 8] = np.sqrt(s[:, 0] * s[:, 1] + 1e-6)
    f[:, 9] = np.power(s[:, 0] * s[:, 1] * s[:, 2] + 1e-9, 1.0 / 3.0)
    f[:, 10] = s[:, 2] / (s[:, 0] + s[:, 1] + s[:, 2] + EPS)
    u = np.einsum("ni,nic->nc", val, np.abs(vec))
    nz = s[:, 0] > 0
    f[nz, 3] = u[nz, 2] / np.linalg.norm(u[nz], axis=1)
    return f


def pca_batch_np(clouds: np.ndarray):
    """PCA of a batch of equal-size neighbourhoods (m, k, 3) in FP64."""
    c = clouds.astype(np.float64)
    d = c - c.mean(axis=1, keepdims=True)
    cov = np.einsum("mki,mkj->mij", d, d) / c.shape[1]
    w, v = np.linalg.eigh(cov)
    w = np.maximum(w, 0.0)
    vec = np.transpose(v, (0, 2, 1)).copy()  # vec[:, i] = eigenvector i
    vec[:, 2] = orient_v2_np(vec[:, 2])
    return w, vec


def compute_features_np(xyz, nn, nn_ptr, k_min=1) -> np.ndarray:
    """pgeof.hpp:75-117 over a CSR; rows grouped by length for vectorisation."""
    xyz = np.asarray(xyz, dtype=np.float32)
    nn = np.asarray(nn, dtype=np.int64)
    ptr = np.asarray(nn_ptr, dtype=np.int64)
    n = len(ptr) - 1
    out = np.zeros((n, 11))
    k = ptr[1:] - ptr[:-1]
    for kk in np.unique(k):
        if kk < k_min or kk == 0:
            continue
        rows = np.nonzero(k == kk)[0]
        gather = nn[ptr[rows][:, None] + np.arange(kk)[None, :]]
        w, vec = pca_batch_np(xyz[gather])
        out[rows] = features_from_eig_np(w, vec)
    return out


def eigentropy_np(val: np.ndarray) -> np.ndarray:
    """pca.hpp:140-150 in FP64."""
    s = val.sum(axis=1, keepdims=True) + EPS
    e = val / s
    return -(e * np.log(e + EPS)).sum(axis=1)


def eigentropy32_np(val: np.ndarray) -> np.ndarray:
    """pca.hpp:140-150 in FP32, term order as written in the reference."""
    v = val.astype(np.float32)
    eps = np.float32(1e-3)
    s = v[:, 0] + v[:, 1] + v[:, 2] + eps
    e = v / s[:, None]
    return (-e[:, 0] * np.log(e[:, 0] + eps) - e[:, 1] * np.log(e[:, 1] + eps)
            - e[:, 2] * np.log(e[:, 2] + eps))


def eigvals_csr_np(xyz, nn, nn_ptr, prefix=None) -> np.ndarray:
    """FP64 ascending covariance eigenvalues of every CSR row (or of its first
    `prefix` entries); rows with no neighbour get zeros."""
    xyz = np.asarray(xyz, dtype=np.float32)
    nn = np.asarray(nn, dtype=np.int64)
    ptr = np.asarray(nn_ptr, dtype=np.int64)
    n = len(ptr) - 1
    k = ptr[1:] - ptr[:-1]
    if prefix is not None:
        k = np.minimum(k, prefix)
    out = np.zeros((n, 3))
    for kk in np.unique(k):
        if kk == 0:
            continue
        rows = np.nonzero(k == kk)[0]
        gather = nn[ptr[rows][:, None] + np.arange(kk)[None, :]]
        c = xyz[gather].astype(np.float64)
        d = c - c.mean(axis=1, keepdims=True)
        cov = np.einsum("mki,mkj->mij", d, d) / kk
        out[rows] = np.linalg.eigvalsh(cov)
    return out
