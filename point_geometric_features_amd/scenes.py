"""Seeded synthetic point clouds for the benchmark workloads and tests.

All generators use numpy's PCG64 ``default_rng(seed)`` and return C-contiguous
float32 (n, 3) arrays; the same numpy build reproduces them bit for bit.
Points are shuffled, as a LiDAR scan is not stored in spatial order.

* ``uniform_cube``     n points uniform in [0, 1)^3 (the reference's README example)
* ``lidar_scene``      200 m x 200 m street scene: ground plane z = 0 with
                       N(0, 0.02) noise (50 %), vertical walls (20 %), cylinders
                       (poles/trunks, 10 %), Gaussian vegetation blobs (20 %)
* ``precision_scene``  exactly planar patches (axis-aligned and tilted), exactly
                       collinear segments and isotropic blobs, offset by +1e5 in
                       x and y before the float32 cast
"""
from __future__ import annotations

import numpy as np


def uniform_cube(n: int, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).random((n, 3), dtype=np.float32)


def _split(n: int, fractions) -> list[int]:
    counts = [int(n * f) for f in fractions]
    counts[0] += n - sum(counts)
    return counts


def _alloc(rng, total: int, weights: np.ndarray) -> np.ndarray:
    """Split `total` points over objects proportionally to `weights`."""
    return rng.multinomial(total, weights / weights.sum()).astype(np.int64)


def lidar_scene(n: int, seed: int = 1, extent: float = 200.0) -> np.ndarray:
    """LiDAR-like scene; object counts scale with the area (40 walls, 400
    cylinders and 1000 blobs at 200 m x 200 m)."""
    rng = np.random.default_rng(seed)
    n_ground, n_wall, n_cyl, n_blob = _split(n, [0.5, 0.2, 0.1, 0.2])
    area = (extent / 200.0) ** 2
    parts = []

    # ground: z = 0 + N(0, 0.02)
    g = np.empty((n_ground, 3))
    g[:, :2] = rng.uniform(0.0, extent, size=(n_ground, 2))
    g[:, 2] = rng.normal(0.0, 0.02, size=n_ground)
    parts.append(g)

    # vertical walls: length U(5,30) m, height U(2,10) m, 1 cm noise along the normal
    nw = max(1, int(round(40 * area)))
    L = rng.uniform(5.0, 30.0, nw)
    H = rng.uniform(2.0, 10.0, nw)
    cnt = _alloc(rng, n_wall, L * H)
    o = rng.uniform(0.0, extent, size=(nw, 2))
    th = rng.uniform(0.0, np.pi, nw)
    wid = np.repeat(np.arange(nw), cnt)
    s = rng.uniform(0.0, 1.0, n_wall) * L[wid]
    z = rng.uniform(0.0, 1.0, n_wall) * H[wid]
    e = rng.normal(0.0, 0.01, n_wall)
    d = np.stack([np.cos(th), np.sin(th)], 1)[wid]
    nrm = np.stack([-d[:, 1], d[:, 0]], 1)
    w = np.empty((n_wall, 3))
    w[:, :2] = o[wid] + d * s[:, None] + nrm * e[:, None]
    w[:, 2] = z
    parts.append(w)

    # cylinders: r U(0.1,0.5) m, height U(2,8) m, 5 mm radial noise
    nc = max(1, int(round(400 * area)))
    R = rng.uniform(0.1, 0.5, nc)
    Hc = rng.uniform(2.0, 8.0, nc)
    cnt = _alloc(rng, n_cyl, R * Hc)
    cc = rng.uniform(0.0, extent, size=(nc, 2))
    cid = np.repeat(np.arange(nc), cnt)
    a = rng.uniform(0.0, 2 * np.pi, n_cyl)
    rr = R[cid] + rng.normal(0.0, 0.005, n_cyl)
    c = np.empty((n_cyl, 3))
    c[:, 0] = cc[cid, 0] + rr * np.cos(a)
    c[:, 1] = cc[cid, 1] + rr * np.sin(a)
    c[:, 2] = rng.uniform(0.0, 1.0, n_cyl) * Hc[cid]
    parts.append(c)

    # vegetation blobs: sigma U(0.3,1.5) m, centre height U(2,12) m
    nb = max(1, int(round(1000 * area)))
    sig = rng.uniform(0.3, 1.5, nb)
    cnt = _alloc(rng, n_blob, np.ones(nb))
    bc = np.empty((nb, 3))
    bc[:, :2] = rng.uniform(0.0, extent, size=(nb, 2))
    bc[:, 2] = rng.uniform(2.0, 12.0, nb)
    bid = np.repeat(np.arange(nb), cnt)
    b = bc[bid] + rng.normal(0.0, 1.0, size=(n_blob, 3)) * sig[bid, None]
    parts.append(b)

    pts = np.concatenate(parts).astype(np.float32)
    pts = pts[rng.permutation(len(pts))]
    return np.ascontiguousarray(pts)


def precision_scene(n: int, seed: int = 3, offset: float = 1e5) -> np.ndarray:
    """Degenerate-geometry scene for the optimal-k / precision workload."""
    rng = np.random.default_rng(seed)
    n_axis, n_tilt, n_line, n_blob = _split(n, [0.25, 0.25, 0.2, 0.3])
    per = 2500
    parts = []

    def centres(m):
        c = np.empty((m, 3))
        c[:, :2] = rng.uniform(0.0, 200.0, size=(m, 2))
        c[:, 2] = rng.uniform(0.0, 20.0, size=m)
        return c

    # axis-aligned planar patches: 5 m x 5 m, z exactly constant per patch
    m = max(1, n_axis // per)
    cen = centres(m)
    pid = rng.integers(0, m, n_axis)
    p = np.empty((n_axis, 3))
    p[:, :2] = cen[pid, :2] + rng.uniform(-2.5, 2.5, size=(n_axis, 2))
    p[:, 2] = np.round(cen[pid, 2], 1)
    parts.append(p)

    # tilted planar patches: p = c + u a + v b, (a, b) orthonormal
    m = max(1, n_tilt // per)
    cen = centres(m)
    nrm = rng.normal(size=(m, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    a = np.cross(nrm, np.array([0.0, 0.0, 1.0]))
    a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-12)
    b = np.cross(nrm, a)
    pid = rng.integers(0, m, n_tilt)
    uv = rng.uniform(-2.5, 2.5, size=(n_tilt, 2))
    parts.append(cen[pid] + uv[:, :1] * a[pid] + uv[:, 1:] * b[pid])

    # collinear segments: 10 m long
    m = max(1, n_line // 1000)
    cen = centres(m)
    dirs = rng.normal(size=(m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pid = rng.integers(0, m, n_line)
    t = rng.uniform(-5.0, 5.0, n_line)
    parts.append(cen[pid] + t[:, None] * dirs[pid])

    # isotropic blobs: sigma U(0.2, 1) m
    m = max(1, n_blob // 1000)
    cen = centres(m)
    sig = rng.uniform(0.2, 1.0, m)
    pid = rng.integers(0, m, n_blob)
    parts.append(cen[pid] + rng.normal(size=(n_blob, 3)) * sig[pid, None])

    pts = np.concatenate(parts)
    pts[:, :2] += offset
    pts = pts.astype(np.float32)
    pts = pts[rng.permutation(len(pts))]
    return np.ascontiguousarray(pts)


def request_sizes(n_requests: int, seed: int = 4, lo: int = 5_000, hi: int = 2_000_000) -> np.ndarray:
    """Log-uniform request sizes in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_requests)).astype(np.int64)


def request_stream(n_requests: int, seed: int = 4, base_points: int = 2_000_000, workers: int = 16) -> list:
    """The serving workload's requests: sizes log-uniform in [5k, 2M]
    (``request_sizes``), each a square window of one LiDAR-like base scene of
    ``base_points`` points at the 10 M-point / 200 m density, its points in a
    seeded random order (a window of a scan is not in spatial order).
    Windows are cut from the base sorted into 128 x 128 ground tiles, with
    ``workers`` threads."""
    from concurrent.futures import ThreadPoolExecutor

    extent = 200.0 * float(np.sqrt(base_points / 10_000_000))
    base = lidar_scene(base_points, seed=seed, extent=extent)
    sizes = request_sizes(n_requests, seed=seed)
    G = 128
    tx = np.minimum((base[:, 0] / extent * G).astype(np.int64), G - 1).clip(0)
    ty = np.minimum((base[:, 1] / extent * G).astype(np.int64), G - 1).clip(0)
    tile = ty * G + tx
    order = np.argsort(tile, kind="stable")
    starts = np.searchsorted(tile[order], np.arange(G * G + 1))
    tiled = base[order]
    rng = np.random.default_rng(seed)
    jobs = []
    for i, sz in enumerate(sizes):
        side = extent * min(1.0, float(np.sqrt(sz / base_points)))
        cx, cy = rng.uniform(side / 2, extent - side / 2, size=2) if side < extent else (extent / 2, extent / 2)
        jobs.append((int(sz), float(cx), float(cy), side, seed * 100_003 + i))

    def cut(job):
        sz, cx, cy, side, s = job
        x0, x1 = int(max(0, (cx - side / 2) / extent * G)), int(min(G - 1, (cx + side / 2) / extent * G))
        y0, y1 = int(max(0, (cy - side / 2) / extent * G)), int(min(G - 1, (cy + side / 2) / extent * G))
        p = np.concatenate([tiled[starts[y * G + x0]:starts[y * G + x1 + 1]] for y in range(y0, y1 + 1)])
        p = p[(np.abs(p[:, 0] - cx) <= side / 2) & (np.abs(p[:, 1] - cy) <= side / 2)][:sz]
        return np.ascontiguousarray(p[np.random.default_rng(s).permutation(len(p))])

    with ThreadPoolExecutor(max(1, workers)) as ex:
        return list(ex.map(cut, jobs))


def request_cloud(size: int, seed: int) -> np.ndarray:
    """One serving request: a LiDAR-like scene of `size` points at the density
    of the 10 M-point / 200 m scene."""
    extent = 200.0 * float(np.sqrt(size / 10_000_000))
    return lidar_scene(int(size), seed=seed, extent=max(extent, 5.0))


def knn_csr(idx: np.ndarray):
    """CSR of a dense kNN result (README.md:136-141)."""
    n, k = idx.shape
    nn_ptr = (np.arange(n + 1, dtype=np.int64) * k).astype(np.uint32)
    return np.ascontiguousarray(idx.reshape(-1).astype(np.uint32)), nn_ptr


def radius_csr(idx: np.ndarray):
    """CSR of a -1-padded radius result (README.md:158-159)."""
    nn_ptr = np.r_[0, (idx >= 0).sum(axis=1).cumsum()].astype(np.uint32)
    nn = idx[idx >= 0].astype(np.uint32)
    return np.ascontiguousarray(nn), np.ascontiguousarray(nn_ptr)
