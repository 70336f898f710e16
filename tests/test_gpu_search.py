"""GPU parity of knn_search / radius_search against the CPU oracle.

Bar: bit-exact indices AND squared distances vs the float32 brute-force /
KD-tree oracle (oracle/pgeof_oracle.c, nanoflann's L2_Simple expression, ties
by index); the reference's own tests (tests/test_pgeof.py:8-27 upstream)
compare indices with scipy's KD-tree, reproduced here with fixed seeds.
"""
import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree

import point_geometric_features_amd as pgeof
from oracle import oracle as O
from point_geometric_features_amd import scenes

pytestmark = pytest.mark.gpu


def _eq_knn(data, query, k, kdtree=True):
    gi, gd = pgeof.knn_search(data, query, k)
    oi, od = O.knn(data, query, k, kdtree=kdtree)
    assert gi.dtype == np.uint32 and gd.dtype == np.float32
    assert gi.shape == (len(query), k)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    return gi, gd


def test_knn_reference_test_vs_scipy():
    # tests/test_pgeof.py:8-15 (reference), seeded
    rng = np.random.default_rng(1234)
    xyz = rng.uniform(0.0, 200.0, size=(1000, 3)).astype(np.float32)
    _, k_legacy = cKDTree(xyz).query(xyz, k=10)
    k_new, _ = pgeof.knn_search(xyz, xyz, 10)
    np.testing.assert_equal(k_legacy, k_new)


def test_knn_uniform_10k_k20_bitexact():
    xyz = scenes.uniform_cube(10_000, seed=0)
    gi, gd = _eq_knn(xyz, xyz, 20, kdtree=False)
    assert (gi[:, 0] == np.arange(len(xyz))).all()
    assert (np.diff(gd, axis=1) >= 0).all()


@pytest.mark.parametrize("k", [1, 2, 4, 5, 6, 7, 20, 21, 22, 26, 33, 64, 100])
def test_knn_various_k(k):
    xyz = scenes.uniform_cube(3000, seed=k)
    _eq_knn(xyz, xyz, k, kdtree=False)


@pytest.mark.parametrize("heap_tail", [1, 0])
def test_knn_large_k_global_heap(heap_tail):
    """k = 1200: with heap_tail 1 the 4-ary heap keeps nodes 21..1199 in global
    memory; with heap_tail 0 the heap cannot fit the LDS budget and the search
    takes the binary heap held wholly in global memory (heapify after a partial
    fill, sort in finish)."""
    from point_geometric_features_amd import capi

    xyz = scenes.uniform_cube(1500, seed=5)
    old = capi.get_tuning("heap_tail")
    try:
        capi.set_tuning("heap_tail", heap_tail)
        _eq_knn(xyz, xyz, 1200, kdtree=False)
    finally:
        capi.set_tuning("heap_tail", old)


def test_knn_k_equals_n():
    xyz = scenes.uniform_cube(300, seed=6)
    _eq_knn(xyz, xyz, 300, kdtree=False)


def test_knn_queries_differ_and_outside():
    rng = np.random.default_rng(7)
    data = rng.uniform(0, 10, size=(5000, 3)).astype(np.float32)
    q = rng.uniform(-20, 30, size=(2000, 3)).astype(np.float32)
    _eq_knn(data, q, 16, kdtree=False)


def test_knn_duplicates_ties_by_index():
    rng = np.random.default_rng(8)
    base = rng.integers(0, 6, size=(400, 3)).astype(np.float32)  # lattice: many exact ties
    xyz = np.concatenate([base, base[::-1]])
    _eq_knn(xyz, xyz, 24, kdtree=False)


def test_knn_lidar_200k_vs_kdtree():
    xyz = scenes.lidar_scene(200_000, seed=11, extent=40.0)
    _eq_knn(xyz, xyz, 20, kdtree=True)


def test_knn_precision_scene_offset():
    xyz = scenes.precision_scene(60_000, seed=12)
    _eq_knn(xyz, xyz, 50, kdtree=True)


def test_knn_empty_query_and_errors():
    xyz = scenes.uniform_cube(100, seed=9)
    gi, gd = pgeof.knn_search(xyz, np.zeros((0, 3), np.float32), 5)
    assert gi.shape == (0, 5)
    with pytest.raises(ValueError):
        pgeof.knn_search(xyz, xyz, 101)
    with pytest.raises(TypeError):
        pgeof.knn_search(xyz.astype(np.float64), xyz, 5)


def test_radius_reference_test_vs_scipy():
    # tests/test_pgeof.py:18-27 (reference), seeded
    rng = np.random.default_rng(4321)
    xyz = rng.random(size=(1000, 3), dtype=np.float32)
    _, k_legacy = cKDTree(xyz).query(xyz, k=10, distance_upper_bound=0.2)
    k_legacy[k_legacy == xyz.shape[0]] = -1
    k_new, _ = pgeof.radius_search(xyz, xyz, 0.2, 10)
    np.testing.assert_equal(k_legacy, k_new)


@pytest.mark.parametrize("r,cap", [(0.05, 20), (0.1, 64), (0.02, 5), (0.3, 150)])
def test_radius_bitexact(r, cap):
    xyz = scenes.uniform_cube(8000, seed=int(r * 1000) + cap)
    gi, gd = pgeof.radius_search(xyz, xyz, r, cap)
    oi, od, cnt = O.radius(xyz, xyz, r, cap, kdtree=False)
    assert gi.dtype == np.int32
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    # nanoflann RKNNResultSet: r^2 in the last slot of rows not filled
    notfull = cnt < cap
    if notfull.any():
        assert (gd[notfull, -1] == np.float32(r) * np.float32(r)).all()


@pytest.mark.parametrize("cap", [9, 12, 30])
def test_radius_rows_of_every_length(cap):
    """Rows that hold 0, 1, ..., 8 hits (sorted in registers at the end of
    the search) and rows of 9 or more (heap-sorted), in one search, bit-exact
    against the oracle; queries partly off the data so that some rows are
    empty."""
    rng = np.random.default_rng(77 + cap)
    data = rng.uniform(0, 4, size=(20_000, 3)).astype(np.float32)
    q = rng.uniform(-0.3, 4.3, size=(20_000, 3)).astype(np.float32)
    r = 0.19  # about 9 points per ball inside the data
    gi, gd = pgeof.radius_search(data, q, r, cap)
    oi, od, cnt = O.radius(data, q, r, cap, kdtree=True)
    assert all((cnt == c).any() for c in range(0, 10)), np.bincount(cnt)[:12]
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


def test_radius_queries_differ():
    rng = np.random.default_rng(13)
    data = rng.uniform(0, 5, size=(6000, 3)).astype(np.float32)
    q = rng.uniform(-1, 6, size=(3000, 3)).astype(np.float32)
    gi, gd = pgeof.radius_search(data, q, 0.4, 32)
    oi, od, _ = O.radius(data, q, 0.4, 32, kdtree=False)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


def test_radius_lidar_vs_kdtree():
    xyz = scenes.lidar_scene(150_000, seed=14, extent=35.0)
    gi, gd = pgeof.radius_search(xyz, xyz, 0.5, 64)
    oi, od, _ = O.radius(xyz, xyz, 0.5, 64, kdtree=True)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


def test_radius_errors():
    xyz = scenes.uniform_cube(50, seed=15)
    with pytest.raises(ValueError):
        pgeof.radius_search(xyz, xyz, 0.1, 51)


def test_knn_sparse_outliers_use_coarse_levels():
    # a dense cluster plus isolated far points: the isolated queries need the
    # coarse grid levels, the dense ones the fine level
    rng = np.random.default_rng(16)
    dense = rng.normal(0, 0.05, size=(20_000, 3))
    sparse = rng.uniform(-50, 50, size=(300, 3))
    xyz = np.concatenate([dense, sparse]).astype(np.float32)
    xyz = xyz[rng.permutation(len(xyz))]
    _eq_knn(xyz, xyz, 20, kdtree=False)


@pytest.mark.parametrize("levels", [1, 2, 3, 4])
def test_knn_level_count_does_not_change_results(levels):
    from point_geometric_features_amd import capi

    xyz = scenes.lidar_scene(60_000, seed=17, extent=25.0)
    old = capi.get_tuning("grid_levels")
    try:
        capi.set_tuning("grid_levels", levels)
        gi, gd = pgeof.knn_search(xyz, xyz, 20)
        x = torch.from_numpy(xyz).cuda()
        nn_d = torch.empty(len(xyz) * 20, dtype=torch.int32, device="cuda")
        ptr_d = torch.empty(len(xyz) + 1, dtype=torch.int32, device="cuda")
        f_d = torch.empty((len(xyz), 11), dtype=torch.float32, device="cuda")
        capi.knn_features_device(x, 20, nn_d, ptr_d, f_d)
        torch.cuda.synchronize()
    finally:
        capi.set_tuning("grid_levels", old)
    oi, od = O.knn(xyz, xyz, 20, kdtree=True)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(nn_d.cpu().numpy().view(np.uint32), gi.reshape(-1))
    nn, ptr = scenes.knn_csr(gi)
    np.testing.assert_array_equal(f_d.cpu().numpy(), pgeof.compute_features(xyz, nn, ptr, 20))


@pytest.mark.parametrize("near", [0, 1])
def test_knn_scan_order_does_not_change_results(near):
    from point_geometric_features_amd import capi

    xyz = scenes.lidar_scene(60_000, seed=23, extent=25.0)
    keys = ("near_first",)
    old = [capi.get_tuning(k) for k in keys]
    try:
        capi.set_tuning("near_first", near)
        gi, gd = pgeof.knn_search(xyz, xyz, 20)
        ri, rd = pgeof.radius_search(xyz, xyz, 0.5, 30)
        x = torch.from_numpy(xyz).cuda()
        nn_d = torch.empty(len(xyz) * 20, dtype=torch.int32, device="cuda")
        ptr_d = torch.empty(len(xyz) + 1, dtype=torch.int32, device="cuda")
        f_d = torch.empty((len(xyz), 11), dtype=torch.float32, device="cuda")
        capi.knn_features_device(x, 20, nn_d, ptr_d, f_d)
        torch.cuda.synchronize()
        fused = f_d.cpu().numpy()
    finally:
        for k, v in zip(keys, old):
            capi.set_tuning(k, v)
    oi, od = O.knn(xyz, xyz, 20, kdtree=True)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    oi, od, _ = O.radius(xyz, xyz, 0.5, 30, kdtree=True)
    np.testing.assert_array_equal(ri, oi)
    np.testing.assert_array_equal(rd, od)
    nn, ptr = scenes.knn_csr(gi)
    np.testing.assert_array_equal(fused, pgeof.compute_features(xyz, nn, ptr, 20))


_KNOB_SETS = [
    {"budget_cells": 1},
    {"budget_cells": 3, "budget_shells": 3},
    {"grid_levels": 4, "budget_levels": 4},
    {"grid_levels": 3, "grid_level_ratio": 4},
    {"grid_cells_per_point": 1, "grid_levels": 1},
    {"grid_cells_per_point": 32, "knn_occupancy": 0.05},
    {"knn_occupancy": 2.0},
    {"knn_cell_points_max": 2.0},
    {"lds_budget": 16384},
    {"lds_budget": 4096},  # heaps in global memory
    {"heap_tail": 0},  # radius cap 24: the whole heap in LDS instead of nodes 21..23 in global memory
    {"staging": 1},  # radius searches staged, the budgeted kNN pass at k >= 32
    {"staging": 2},
]


@pytest.mark.parametrize("knobs", _KNOB_SETS, ids=lambda d: ",".join(f"{k}={v}" for k, v in d.items()))
def test_tuning_knobs_change_speed_only(knobs):
    """Every knob setting gives the oracle's neighbours bit for bit, and the
    fused path the two-call path's features bit for bit."""
    from point_geometric_features_amd import capi

    xyz = scenes.lidar_scene(50_000, seed=29, extent=20.0)
    old = {k: capi.get_tuning(k) for k in knobs}
    try:
        for k, v in knobs.items():
            capi.set_tuning(k, v)
        gi, gd = pgeof.knn_search(xyz, xyz, 16)
        x = torch.from_numpy(xyz).cuda()
        nn_d = torch.empty(len(xyz) * 16, dtype=torch.int32, device="cuda")
        ptr_d = torch.empty(len(xyz) + 1, dtype=torch.int32, device="cuda")
        f_d = torch.empty((len(xyz), 11), dtype=torch.float32, device="cuda")
        ws = torch.empty(capi.knn_workspace_size(len(xyz), len(xyz), 16), dtype=torch.uint8, device="cuda")
        capi.knn_features_device(x, 16, nn_d, ptr_d, f_d, workspace=ws)
        torch.cuda.synchronize()
        ri, rd = pgeof.radius_search(xyz, xyz, 0.3, 24)
    finally:
        for k, v in old.items():
            capi.set_tuning(k, v)
    oi, od = O.knn(xyz, xyz, 16, kdtree=True)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(nn_d.cpu().numpy().view(np.uint32), gi.reshape(-1))
    nn, ptr = scenes.knn_csr(gi)
    np.testing.assert_array_equal(f_d.cpu().numpy(), pgeof.compute_features(xyz, nn, ptr))
    oi, od, _ = O.radius(xyz, xyz, 0.3, 24, kdtree=True)
    np.testing.assert_array_equal(ri, oi)
    np.testing.assert_array_equal(rd, od)


def test_host_search_after_outputs_of_odd_length():
    """The host entry points put the search workspace right after their
    outputs; with m * k odd its base is 4 mod 8, and every workspace region
    must still be aligned for the kernels' 64-bit atomics (regression: the
    occupancy probe's sum of squares faulted there)."""
    data = scenes.lidar_scene(80_000, seed=101, extent=25.0)
    q = data[::7].copy()  # 11,429 queries
    for k in (9, 7):
        assert (len(q) * k) % 2 == 1
        _eq_knn(data, q, k, kdtree=True)
    ri, rd = pgeof.radius_search(data, q, 0.4, 9)
    oi, od, _ = O.radius(data, q, 0.4, 9, kdtree=True)
    np.testing.assert_array_equal(ri, oi)
    np.testing.assert_array_equal(rd, od)


@pytest.mark.parametrize("staging", [0, 1, 2])
def test_knn_k50_row_staging_changes_speed_only(staging):
    """k = 50 (heap tails, the budgeted pass staged in LDS with staging 1 and
    2, per-lane row streams with 0): the oracle's neighbours bit for bit, and
    the fused features equal compute_features on that CSR."""
    from point_geometric_features_amd import capi

    xyz = scenes.lidar_scene(60_000, seed=31, extent=25.0)
    old = capi.get_tuning("staging")
    try:
        capi.set_tuning("staging", staging)
        gi, gd = pgeof.knn_search(xyz, xyz, 50)
        x = torch.from_numpy(xyz).cuda()
        nn_d = torch.empty(len(xyz) * 50, dtype=torch.int32, device="cuda")
        ptr_d = torch.empty(len(xyz) + 1, dtype=torch.int32, device="cuda")
        f_d = torch.empty((len(xyz), 11), dtype=torch.float32, device="cuda")
        capi.knn_features_device(x, 50, nn_d, ptr_d, f_d)
        torch.cuda.synchronize()
    finally:
        capi.set_tuning("staging", old)
    oi, od = O.knn(xyz, xyz, 50, kdtree=True)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(nn_d.cpu().numpy().view(np.uint32), gi.reshape(-1))
    nn, ptr = scenes.knn_csr(gi)
    np.testing.assert_array_equal(f_d.cpu().numpy(), pgeof.compute_features(xyz, nn, ptr, 50))


@pytest.mark.parametrize("k", [26, 50])
def test_fused_heap_tail_lattice_ties(k):
    """The fused path at k > 21 keys its budgeted pass by sorted position and
    sends every query with an exact distance tie to the index-keyed pass: on
    a lattice (ties everywhere) the rows must still be the oracle's (ties by
    index) and the features those of compute_features on that CSR."""
    from point_geometric_features_amd import capi

    g = np.stack(np.meshgrid(*[np.arange(16)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    rng = np.random.default_rng(37)
    xyz = np.ascontiguousarray(g[rng.permutation(len(g))] * np.float32(0.25))
    x = torch.from_numpy(xyz).cuda()
    nn_d = torch.empty(len(xyz) * k, dtype=torch.int32, device="cuda")
    ptr_d = torch.empty(len(xyz) + 1, dtype=torch.int32, device="cuda")
    f_d = torch.empty((len(xyz), 11), dtype=torch.float32, device="cuda")
    d_d = torch.empty(len(xyz) * k, dtype=torch.float32, device="cuda")
    old = capi.get_tuning("heap_tail")
    try:
        capi.set_tuning("heap_tail", 2)  # the wave-slot tails of large searches, at this size too
        capi.knn_features_device(x, k, nn_d, ptr_d, f_d, sqr_dist=d_d)
        torch.cuda.synchronize()
    finally:
        capi.set_tuning("heap_tail", old)
    oi, od = O.knn(xyz, xyz, k, kdtree=False)
    np.testing.assert_array_equal(nn_d.cpu().numpy().view(np.uint32).reshape(-1, k), oi)
    np.testing.assert_array_equal(d_d.cpu().numpy().reshape(-1, k), od)
    nn, ptr = scenes.knn_csr(oi)
    np.testing.assert_array_equal(f_d.cpu().numpy(), pgeof.compute_features(xyz, nn, ptr, k))
