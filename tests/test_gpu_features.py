"""GPU parity of compute_features / _multiscale / _optimal against the oracle.

Reference semantics: include/pca.hpp:71-200, include/pgeof.hpp:75-310.
Tolerances and the excluded (unpinned) entries are documented in tests/parity.py.
"""
import numpy as np
import pytest

import point_geometric_features_amd as pgeof
from oracle import oracle as O
from point_geometric_features_amd import scenes
from tests.parity import assert_features_close, optimal_disagreement_is_tie

pytestmark = pytest.mark.gpu


def _knn_csr(xyz, k):
    idx, _ = O.knn(xyz, xyz, k)
    return scenes.knn_csr(idx)


def test_features_config1_uniform_10k_k20():
    xyz = scenes.uniform_cube(10_000, seed=0)
    idx, _ = pgeof.knn_search(xyz, xyz, 20)
    nn, ptr = scenes.knn_csr(idx)
    f = pgeof.compute_features(xyz, nn, ptr)
    assert f.shape == (10_000, 11) and f.dtype == np.float32
    ref = O.compute_features(xyz, nn, ptr)
    val = O.eigvals_csr_np(xyz, nn, ptr)
    max_err, _ = assert_features_close(f, ref, val, "config1")
    assert max_err.max() < 1e-4
    assert (f[:, 6] >= 0).all()


def test_features_lidar_k50():
    xyz = scenes.lidar_scene(120_000, seed=21, extent=30.0)
    nn, ptr = _knn_csr(xyz, 50)
    f = pgeof.compute_features(xyz, nn, ptr)
    ref = O.compute_features(xyz, nn, ptr)
    val = O.eigvals_csr_np(xyz, nn, ptr)
    assert_features_close(f, ref, val, "lidar k50")


def test_features_precision_scene_offset():
    xyz = scenes.precision_scene(50_000, seed=22)
    nn, ptr = _knn_csr(xyz, 30)
    f = pgeof.compute_features(xyz, nn, ptr)
    ref = O.compute_features(xyz, nn, ptr)
    val = O.eigvals_csr_np(xyz, nn, ptr)
    assert_features_close(f, ref, val, "precision scene")


def test_features_known_answers():
    # closed forms with Eigen's ascending eigen-order (include/pca.hpp:79-96)
    sigma = 0.5
    line_x = np.zeros((4, 3), np.float32)
    line_x[:, 0] = [-1.5, -0.5, 0.5, 1.5]  # variance 1.25 along x, exactly collinear
    plane = np.array([[a, b, 2.0] for a in (-1.0, 1.0) for b in (-2.0, 2.0)], np.float32)  # var 1, 4, z const
    single = np.array([[3.0, 4.0, 5.0]], np.float32)
    dup = np.array([[1.0, 1.0, 1.0]] * 5, np.float32)
    xyz = np.concatenate([line_x, plane, single, dup])
    rows = [np.arange(0, 4), np.arange(4, 8), np.array([8]), np.arange(9, 14)]
    nn = np.concatenate(rows).astype(np.uint32)
    ptr = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.uint32)
    f = pgeof.compute_features(xyz, nn, ptr)
    eps = 1e-3
    s = np.sqrt(1.25)
    expect_line = [0.0, -s / eps, s / eps, 0.0, 1.0, 0.0, 0.0, 0.0, np.sqrt(1e-6), 1e-3, s / (s + eps)]
    a, b = 1.0, 2.0
    expect_plane = [-a / eps, (a - b) / eps, b / eps, 0.0, 0.0, 1.0, 0.0, 0.0, np.sqrt(1e-6), 1e-3,
                    b / (a + b + eps)]
    zero_row = [0, 0, 0, 0, 0, 0, 1, 0, 1e-3, 1e-3, 0]
    np.testing.assert_allclose(f[0], expect_line, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(f[1, [0, 1, 2, 3, 6, 7, 8, 9, 10]],
                               np.array(expect_plane)[[0, 1, 2, 3, 6, 7, 8, 9, 10]], rtol=1e-5, atol=1e-5)
    assert abs(abs(f[1, 5]) - 1.0) < 1e-6 and abs(f[1, 4]) < 1e-6  # in-plane major axis (y), sign unpinned
    np.testing.assert_allclose(f[2], zero_row, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(f[3], zero_row, rtol=1e-5, atol=1e-6)


def test_features_k_min_and_empty_rows():
    xyz = scenes.uniform_cube(2000, seed=23)
    idx, _ = O.radius(xyz, xyz, 0.08, 30)[:2]
    nn, ptr = scenes.radius_csr(idx)
    for k_min in (1, 5, 12):
        f = pgeof.compute_features(xyz, nn, ptr, k_min=k_min)
        ref = O.compute_features(xyz, nn, ptr, k_min=k_min)
        val = O.eigvals_csr_np(xyz, nn, ptr)
        assert_features_close(f, ref, val, f"radius csr k_min={k_min}")
        counts = np.diff(ptr.astype(np.int64))
        assert (f[counts < k_min] == 0).all()
    # rows past the end of xyz are not needed: n points = len(nn_ptr) - 1
    f = pgeof.compute_features(xyz, nn, ptr[:11])
    assert f.shape == (10, 11)
    f = pgeof.compute_features(xyz, nn, ptr[:1])
    assert f.shape == (0, 11)


def test_features_invalid_csr_raises():
    xyz = scenes.uniform_cube(100, seed=24)
    nn = np.array([0, 1, 2, 100], np.uint32)  # 100 is out of range
    ptr = np.array([0, 2, 4], np.uint32)
    with pytest.raises(ValueError):
        pgeof.compute_features(xyz, nn, ptr)
    with pytest.raises(ValueError):
        pgeof.compute_features(xyz, nn[:3], np.array([0, 2, 1], np.uint32))
    with pytest.raises(ValueError):
        pgeof.compute_features(xyz, nn[:3], np.array([0, 2, 9], np.uint32))
    with pytest.raises(ValueError):
        pgeof.compute_features(xyz, nn[:3], np.array([0, 3], np.uint32), k_min=0)
    with pytest.raises(TypeError):
        pgeof.compute_features(xyz, nn.astype(np.int64), ptr)
    # an index past the end in the middle of a row, for every CSR driver
    with pytest.raises(ValueError):
        pgeof.compute_features_multiscale(xyz, nn, ptr, [1, 2])
    with pytest.raises(ValueError):
        pgeof.compute_features_optimal(xyz, nn, ptr)
    # rows that point into an empty cloud
    with pytest.raises(ValueError):
        pgeof.compute_features(np.zeros((0, 3), np.float32), nn[:2], np.array([0, 2], np.uint32))


def test_multiscale_reference_test():
    # tests/test_pgeof.py:30-46 (reference): self-consistency, rtol 1e-1 atol 1e-5
    rng = np.random.default_rng(31)
    xyz = rng.uniform(0.0, 200.0, size=(10_000, 3)).astype(np.float32)
    nn, ptr = _knn_csr(xyz, 50)
    scales = np.array([50, 20])
    multi = pgeof.compute_features_multiscale(xyz, nn, ptr, np.flip(scales), False)
    simple = pgeof.compute_features(xyz, nn, ptr, 50, False)
    multi_simple = pgeof.compute_features_multiscale(xyz, nn, ptr, [20], False)
    np.testing.assert_allclose(multi[:, 0], multi_simple[:, 0], 1e-1, 1e-5)
    np.testing.assert_allclose(multi[:, 1], simple, 1e-1, 1e-5)
    # stronger: one walk of the row gives every prefix exactly
    np.testing.assert_array_equal(multi[:, 1], simple)
    nn20 = nn.reshape(-1, 50)[:, :20].ravel().copy()
    ptr20 = (np.arange(len(xyz) + 1) * 20).astype(np.uint32)
    np.testing.assert_array_equal(multi[:, 0], pgeof.compute_features(xyz, nn20, ptr20))


def test_multiscale_vs_oracle_ragged():
    xyz = scenes.lidar_scene(40_000, seed=32, extent=20.0)
    idx, _ = pgeof.radius_search(xyz, xyz, 0.5, 64)
    nn, ptr = scenes.radius_csr(idx)
    sc = [10, 20, 20, 50, 100]
    m = pgeof.compute_features_multiscale(xyz, nn, ptr, sc)
    assert m.shape == (len(xyz), 5, 11)
    ref = O.compute_features_multiscale(xyz, nn, ptr, sc)
    for s, k in enumerate(sc):
        val = O.eigvals_csr_np(xyz, nn, ptr, prefix=k)
        assert_features_close(m[:, s], ref[:, s], val, f"scale {k}")
    counts = np.diff(ptr.astype(np.int64))
    assert (m[counts < 100, 4] == 0).all()
    np.testing.assert_array_equal(m[:, 1], m[:, 2])
    with pytest.raises(ValueError):
        pgeof.compute_features_multiscale(xyz, nn, ptr, [20, 10])
    with pytest.raises(ValueError):
        pgeof.compute_features_multiscale(xyz, nn, ptr, [0, 10])
    assert pgeof.compute_features_multiscale(xyz, nn, ptr, []).shape == (len(xyz), 0, 11)


@pytest.mark.parametrize("k_min,k_step,k_min_search", [(10, 1, 10), (1, 3, 5), (5, 7, 1)])
def test_optimal_vs_oracle(k_min, k_step, k_min_search):
    xyz = scenes.precision_scene(20_000, seed=41)
    nn, ptr = _knn_csr(xyz, 40)
    g = pgeof.compute_features_optimal(xyz, nn, ptr, k_min, k_step, k_min_search)
    assert g.shape == (len(xyz), 12)
    ref = O.compute_features_optimal(xyz, nn, ptr, k_min, k_step, k_min_search, eigentropy_fp32=True)
    agree = g[:, 11] == ref[:, 11]
    # k_opt is an argmin over FP32 eigentropies; near-ties may resolve differently
    assert agree.mean() > 0.995, agree.mean()
    assert optimal_disagreement_is_tie(xyz, nn, 40, np.nonzero(~agree)[0], g[~agree, 11], ref[~agree, 11])
    kopt = ref[agree, 11].astype(np.int64)
    rows = np.nonzero(agree)[0]
    val = np.zeros((len(rows), 3))
    for kk in np.unique(kopt):
        sel = kopt == kk
        r = rows[sel]
        sub_ptr = np.r_[0, np.full(len(r), kk).cumsum()].astype(np.uint32)
        sub_nn = nn.reshape(-1, 40)[r, :kk].ravel().copy()
        val[sel] = O.eigvals_csr_np(xyz, sub_nn, sub_ptr)
    assert_features_close(g[agree], ref[agree], val, "optimal")
    ks = g[:, 11]
    k0 = max(k_min, k_min_search)
    assert ((ks == k0) | (ks % k_step == 0) | (ks == 40)).all()


def test_optimal_errors():
    xyz = scenes.uniform_cube(100, seed=42)
    nn, ptr = _knn_csr(xyz, 5)
    with pytest.raises(ValueError):
        pgeof.compute_features_optimal(xyz, nn, ptr, 0, 1, 0)
    with pytest.raises(ValueError):
        pgeof.compute_features_optimal(xyz, nn, ptr, 1, 0, 1)
    # k_min = 0 alone is accepted (the reference checks k_min < 1 && k_min_search < 1)
    pgeof.compute_features_optimal(xyz, nn, ptr, 0, 1, 1)


def test_csr_spatial_row_order_changes_nothing():
    """Plain and optimal-k features on 65536 rows or more visit the rows in
    spatial order; the results are the same bits as in index order, and bad
    rows still raise."""
    from point_geometric_features_amd import capi

    xyz = scenes.lidar_scene(80_000, seed=33, extent=25.0)
    idx, _ = pgeof.knn_search(xyz, xyz, 24)
    nn, ptr = scenes.knn_csr(idx)
    old = capi.get_tuning("csr_order")
    try:
        res = {}
        for order in (1, 0):
            capi.set_tuning("csr_order", order)
            res[order] = (pgeof.compute_features(xyz, nn, ptr),
                          pgeof.compute_features_multiscale(xyz, nn, ptr, [8, 16, 24]),
                          pgeof.compute_features_optimal(xyz, nn, ptr, 8, 2, 8))
        capi.set_tuning("csr_order", 1)
        bad = nn.copy()
        bad[len(bad) // 2] = len(xyz)
        with pytest.raises(ValueError):
            pgeof.compute_features_optimal(xyz, bad, ptr, 8, 2, 8)
        with pytest.raises(ValueError):
            pgeof.compute_features(xyz, bad, ptr)
    finally:
        capi.set_tuning("csr_order", old)
    for a, b in zip(res[1], res[0]):
        np.testing.assert_array_equal(a, b)
    ref = O.compute_features(xyz, nn, ptr)
    assert_features_close(res[1][0], ref, O.eigvals_csr_np(xyz, nn, ptr), "spatial row order")


def test_row_order_one_bucket():
    """Every row in one bucket of the spatial row order (all points at one
    place, plus a far cluster so that the box is not empty): the last pass
    places 300,000 rows of one bucket; the results equal index order's."""
    from point_geometric_features_amd import capi

    rng = np.random.default_rng(35)
    xyz = np.zeros((300_000, 3), np.float32)
    xyz[:1000] = rng.random((1000, 3), dtype=np.float32) + 100.0
    nn = ((np.arange(len(xyz))[:, None] + np.arange(8)[None, :]) % len(xyz)).astype(np.uint32).ravel()
    ptr = (np.arange(len(xyz) + 1) * 8).astype(np.uint32)
    old = capi.get_tuning("csr_order")
    try:
        res = {}
        for order in (1, 0):
            capi.set_tuning("csr_order", order)
            res[order] = pgeof.compute_features(xyz, nn, ptr)
    finally:
        capi.set_tuning("csr_order", old)
    np.testing.assert_array_equal(res[1], res[0])


def test_row_order_more_rows_than_points():
    """A CSR of more rows than points (the kNN rows of 100,000 queries into
    40,000 data points, include/pgeof.hpp:75-117 takes the row count from
    nn_ptr): rows past the end of xyz are ordered by their first neighbour.
    The features equal index order's bit for bit (plain, multiscale and
    optimal k) and the oracle's within the parity bar."""
    from point_geometric_features_amd import capi

    data = scenes.lidar_scene(40_000, seed=36, extent=20.0)
    q = scenes.lidar_scene(100_000, seed=37, extent=20.0)
    idx, _ = pgeof.knn_search(data, q, 16)
    nn, ptr = scenes.knn_csr(idx)
    old = capi.get_tuning("csr_order")
    try:
        res = {}
        for order in (1, 0):
            capi.set_tuning("csr_order", order)
            res[order] = (pgeof.compute_features(data, nn, ptr),
                          pgeof.compute_features_multiscale(data, nn, ptr, [8, 16]),
                          pgeof.compute_features_optimal(data, nn, ptr, 8, 2, 8))
    finally:
        capi.set_tuning("csr_order", old)
    for a, b in zip(res[1], res[0]):
        np.testing.assert_array_equal(a, b)
    assert res[1][0].shape == (100_000, 11)
    ref = O.compute_features(data, nn, ptr)
    val = O.eigvals_csr_np(data, nn, ptr)
    assert_features_close(res[1][0], ref, val, "more rows than points")
