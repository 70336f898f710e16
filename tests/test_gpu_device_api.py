"""GPU parity of the device-memory entry points and of the paths the
benchmark times, against the oracle and the host API.

* pgeof_knn_features_device (knn_search -> CSR -> compute_features fused,
  README.md:136-141): neighbours bit-exact vs the oracle's KD-tree, nn_ptr =
  arange * k, features bit-equal to the two-call path and within the bar of
  the FP64 oracle; k = 22, 50 and 100 use the heap's global tail, k = 50 is
  the 10 M-point benchmark's second configuration; the tails in hardware
  wave slots equal the tails per launch lane and the all-LDS heap bit for
  bit.
* pgeof_radius_csr_device (README.md:158-159): nn_ptr = [0, cumsum(count)],
  nn = idx[idx >= 0] of the oracle's radius search.
* the device forms of compute_features / _multiscale / _optimal with and
  without the error word, and of knn_search / radius_search, equal to the
  host API bit for bit.
* compute_features_optimal(10, 1, 10) on rows of 100 neighbours
  (include/pgeof.hpp:243-310), the precision workload's setting.
"""
import numpy as np
import pytest
import torch

import point_geometric_features_amd as pgeof
from oracle import oracle as O
from point_geometric_features_amd import capi, scenes
from tests.parity import assert_features_close, optimal_disagreement_is_tie

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _fused(xyz, k, workspace=True):
    n = len(xyz)
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.full((n * k,), -7, dtype=torch.int32, device=DEV)
    ptr = torch.full((n + 1,), -7, dtype=torch.int32, device=DEV)
    d2 = torch.empty(n * k, dtype=torch.float32, device=DEV)
    f = torch.empty((n, 11), dtype=torch.float32, device=DEV)
    ws = torch.empty(capi.knn_workspace_size(n, n, k), dtype=torch.uint8, device=DEV) if workspace else None
    capi.knn_features_device(x, k, nn, ptr, f, sqr_dist=d2, workspace=ws)
    torch.cuda.synchronize()
    return (nn.cpu().numpy().view(np.uint32), ptr.cpu().numpy().view(np.uint32), d2.cpu().numpy(),
            f.cpu().numpy())


@pytest.fixture(scope="module")
def lidar250k():
    return scenes.lidar_scene(250_000, seed=71, extent=45.0)


@pytest.mark.parametrize("k", [22, 50, 100])
def test_knn_features_device_lidar(lidar250k, k):
    xyz = lidar250k
    n = len(xyz)
    nn, ptr, d2, f = _fused(xyz, k)
    oi, od = O.knn(xyz, xyz, k, kdtree=True)
    np.testing.assert_array_equal(nn, oi.ravel())
    np.testing.assert_array_equal(d2, od.ravel())
    np.testing.assert_array_equal(ptr, np.arange(n + 1, dtype=np.uint64) * k)
    # the two-call path, bit for bit
    gi, gd = pgeof.knn_search(xyz, xyz, k)
    np.testing.assert_array_equal(gi.ravel(), nn)
    hnn, hptr = scenes.knn_csr(gi)
    np.testing.assert_array_equal(f, pgeof.compute_features(xyz, hnn, hptr))
    # and the FP64 oracle
    ref = O.compute_features(xyz, hnn, hptr)
    val = O.eigvals_csr_np(xyz, hnn, hptr)
    assert_features_close(f, ref, val, f"fused lidar k={k}")


def test_knn_features_device_heap_tail_on_off(lidar250k):
    """k = 50: nodes 21..49 of every heap in global memory, one slot per
    hardware wave slot (heap_tail 2, the layout of searches of millions of
    queries) or per lane of the launch (heap_tail 1 at this size), or the
    whole heap in LDS (0); the same bits every way."""
    n, k = len(lidar250k), 50
    old = capi.get_tuning("heap_tail")
    try:
        capi.set_tuning("heap_tail", 1)
        a = _fused(lidar250k, k)
        per_lane_ws = capi.knn_workspace_size(n, n, k)
        capi.set_tuning("heap_tail", 2)
        c = _fused(lidar250k, k)
        per_slot_ws = capi.knn_workspace_size(n, n, k)
        capi.set_tuning("heap_tail", 0)
        b = _fused(lidar250k, k, workspace=False)
    finally:
        capi.set_tuning("heap_tail", old)
    # the wave-slot tails: 65536 slots of 64 lanes, 16 pairs of 16 bytes
    assert per_slot_ws - per_lane_ws >= 65536 * 64 * 16 * 16 - (n + 256) * 16 * 16
    for x, y, z in zip(a, b, c):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x, z)


def test_knn_search_wave_slot_heap_tail_vs_oracle():
    """kNN at k = 64 and 130 with the heap tails in wave slots, uneven
    density (a dense cluster in a sparse cube): the budgeted and unbudgeted
    passes both sift into the tail; bit-exact against the oracle."""
    rng = np.random.default_rng(17)
    xyz = np.concatenate([rng.random((30_000, 3)) * 40.0,
                          rng.normal(20.0, 0.3, (30_000, 3))]).astype(np.float32)
    old = capi.get_tuning("heap_tail")
    try:
        capi.set_tuning("heap_tail", 2)
        res = {k: pgeof.knn_search(xyz, xyz, k) for k in (64, 130)}
    finally:
        capi.set_tuning("heap_tail", old)
    for k, (gi, gd) in res.items():
        oi, od = O.knn(xyz, xyz, k, kdtree=True)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)


def test_knn_large_k_binary_global_heap():
    """k = 1200 with the heap tail off: the binary heap in global memory
    (heapify after a partial fill, heap sort at the end)."""
    xyz = scenes.uniform_cube(1500, seed=5)
    old = capi.get_tuning("heap_tail")
    try:
        capi.set_tuning("heap_tail", 0)
        gi, gd = pgeof.knn_search(xyz, xyz, 1200)
    finally:
        capi.set_tuning("heap_tail", old)
    oi, od = O.knn(xyz, xyz, 1200, kdtree=False)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


def _radius_csr_device(xyz, r, cap, workspace=True):
    n = len(xyz)
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.full((n * cap,), -7, dtype=torch.int32, device=DEV)
    ptr = torch.full((n + 1,), -7, dtype=torch.int32, device=DEV)
    ws = torch.empty(capi.radius_csr_workspace_size(n, cap), dtype=torch.uint8, device=DEV) if workspace else None
    nnz = capi.radius_csr_device(x, r, cap, nn, ptr, workspace=ws)
    torch.cuda.synchronize()
    return nnz, nn.cpu().numpy().view(np.uint32), ptr.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("workspace", [True, False])
def test_radius_csr_device_lidar(lidar250k, workspace):
    xyz = lidar250k
    nnz, nn, ptr = _radius_csr_device(xyz, 0.5, 64, workspace)
    oi, _, cnt = O.radius(xyz, xyz, 0.5, 64, kdtree=True)
    np.testing.assert_array_equal(ptr, np.r_[0, np.cumsum(np.minimum(cnt, 64))].astype(np.uint32))
    assert nnz == int(ptr[-1])
    np.testing.assert_array_equal(nn[:nnz], oi[oi >= 0].astype(np.uint32))
    # README.md:158-159 on the host API's output gives the same CSR
    hi, _ = pgeof.radius_search(xyz, xyz, 0.5, 64)
    hnn, hptr = scenes.radius_csr(hi)
    np.testing.assert_array_equal(ptr, hptr)
    np.testing.assert_array_equal(nn[:nnz], hnn)


def test_radius_csr_device_small_radius_and_empty():
    xyz = scenes.uniform_cube(5000, seed=72)
    nnz, nn, ptr = _radius_csr_device(xyz, 0.01, 8)
    oi, _, cnt = O.radius(xyz, xyz, 0.01, 8, kdtree=False)
    np.testing.assert_array_equal(ptr, np.r_[0, np.cumsum(np.minimum(cnt, 8))].astype(np.uint32))
    np.testing.assert_array_equal(nn[:nnz], oi[oi >= 0].astype(np.uint32))
    # max_knn = 0: every row empty
    nnz0, _, ptr0 = _radius_csr_device(xyz, 0.1, 0, workspace=False)
    assert nnz0 == 0 and (ptr0 == 0).all()


def _csr_device_all(xyz, nn, ptr, with_err):
    x = torch.from_numpy(xyz).to(DEV)
    nn_d = torch.from_numpy(nn.view(np.int32)).to(DEV)
    ptr_d = torch.from_numpy(ptr.view(np.int32)).to(DEV)
    n = len(ptr) - 1
    err = torch.zeros(1, dtype=torch.int32, device=DEV) if with_err else None
    f = torch.empty((n, 11), dtype=torch.float32, device=DEV)
    m = torch.empty((n, 3, 11), dtype=torch.float32, device=DEV)
    o = torch.empty((n, 12), dtype=torch.float32, device=DEV)
    capi.compute_features_device(x, nn_d, ptr_d, f, k_min=3, err=err)
    capi.compute_features_multiscale_device(x, nn_d, ptr_d, [5, 12, 30], m, err=err)
    capi.compute_features_optimal_device(x, nn_d, ptr_d, o, 4, 2, 6, err=err)
    torch.cuda.synchronize()
    e = int(err.item()) if with_err else 0
    return f.cpu().numpy(), m.cpu().numpy(), o.cpu().numpy(), e


@pytest.mark.parametrize("with_err", [True, False])
def test_csr_device_entry_points_equal_host(with_err):
    xyz = scenes.lidar_scene(90_000, seed=73, extent=28.0)
    idx, _ = pgeof.radius_search(xyz, xyz, 0.45, 40)
    nn, ptr = scenes.radius_csr(idx)
    f, m, o, e = _csr_device_all(xyz, nn, ptr, with_err)
    assert e == 0
    np.testing.assert_array_equal(f, pgeof.compute_features(xyz, nn, ptr, k_min=3))
    np.testing.assert_array_equal(m, pgeof.compute_features_multiscale(xyz, nn, ptr, [5, 12, 30]))
    np.testing.assert_array_equal(o, pgeof.compute_features_optimal(xyz, nn, ptr, 4, 2, 6))


def test_csr_device_error_word_and_raise():
    xyz = scenes.uniform_cube(3000, seed=74)
    idx, _ = pgeof.knn_search(xyz, xyz, 10)
    nn, ptr = scenes.knn_csr(idx)
    bad = nn.copy()
    bad[1234] = len(xyz) + 5
    _, _, _, e = _csr_device_all(xyz, bad, ptr, True)
    assert e & 1
    with pytest.raises(ValueError):
        _csr_device_all(xyz, bad, ptr, False)
    bad_ptr = ptr.copy()
    bad_ptr[10] = bad_ptr[11] + 1  # decreasing
    _, _, _, e = _csr_device_all(xyz, nn, bad_ptr, True)
    assert e & 2


def test_search_device_entry_points_equal_host():
    rng = np.random.default_rng(75)
    data = scenes.lidar_scene(60_000, seed=75, extent=22.0)
    q = np.concatenate([data[rng.choice(len(data), 20_000, replace=False)],
                        rng.uniform(-5, 27, size=(5000, 3)).astype(np.float32)])
    x = torch.from_numpy(data).to(DEV)
    qd = torch.from_numpy(q).to(DEV)
    k = 24
    idx = torch.empty((len(q), k), dtype=torch.int32, device=DEV)
    d2 = torch.empty((len(q), k), dtype=torch.float32, device=DEV)
    capi.knn_search_device(x, qd, k, idx, d2)
    ridx = torch.empty((len(q), 32), dtype=torch.int32, device=DEV)
    rd2 = torch.empty((len(q), 32), dtype=torch.float32, device=DEV)
    capi.radius_search_device(x, qd, 0.6, 32, ridx, rd2)
    torch.cuda.synchronize()
    gi, gd = pgeof.knn_search(data, q, k)
    np.testing.assert_array_equal(idx.cpu().numpy().view(np.uint32), gi)
    np.testing.assert_array_equal(d2.cpu().numpy(), gd)
    hi, hd = pgeof.radius_search(data, q, 0.6, 32)
    np.testing.assert_array_equal(ridx.cpu().numpy(), hi)
    np.testing.assert_array_equal(rd2.cpu().numpy(), hd)


def test_optimal_k100_precision_rows():
    """The precision workload's setting: kNN k = 100 rows, optimal k over
    10..100 (k_min = 10, k_step = 1, k_min_search = 10)."""
    xyz = scenes.precision_scene(30_000, seed=76)
    idx, _ = pgeof.knn_search(xyz, xyz, 100)
    oi, _ = O.knn(xyz, xyz, 100, kdtree=True)
    np.testing.assert_array_equal(idx, oi)
    nn, ptr = scenes.knn_csr(idx)
    g = pgeof.compute_features_optimal(xyz, nn, ptr, 10, 1, 10)
    ref = O.compute_features_optimal(xyz, nn, ptr, 10, 1, 10, eigentropy_fp32=True)
    agree = g[:, 11] == ref[:, 11]
    assert agree.mean() > 0.995, agree.mean()
    # every disagreement is a tie of the FP32 eigentropies of the two k
    assert optimal_disagreement_is_tie(xyz, nn, 100, np.nonzero(~agree)[0], g[~agree, 11], ref[~agree, 11])
    rows = np.nonzero(agree)[0]
    kopt = ref[agree, 11].astype(np.int64)
    val = np.zeros((len(rows), 3))
    for kk in np.unique(kopt):
        sel = kopt == kk
        r = rows[sel]
        sub_ptr = np.r_[0, np.full(len(r), kk).cumsum()].astype(np.uint32)
        sub_nn = nn.reshape(-1, 100)[r, :kk].ravel().copy()
        val[sel] = O.eigvals_csr_np(xyz, sub_nn, sub_ptr)
    assert_features_close(g[agree], ref[agree], val, "optimal k=100")
    assert ((g[:, 11] >= 10) & (g[:, 11] <= 100)).all()


@pytest.mark.parametrize("off_i,off_d", [(0, 0), (1, 1), (1, 2), (3, 0), (2, 2), (2, 0)])
def test_search_outputs_at_any_alignment(off_i, off_d):
    """Index and distance rows are written 16 bytes per store from each
    row's first 16-byte boundary, one by one when the two arrays are not
    aligned alike; outputs that start off a 16-byte boundary (views into
    larger buffers) give the host API's results, for kNN (k = 24, 50),
    radius search (max 30) and the fused kNN + features path (k = 50, an
    even k whose rows are written two entries per store when 8-byte aligned)."""
    data = scenes.lidar_scene(40_000, seed=76, extent=18.0)
    n = len(data)
    x = torch.from_numpy(data).to(DEV)

    def view(dtype, shape, off):
        buf = torch.full((int(np.prod(shape)) + 4,), -7, dtype=dtype, device=DEV)
        return buf[off:off + int(np.prod(shape))].view(*shape)

    for k in (24, 50):
        idx, d2 = view(torch.int32, (n, k), off_i), view(torch.float32, (n, k), off_d)
        capi.knn_search_device(x, x, k, idx, d2)
        gi, gd = pgeof.knn_search(data, data, k)
        np.testing.assert_array_equal(idx.cpu().numpy().view(np.uint32), gi)
        np.testing.assert_array_equal(d2.cpu().numpy(), gd)
    ridx, rd2 = view(torch.int32, (n, 30), off_i), view(torch.float32, (n, 30), off_d)
    capi.radius_search_device(x, x, 0.3, 30, ridx, rd2)
    hi, hd = pgeof.radius_search(data, data, 0.3, 30)
    np.testing.assert_array_equal(ridx.cpu().numpy(), hi)
    np.testing.assert_array_equal(rd2.cpu().numpy(), hd)
    k = 50
    nn, fd2 = view(torch.int32, (n * k,), off_i), view(torch.float32, (n * k,), off_d)
    ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    f = view(torch.float32, (n, 11), off_i)
    capi.knn_features_device(x, k, nn, ptr, f, sqr_dist=fd2)
    gi, gd = pgeof.knn_search(data, data, k)
    np.testing.assert_array_equal(nn.cpu().numpy().view(np.uint32), gi.ravel())
    np.testing.assert_array_equal(fd2.cpu().numpy(), gd.ravel())
    nn_h, ptr_h = scenes.knn_csr(gi)
    np.testing.assert_array_equal(f.cpu().numpy(), pgeof.compute_features(data, nn_h, ptr_h))
