"""GPU tests of the one-call pipelines and of capturing them into HIP graphs.

* pgeof_radius_multiscale_device: radius_search -> CSR -> multiscale
  (include/pgeof.hpp:159-211, README.md:158-159) in one call, with the
  features from the search kernel (fuse_multiscale 1) or from the CSR kernel
  after it (0): nn / nn_ptr equal the oracle's radius CSR, features equal the
  two-call path bit for bit.
* pgeof_knn_optimal_device: knn_search -> CSR -> optimal k
  (include/pgeof.hpp:243-310) in one call, equal to the two-call path.
* The searches plan their grids on the device and queue every launch without
  a read-back: pgeof_knn_features_device and both pipelines, given a
  workspace, are captured into a graph (torch.cuda.CUDAGraph) and replayed
  several times, with eager work and synchronisations between the replays;
  every replay equals the eager result bit for bit.
"""
import numpy as np
import pytest
import torch

import point_geometric_features_amd as pgeof
from oracle import oracle as O
from point_geometric_features_amd import capi, scenes

pytestmark = pytest.mark.gpu

DEV = "cuda"
SCALES = [10, 20, 50, 100]


@pytest.fixture(scope="module")
def lidar120k():
    return scenes.lidar_scene(120_000, seed=91, extent=200.0 * np.sqrt(120_000 / 10_000_000))


def _radius_multiscale(xyz, r, cap, scales, fused, workspace=True):
    n = len(xyz)
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.full((n * cap,), -7, dtype=torch.int32, device=DEV)
    ptr = torch.full((n + 1,), -7, dtype=torch.int32, device=DEV)
    f = torch.full((n, len(scales), 11), float("nan"), dtype=torch.float32, device=DEV)
    ws = (torch.empty(capi.radius_multiscale_workspace_size(n, cap, len(scales)), dtype=torch.uint8, device=DEV)
          if workspace else None)
    old = capi.get_tuning("fuse_multiscale")
    try:
        capi.set_tuning("fuse_multiscale", fused)
        capi.radius_multiscale_device(x, r, cap, scales, nn, ptr, f, workspace=ws)
        torch.cuda.synchronize()
    finally:
        capi.set_tuning("fuse_multiscale", old)
    p = ptr.cpu().numpy().view(np.uint32)
    return nn.cpu().numpy().view(np.uint32)[: p[-1]], p, f.cpu().numpy()


@pytest.mark.parametrize("fused", [0, 1])
def test_radius_multiscale_device_equals_two_calls(lidar120k, fused):
    xyz = lidar120k
    nn, ptr, f = _radius_multiscale(xyz, 0.5, 64, SCALES, fused)
    oi, _, cnt = O.radius(xyz, xyz, 0.5, 64, kdtree=True)
    np.testing.assert_array_equal(ptr, np.r_[0, np.cumsum(np.minimum(cnt, 64))].astype(np.uint32))
    np.testing.assert_array_equal(nn, oi[oi >= 0].astype(np.uint32))
    np.testing.assert_array_equal(f, pgeof.compute_features_multiscale(xyz, nn, ptr, SCALES))
    # scale 100 exceeds every row (at most 64 hits): identically zero
    assert not f[:, 3].any()


@pytest.mark.parametrize("fused", [0, 1])
def test_radius_multiscale_device_edge_cases(fused):
    xyz = scenes.uniform_cube(4000, seed=92)
    # repeated scales, a scale of 1, no workspace (stream-ordered allocation)
    sc = [1, 3, 3, 8]
    nn, ptr, f = _radius_multiscale(xyz, 0.08, 12, sc, fused, workspace=False)
    np.testing.assert_array_equal(f, pgeof.compute_features_multiscale(xyz, nn, ptr, sc))
    # more than 16 scales (the values travel as kernel arguments, 16 per launch)
    sc = list(range(1, 21))
    nn, ptr, f = _radius_multiscale(xyz, 0.1, 24, sc, fused)
    np.testing.assert_array_equal(f, pgeof.compute_features_multiscale(xyz, nn, ptr, sc))
    with pytest.raises(ValueError):
        _radius_multiscale(xyz, 0.1, 24, [5, 3], fused)


def test_knn_optimal_device_equals_two_calls():
    xyz = scenes.precision_scene(80_000, seed=93)
    n, k = len(xyz), 40
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.empty(n * k, dtype=torch.int32, device=DEV)
    ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    f = torch.empty((n, 12), dtype=torch.float32, device=DEV)
    ws = torch.empty(capi.knn_optimal_workspace_size(n, k), dtype=torch.uint8, device=DEV)
    capi.knn_optimal_device(x, k, nn, ptr, f, 10, 1, 10, workspace=ws)
    torch.cuda.synchronize()
    gi, _ = pgeof.knn_search(xyz, xyz, k)
    hnn, hptr = scenes.knn_csr(gi)
    np.testing.assert_array_equal(nn.cpu().numpy().view(np.uint32), hnn)
    np.testing.assert_array_equal(ptr.cpu().numpy().view(np.uint32), hptr)
    np.testing.assert_array_equal(f.cpu().numpy(), pgeof.compute_features_optimal(xyz, hnn, hptr, 10, 1, 10))


def _replays_equal_eager(run, outs, n_replays=3):
    """Eager result, then a graph of the same call replayed n_replays times
    with the outputs poisoned by eager work before each replay."""
    run()
    torch.cuda.synchronize()
    eager = [o.clone() for o in outs]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run()  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    for _ in range(n_replays):
        for o in outs:
            o.fill_(-3)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e)


def test_graph_knn_features(lidar120k):
    xyz = lidar120k
    n, k = len(xyz), 20
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.empty(n * k, dtype=torch.int32, device=DEV)
    ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    f = torch.empty((n, 11), dtype=torch.float32, device=DEV)
    ws = torch.empty(capi.knn_workspace_size(n, n, k), dtype=torch.uint8, device=DEV)
    _replays_equal_eager(lambda: capi.knn_features_device(x, k, nn, ptr, f, workspace=ws), [nn, ptr, f])
    oi, _ = O.knn(xyz, xyz, k, kdtree=True)
    np.testing.assert_array_equal(nn.cpu().numpy().view(np.uint32), oi.ravel())


@pytest.mark.parametrize("fused", [0, 1])
def test_graph_radius_multiscale(lidar120k, fused):
    xyz = lidar120k
    n, cap = len(xyz), 64
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.empty(n * cap, dtype=torch.int32, device=DEV)
    ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    f = torch.empty((n, len(SCALES), 11), dtype=torch.float32, device=DEV)
    ws = torch.empty(capi.radius_multiscale_workspace_size(n, cap, len(SCALES)), dtype=torch.uint8, device=DEV)
    old = capi.get_tuning("fuse_multiscale")
    try:
        capi.set_tuning("fuse_multiscale", fused)
        _replays_equal_eager(lambda: capi.radius_multiscale_device(x, 0.5, cap, SCALES, nn, ptr, f, workspace=ws),
                             [ptr, f])
    finally:
        capi.set_tuning("fuse_multiscale", old)


def test_graph_knn_optimal():
    xyz = scenes.precision_scene(60_000, seed=94)
    n, k = len(xyz), 30
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.empty(n * k, dtype=torch.int32, device=DEV)
    ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    f = torch.empty((n, 12), dtype=torch.float32, device=DEV)
    ws = torch.empty(capi.knn_optimal_workspace_size(n, k), dtype=torch.uint8, device=DEV)
    _replays_equal_eager(lambda: capi.knn_optimal_device(x, k, nn, ptr, f, 10, 1, 10, workspace=ws), [nn, ptr, f])


def test_autotune_keeps_results(lidar120k):
    """The autotuner times a sweep of knob sets, keeps the fastest for the
    workload's family, k and size class; results with it are unchanged."""
    xyz = lidar120k
    n = len(xyz)
    x = torch.from_numpy(xyz).to(DEV)
    try:
        rep = capi.autotune("knn_features", x, 20, reps=2)
        assert rep["workload"] == "knn_features" and rep["k"] == 20 and rep["n"] == n
        assert len(rep["ms"]) == len(rep["candidates"]) > 1
        assert rep["best_ms"] == min(rep["ms"]) and rep["best_ms"] <= rep["default_ms"]
        nn = torch.empty(n * 20, dtype=torch.int32, device=DEV)
        ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
        f = torch.empty((n, 11), dtype=torch.float32, device=DEV)
        ws = torch.empty(capi.knn_workspace_size(n, n, 20), dtype=torch.uint8, device=DEV)
        capi.knn_features_device(x, 20, nn, ptr, f, workspace=ws)
        torch.cuda.synchronize()
        oi, _ = O.knn(xyz, xyz, 20, kdtree=True)
        np.testing.assert_array_equal(nn.cpu().numpy().view(np.uint32), oi.ravel())
        rep = capi.autotune("radius_multiscale", x, 64, radius=0.5, k_scales=SCALES, reps=1)
        assert rep["workload"] == "radius_multiscale" and len(rep["ms"]) > 1
        nn, ptr, f = _radius_multiscale(xyz, 0.5, 64, SCALES, rep["knobs"]["fuse_multiscale"])
        np.testing.assert_array_equal(f, pgeof.compute_features_multiscale(xyz, nn, ptr, SCALES))
        with pytest.raises(ValueError):
            capi.autotune("nope", x, 20)
    finally:
        capi.autotune_clear()
