"""GPU test of pgeof_compute_features_device_noalloc, the entry point that
allocates nothing (its row-order scratch is the library's static device
array): through the C ABI, against pgeof_compute_features_device on the same
inputs, bit for bit, on two calls in a row (the second on a smaller input,
which must not see the first call's scratch contents), plus its argument
checks (err required, row limit) and error bits."""
import numpy as np
import pytest
import torch

from point_geometric_features_amd import capi, scenes

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _knn_csr(n, k, seed):
    xyz = scenes.lidar_scene(n, seed=seed, extent=200.0 * float(np.sqrt(n / 10_000_000)))
    x = torch.from_numpy(xyz).to(DEV)
    nn = torch.empty(n * k, dtype=torch.int32, device=DEV)
    ptr = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    f = torch.empty((n, 11), dtype=torch.float32, device=DEV)
    capi.knn_features_device(x, k, nn, ptr, f)
    torch.cuda.synchronize()
    return x, nn, ptr, f


def _both(x, nn, ptr, k_min=1):
    n = ptr.shape[0] - 1
    fa = torch.full((n, 11), float("nan"), dtype=torch.float32, device=DEV)
    fb = torch.full((n, 11), float("nan"), dtype=torch.float32, device=DEV)
    ea = torch.zeros(1, dtype=torch.int32, device=DEV)
    eb = torch.zeros(1, dtype=torch.int32, device=DEV)
    capi.compute_features_device(x, nn, ptr, fa, k_min=k_min, err=ea)
    capi.compute_features_device_noalloc(x, nn, ptr, fb, eb, k_min=k_min)
    torch.cuda.synchronize()
    return fa.cpu().numpy(), fb.cpu().numpy(), int(ea.item()), int(eb.item())


def test_noalloc_equals_compute_features_device_twice():
    # 400 k rows, then 90 k rows: both take the spatial row order (>= 65536
    # rows), i.e. the static scratch, the second with fewer rows
    for n, k, seed in [(400_000, 20, 81), (90_000, 16, 82)]:
        print(f"noalloc: n={n} k={k} seed={seed}", flush=True)
        x, nn, ptr, f_fused = _knn_csr(n, k, seed)
        a, b, ea, eb = _both(x, nn, ptr)
        assert ea == 0 and eb == 0
        np.testing.assert_array_equal(b, a)
        np.testing.assert_array_equal(b, f_fused.cpu().numpy())
        # k_min above the row length: every row zero
        a, b, _, _ = _both(x, nn, ptr, k_min=k + 1)
        np.testing.assert_array_equal(b, a)
        assert not b.any()


def test_noalloc_small_and_bad_index():
    # below 65536 rows (no row order), one neighbour index out of range: the
    # row is zeroed and bit 1 raised, as in the allocating entry point
    x, nn, ptr, _ = _knn_csr(5_000, 8, 83)
    nn[8 * 123 + 3] = 5_000
    a, b, ea, eb = _both(x, nn, ptr)
    assert ea == 1 and eb == 1
    np.testing.assert_array_equal(b, a)
    assert not b[123].any()


def test_noalloc_argument_checks():
    x, nn, ptr, _ = _knn_csr(2_000, 8, 84)
    f = torch.empty((2_000, 11), dtype=torch.float32, device=DEV)
    with pytest.raises(ValueError, match="err"):
        capi.check(capi.lib().pgeof_compute_features_device_noalloc(
            x.data_ptr(), 2_000, nn.data_ptr(), nn.shape[0], ptr.data_ptr(), ptr.shape[0], 1, f.data_ptr(), None,
            None))
    err = torch.zeros(1, dtype=torch.int32, device=DEV)
    # a row count above the limit is refused before anything is launched
    with pytest.raises(ValueError, match="PGEOF_NOALLOC_MAX_ROWS"):
        capi.check(capi.lib().pgeof_compute_features_device_noalloc(
            x.data_ptr(), 2_000, nn.data_ptr(), nn.shape[0], ptr.data_ptr(), capi.NOALLOC_MAX_ROWS + 2, 1,
            f.data_ptr(), err.data_ptr(), None))
    with pytest.raises(ValueError, match="k_min"):
        capi.compute_features_device_noalloc(x, nn, ptr, f, err, k_min=0)
