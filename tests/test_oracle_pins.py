"""Pins the CPU oracle before it is trusted as the checker.

The reference cannot be built here (its Eigen / nanoflann / Taskflow
submodules are empty and nanobind is absent), and its own tests hold no golden
values, so the oracle is pinned by:
  * closed forms of the eigen-decomposition and of the feature formulas
    (include/pca.hpp:71-200) on covariances with known spectra;
  * two independent eigensolvers: the C oracle's cyclic Jacobi vs LAPACK
    (numpy.linalg.eigh) in the numpy restatement;
  * brute force vs the KD-tree search vs scipy's KD-tree -- the comparator of
    the reference's own tests (tests/test_pgeof.py:8-27 upstream);
  * a pure-Python loop restatement of the multiscale and optimal drivers
    (include/pgeof.hpp:159-211, 243-310) on small inputs.
Each check is chosen so that a plausible mistake (descending instead of
ascending eigen order, a dropped epsilon, a swapped eigenvector, a wrong tie
rule, an off-by-one in the k loop) fails it.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from oracle import oracle as O

EPS = 1e-3


def _pack(C):
    return np.array([C[0, 0], C[0, 1], C[0, 2], C[1, 1], C[1, 2], C[2, 2]])


@pytest.mark.parametrize("seed", range(8))
def test_eig3_rotated_known_spectrum(seed):
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(0.01, 3.0, 3))
    R = Rotation.random(random_state=seed).as_matrix()
    C = R @ np.diag(lam) @ R.T
    val, vec = O.eig3(_pack(C))
    np.testing.assert_allclose(val, lam, rtol=1e-12, atol=1e-14)  # ascending (Eigen's order)
    for i in range(3):
        # eigenvector i spans column i of R (up to sign)
        assert abs(abs(vec[i] @ R[:, i]) - 1.0) < 1e-10
    assert vec[2, 2] >= 0  # pca.hpp:96


def test_eig3_degenerate_bases_follow_eigen():
    # zero matrix: identity basis, v2 = e_z
    val, vec = O.eig3(np.zeros(6))
    np.testing.assert_array_equal(val, [0, 0, 0])
    np.testing.assert_array_equal(vec, np.eye(3))
    # diag(s, 0, 0): Eigen's selection sort yields (e_y, e_z, e_x)
    val, vec = O.eig3([2.0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(val, [0, 0, 2.0])
    np.testing.assert_array_equal(np.abs(vec), [[0, 1, 0], [0, 0, 1], [1, 0, 0]])


def _features_closed_form(lam, v2):
    s = np.sqrt(lam)
    f = np.zeros(11)
    fact = 1.0 / (s[0] + EPS)
    f[0] = (s[0] - s[1]) * fact
    f[1] = (s[1] - s[2]) * fact
    f[2] = s[2] * fact
    f[4:7] = v2
    f[7] = s[0]
    f[8] = np.sqrt(s[0] * s[1] + 1e-6)
    f[9] = (s[0] * s[1] * s[2] + 1e-9) ** (1.0 / 3.0)
    f[10] = s[2] / (s.sum() + EPS)
    return f


def test_features_closed_forms_collinear_planar_zero():
    # collinear along x, variance 1.25 (closed form with Eigen's ascending order: lambda = (0, 0, 1.25))
    line = np.zeros((4, 3), np.float32)
    line[:, 0] = [-1.5, -0.5, 0.5, 1.5]
    # planar z = 2 with variances 1 (x) and 4 (y)
    plane = np.array([[a, b, 2.0] for a in (-1.0, 1.0) for b in (-2.0, 2.0)], np.float32)
    xyz = np.concatenate([line, plane, [[7.0, 8.0, 9.0]]]).astype(np.float32)
    nn = np.arange(9, dtype=np.uint32)
    ptr = np.array([0, 4, 8, 9], np.uint32)
    f = O.compute_features(xyz, nn, ptr)
    np.testing.assert_allclose(f[0], _features_closed_form(np.array([0, 0, 1.25]), [1, 0, 0]), atol=1e-12)
    np.testing.assert_allclose(f[1], _features_closed_form(np.array([0, 1.0, 4.0]), [0, 1, 0]), atol=1e-9)
    np.testing.assert_allclose(f[2], [0, 0, 0, 0, 0, 0, 1, 0, 1e-3, 1e-3, 0], atol=1e-15)


def test_features_full_rank_closed_form_with_verticality():
    # axis-aligned box of 8 points: covariance diag(a^2, b^2, c^2), eigenvectors = axes
    a, b, c = 0.5, 1.0, 2.0
    xyz = np.array([[x, y, z] for x in (-a, a) for y in (-b, b) for z in (-c, c)], np.float32)
    f = O.compute_features(xyz, np.arange(8, dtype=np.uint32), np.array([0, 8], np.uint32))[0]
    lam = np.array([a * a, b * b, c * c])
    expect = _features_closed_form(lam, [0, 0, 1])
    # VerticalityPGEOF: u = sum_i lam_i |v_i| = (a^2, b^2, c^2) -> u_z / |u|
    u = lam
    expect[3] = u[2] / np.linalg.norm(u)
    np.testing.assert_allclose(f, expect, rtol=1e-12, atol=1e-12)


def test_c_oracle_vs_numpy_eigh_random_neighbourhoods():
    rng = np.random.default_rng(3)
    xyz = (rng.normal(size=(6000, 3)) * [3.0, 1.0, 0.2]).astype(np.float32)
    idx = rng.integers(0, len(xyz), size=(2000, 25)).astype(np.uint32)
    nn = idx.ravel()
    ptr = (np.arange(2001) * 25).astype(np.uint32)
    fc = O.compute_features(xyz, nn, ptr)
    fn = O.compute_features_np(xyz, nn, ptr)
    np.testing.assert_allclose(fc, fn, rtol=1e-9, atol=1e-9)


def test_feature_invariants_translation_and_row_order():
    rng = np.random.default_rng(4)
    xyz = rng.normal(size=(500, 3)).astype(np.float32)
    idx = np.stack([rng.permutation(500)[:30] for _ in range(200)]).astype(np.uint32)
    ptr = (np.arange(201) * 30).astype(np.uint32)
    f = O.compute_features(xyz, idx.ravel(), ptr)
    g = O.compute_features(xyz, idx[:, ::-1].ravel().copy(), ptr)
    np.testing.assert_allclose(f, g, rtol=1e-9, atol=1e-9)
    h = O.compute_features((xyz + np.float32(4.0)).astype(np.float32), idx.ravel(), ptr)
    np.testing.assert_allclose(f, h, rtol=1e-4, atol=1e-5)  # float32 translation rounding only


def test_knn_bruteforce_kdtree_scipy_agree():
    rng = np.random.default_rng(5)
    xyz = rng.uniform(0.0, 200.0, size=(5000, 3)).astype(np.float32)
    bi, bd = O.knn(xyz, xyz, 12, kdtree=False)
    ki, kd = O.knn(xyz, xyz, 12, kdtree=True)
    np.testing.assert_array_equal(bi, ki)
    np.testing.assert_array_equal(bd, kd)
    _, si = cKDTree(xyz).query(xyz, k=12)
    np.testing.assert_array_equal(bi, si)
    assert (bi[:, 0] == np.arange(len(xyz))).all()
    # distances are the float32 expression ((dx*dx + dy*dy) + dz*dz)
    q = xyz[:, None, :]
    p = xyz[bi]
    d = (q - p).astype(np.float32)
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]).astype(np.float32) + d[..., 2] * d[..., 2]
    np.testing.assert_array_equal(bd, d2.astype(np.float32))


def test_knn_ties_broken_by_index():
    xyz = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, 0, -1], [5, 5, 5]], np.float32)
    idx, d = O.knn(xyz, xyz[:1], 5, kdtree=False)
    np.testing.assert_array_equal(idx[0], [0, 1, 2, 3, 4])
    idx2, _ = O.knn(xyz, xyz[:1], 5, kdtree=True)
    np.testing.assert_array_equal(idx2, idx)


def test_radius_semantics_and_scipy():
    rng = np.random.default_rng(6)
    xyz = rng.random(size=(3000, 3), dtype=np.float32)
    bi, bd, cnt = O.radius(xyz, xyz, 0.1, 16, kdtree=False)
    ki, kd, _ = O.radius(xyz, xyz, 0.1, 16, kdtree=True)
    np.testing.assert_array_equal(bi, ki)
    np.testing.assert_array_equal(bd, kd)
    _, si = cKDTree(xyz).query(xyz, k=16, distance_upper_bound=0.1)
    si[si == len(xyz)] = -1
    np.testing.assert_array_equal(bi, si)
    r2 = np.float32(0.1) * np.float32(0.1)
    assert (bd[bi >= 0] < r2).all()
    notfull = cnt < 16
    assert (bd[notfull, -1] == r2).all()  # RKNNResultSet::init leaves r^2 in the last slot
    for row in np.nonzero(notfull)[0][:50]:
        c = cnt[row]
        assert (bi[row, c:] == -1).all() and (bd[row, c:-1] == 0).all()


def _py_optimal(xyz, nn, ptr, k_min, k_step, k_min_search):
    """Pure-Python loop restatement of include/pgeof.hpp:262-299 with FP32
    eigentropy (pca.hpp:140-150) on numpy-eigh eigenvalues."""
    n = len(ptr) - 1
    out = np.zeros((n, 12))
    for i in range(n):
        k_nn = int(ptr[i + 1] - ptr[i])
        if not (k_nn >= k_min and k_nn >= k_min_search):
            continue
        k0 = min(max(k_min, k_min_search), k_nn)
        best = None
        for k in range(k0, k_nn + 1):
            if k > k0 and k % k_step != 0 and k != k_nn:
                continue
            pts = xyz[nn[ptr[i]:ptr[i] + k]].astype(np.float64)
            d = pts - pts.mean(0)
            w = np.maximum(np.linalg.eigvalsh(d.T @ d / k), 0)
            h = O.eigentropy32_np(w[None])[0]
            if k == k0 or h < best[0]:
                best = (h, k)
        k = best[1]
        out[i, :11] = O.compute_features_np(xyz, nn[ptr[i]:ptr[i] + k], np.array([0, k], np.uint32))[0]
        out[i, 11] = k
    return out


def test_optimal_driver_vs_python_loop():
    rng = np.random.default_rng(7)
    xyz = (rng.normal(size=(400, 3)) * [1.0, 0.5, 0.05]).astype(np.float32)
    idx, _ = O.knn(xyz, xyz, 20, kdtree=False)
    nn, ptr = idx.ravel(), (np.arange(401) * 20).astype(np.uint32)
    for args in [(5, 1, 5), (1, 3, 4), (8, 4, 2)]:
        got = O.compute_features_optimal(xyz, nn, ptr, *args, eigentropy_fp32=True)
        exp = _py_optimal(xyz, nn, ptr, *args)
        np.testing.assert_array_equal(got[:, 11], exp[:, 11])
        np.testing.assert_allclose(got[:, :11], exp[:, :11], rtol=1e-8, atol=1e-8)


def test_multiscale_driver_semantics():
    rng = np.random.default_rng(8)
    xyz = rng.normal(size=(300, 3)).astype(np.float32)
    lens = rng.integers(0, 15, size=100)
    ptr = np.r_[0, np.cumsum(lens)].astype(np.uint32)
    nn = rng.integers(0, 300, size=int(ptr[-1])).astype(np.uint32)
    sc = [2, 5, 5, 9]
    m = O.compute_features_multiscale(xyz, nn, ptr, sc)
    for s, k in enumerate(sc):
        for i in range(100):
            if lens[i] < k:
                # pgeof.hpp:193-196: stop at the first scale the row cannot fill
                assert (m[i, s:] == 0).all()
                break
            exp = O.compute_features_np(xyz, nn[ptr[i]:ptr[i] + k], np.array([0, k], np.uint32))[0]
            # k = 2 rows are rank-deficient: lambda_0 is rounding noise (1e-17)
            np.testing.assert_allclose(m[i, s], exp, rtol=1e-9, atol=1e-4 if k < 4 else 1e-9)
    with pytest.raises(ValueError):
        O.compute_features_multiscale(xyz, nn, ptr, [5, 2])
    with pytest.raises(ValueError):
        O.compute_features_multiscale(xyz, nn, ptr, [0, 2])


def test_plausible_mistakes_are_caught():
    """A descending eigen-order or a dropped epsilon in the feature formulas
    must fail the closed-form check above."""
    lam = np.array([0.25, 1.0, 4.0])
    good = _features_closed_form(lam, [0, 0, 1])
    swapped = _features_closed_form(lam[::-1].copy(), [0, 0, 1])
    assert np.abs(good - swapped).max() > 1.0
    s = np.sqrt(lam)
    no_eps = good.copy()
    no_eps[2] = s[2] / s[0]
    assert abs(no_eps[2] - good[2]) > 1e-3


def test_eigentropy_closed_forms_and_shannon_entropy():
    """Eigentropy (include/pca.hpp:140-150) against what the mathematics
    fixes: for lambda = (1, 1, 1) it is ln 3 shifted by the epsilons, for a
    single non-zero eigenvalue it is -e ln(e + eps) of e = s / (s + eps), and
    on any spectrum it is the Shannon entropy of the normalised eigenvalues
    (scipy.stats.entropy, natural log) up to terms of the order of eps.  A
    log10, a dropped minus sign or a missing epsilon in the sum fails one of
    these."""
    from scipy.stats import entropy

    e = 1.0 / (3.0 + EPS)
    want = -3.0 * e * np.log(e + EPS)
    for h in (O.eigentropy_np(np.array([[1.0, 1.0, 1.0]]))[0], O.eigentropy32_np(np.array([[1.0, 1.0, 1.0]]))[0]):
        assert abs(h - want) < 2e-7
    assert abs(want - np.log(3.0)) < 5e-3  # epsilons aside, ln 3
    s2 = 0.7
    es = s2 / (s2 + EPS)
    want1 = -es * np.log(es + EPS)  # the two zero eigenvalues contribute 0 * ln(eps) = 0
    assert abs(O.eigentropy_np(np.array([[0.0, 0.0, s2]]))[0] - want1) < 1e-15
    rng = np.random.default_rng(11)
    lam = np.sort(rng.uniform(0.05, 4.0, size=(500, 3)), axis=1)
    h = O.eigentropy_np(lam)
    hs = entropy(lam.T)
    assert np.abs(h - hs).max() < 5e-3
    assert np.abs(O.eigentropy32_np(lam) - h).max() < 1e-6


def test_selected_verticality_and_eigentropy_closed_forms():
    """compute_features_selected's own outputs (include/pca.hpp:284 Verticality
    = 1 - |v2.z|, :287 Eigentropy) on a rotated box of 8 points whose
    covariance is R diag(a^2, b^2, c^2) R^T, in the float and the double
    overload of the oracle: every point's neighbourhood is the whole box."""
    a, b, c = 0.3, 0.8, 1.6
    box = np.array([[x, y, z] for x in (-a, a) for y in (-b, b) for z in (-c, c)])
    R = Rotation.from_euler("xyz", [0.4, -0.7, 0.25]).as_matrix()
    pts = box @ R.T + [10.0, -3.0, 2.0]
    lam = np.array([a * a, b * b, c * c])
    v2 = R[:, 2] * np.sign(R[2, 2])  # major axis, z >= 0 (pca.hpp:96)
    es = lam / (lam.sum() + EPS)
    want = {12: 1.0 - abs(v2[2]), 13: -(es * np.log(es + EPS)).sum(), 4: v2[0], 5: v2[1], 6: v2[2]}
    for dt, tol in ((np.float64, 1e-9), (np.float32, 2e-5)):
        got = O.compute_features_selected(pts.astype(dt), 100.0, 8, list(want))
        for j, (fid, w) in enumerate(want.items()):
            np.testing.assert_allclose(got[:, j], w, atol=tol, err_msg=f"feature {fid} {dt.__name__}")


def test_selected_rows_equal_the_full_computation():
    """The oracle's row-sampled compute_features_selected (used by the bench's
    error check) gives the rows of the full computation bit for bit, in both
    precisions, including rows with fewer than 2 hits (zeros)."""
    rng = np.random.default_rng(61)
    xyz = rng.random((1500, 3)) * [1.0, 1.0, 0.3]
    fids = [12, 10, 13, 0, 6]
    rows = np.array([0, 7, 1499, 700, 7, 31], dtype=np.uint32)
    for dt in (np.float64, np.float32):
        x = xyz.astype(dt)
        full = O.compute_features_selected(x, 0.08, 16, fids)
        assert (full == 0).all(axis=1).any() and not (full == 0).all()
        np.testing.assert_array_equal(O.compute_features_selected(x, 0.08, 16, fids, rows=rows), full[rows])
