"""Writes tests/golden/pgeof_golden.npz: inputs and expected outputs of the
hot path on a small seeded cloud, computed by the CPU oracle.

The reference holds no golden values of its own (its tests compare against
scipy or against itself, tests/test_pgeof.py upstream) and it cannot be built
here, so these vectors pin the oracle and the GPU path to each other across
machines.  Regenerate with: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    xyz = (rng.random((1000, 3), dtype=np.float32) * np.float32(2.0)).astype(np.float32)
    k = 16
    knn_idx, knn_d2 = O.knn(xyz, xyz, k, kdtree=False)
    nn = knn_idx.ravel().copy()
    ptr = (np.arange(len(xyz) + 1) * k).astype(np.uint32)
    feat = O.compute_features(xyz, nn, ptr)
    scales = np.array([4, 8, 16], np.uint32)
    multi = O.compute_features_multiscale(xyz, nn, ptr, scales)
    opt = O.compute_features_optimal(xyz, nn, ptr, 4, 2, 6, eigentropy_fp32=True)
    r, cap = 0.16, 12
    rad_idx, rad_d2, _ = O.radius(xyz, xyz, r, cap, kdtree=False)
    rnn = rad_idx[rad_idx >= 0].astype(np.uint32)
    rptr = np.r_[0, (rad_idx >= 0).sum(1).cumsum()].astype(np.uint32)
    rfeat = O.compute_features(xyz, rnn, rptr)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pgeof_golden.npz")
    np.savez_compressed(out, xyz=xyz, k=k, knn_idx=knn_idx, knn_d2=knn_d2, nn=nn, nn_ptr=ptr, features=feat,
                        scales=scales, multiscale=multi, optimal=opt, optimal_args=np.array([4, 2, 6]),
                        radius=np.float32(r), max_knn=cap, radius_idx=rad_idx, radius_d2=rad_d2,
                        radius_nn=rnn, radius_nn_ptr=rptr, radius_features=rfeat)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
