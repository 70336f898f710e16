"""Benchmark of pgeof's hot path on one MI355X.

Headline (default): knn_search(k=20) -> CSR -> compute_features on a 10 M-point
synthetic LiDAR scene, inputs resident in HBM, through the fused device entry
point ``pgeof_knn_features_device`` (results equal the two-call path bit for
bit).  One step = one pass over the whole cloud (one "cloud" of the per-cloud
latency).  Prints ONE JSON line.

Other workloads (--workload): lidar10m_k50 (k=50), radius_multiscale_2m
(radius r=0.5 max 64 -> CSR -> multiscale [10,20,50,100]), optimal_1m
(precision scene, knn 100 -> optimal k 10..100), features_lidar10m_k20 (the
CSR feature kernel alone on a precomputed kNN CSR).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public spec


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lidar10m_k20")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--tune", action="append", default=[], help="library tuning knob key=value")
    return ap.parse_args()


def _pmc_traffic(kernel: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    calibrated on a copy kernel; committed summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    e = d.get(kernel)
    return None if e is None else float(e["bytes_per_launch"])


def _cpu_baseline(xyz: np.ndarray, k: int, sample: int):
    """The oracle's multithreaded C restatement (KD-tree kNN + FP64 features)
    on the host cores: KD-tree over all points + `sample` queries, scaled to
    the whole cloud.  Rank 0 only."""
    from oracle import oracle as O

    cores = min(16, len(os.sched_getaffinity(0)))
    rng = np.random.default_rng(0)
    s = np.sort(rng.choice(len(xyz), size=min(sample, len(xyz)), replace=False)).astype(np.uint32)
    _, _, t_build, t_query = O.knn_features_timed(xyz, s, k, n_threads=cores)
    total = t_build + t_query * len(xyz) / len(s)
    return {
        "value": len(xyz) / total / 1e6,
        "unit": "Mpoints/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"KD-tree (leaf 10) over all {len(xyz)} points ({t_build:.2f} s, 1 thread) + kNN k={k} and "
                   f"FP64 features for {len(s)} sampled queries ({t_query:.2f} s, {cores} threads), "
                   f"query time scaled to all points"),
    }


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported "
                                   "(the multi-GPU path is a design only, see DESIGN.md)"}))
        sys.exit(2)

    import point_geometric_features_amd as pgeof  # noqa: F401  (selects the shared HIP runtime)
    import torch

    from point_geometric_features_amd import capi, scenes

    for kv in args.tune:
        key, val = kv.split("=")
        capi.set_tuning(key, float(val))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    wl = args.workload
    cfg = {"workload": wl}
    if args.tune:
        cfg["tuning"] = dict(kv.split("=") for kv in args.tune)
    extra = {}

    if wl in ("lidar10m_k20", "lidar10m_k50"):
        k = 20 if wl.endswith("k20") else 50
        n = 10_000_000
        xyz = scenes.lidar_scene(n, seed=1)
        cfg.update({"points": n, "k": k, "scene": "synthetic LiDAR 200 m x 200 m (seed 1)",
                    "path": "knn_search + compute_features, fused, device-resident"})
        x = torch.from_numpy(xyz).to(dev)
        nn = torch.empty(n * k, dtype=torch.int32, device=dev)
        ptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        feat = torch.empty((n, 11), dtype=torch.float32, device=dev)
        ws = torch.empty(capi.knn_workspace_size(n, n, k), dtype=torch.uint8, device=dev)

        def step():
            capi.knn_features_device(x, k, nn, ptr, feat, workspace=ws, stream=stream)

        kernel = "knn_features_fused"
        bytes_per_point = 60 + 4 * k  # xyz 12 + nn_ptr 4 + nn 4k + features 44
        units = n
        metric_unit = "Mpoints/s"
    elif wl == "features_lidar10m_k20":
        k, n = 20, 10_000_000
        xyz = scenes.lidar_scene(n, seed=1)
        cfg.update({"points": n, "k": k, "path": "compute_features on a kNN CSR (CSR feature kernel only)"})
        x = torch.from_numpy(xyz).to(dev)
        nn = torch.empty(n * k, dtype=torch.int32, device=dev)
        ptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        feat = torch.empty((n, 11), dtype=torch.float32, device=dev)
        capi.knn_features_device(x, k, nn, ptr, feat)
        torch.cuda.synchronize()
        err = torch.zeros(1, dtype=torch.int32, device=dev)

        def step():
            capi.compute_features_device(x, nn, ptr, feat, k_min=1, err=err, stream=stream)

        kernel = "csr_features"
        bytes_per_point = 60 + 4 * k
        units = n
        metric_unit = "Mpoints/s"
    elif wl == "radius_multiscale_2m":
        n, r, cap, sc = 2_000_000, 0.5, 64, [10, 20, 50, 100]
        xyz = scenes.lidar_scene(n, seed=2, extent=200.0 * np.sqrt(n / 10_000_000))
        cfg.update({"points": n, "radius": r, "max_knn": cap, "scales": sc,
                    "path": "radius CSR + compute_features_multiscale, device-resident"})
        x = torch.from_numpy(xyz).to(dev)
        nn = torch.empty(n * cap, dtype=torch.int32, device=dev)
        ptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        feat = torch.empty((n, len(sc), 11), dtype=torch.float32, device=dev)
        ws = torch.empty(capi.radius_csr_workspace_size(n, cap), dtype=torch.uint8, device=dev)
        err = torch.zeros(1, dtype=torch.int32, device=dev)
        nnz_box = [0]

        def step():
            nnz = capi.radius_csr_device(x, r, cap, nn, ptr, workspace=ws, stream=stream)
            nnz_box[0] = nnz
            capi.compute_features_multiscale_device(x, nn[:nnz], ptr, sc, feat, err=err, stream=stream)

        kernel = "csr_features_multiscale"
        bytes_per_point = None
        units = n
        metric_unit = "Mpoints/s"
    elif wl == "optimal_1m":
        n, k = 1_000_000, 100
        xyz = scenes.precision_scene(n, seed=3)
        cfg.update({"points": n, "k": k, "k_min": 10, "k_step": 1, "k_min_search": 10,
                    "path": "knn k=100 -> compute_features_optimal, device-resident"})
        x = torch.from_numpy(xyz).to(dev)
        nn = torch.empty(n * k, dtype=torch.int32, device=dev)
        ptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        f11 = torch.empty((n, 11), dtype=torch.float32, device=dev)
        feat = torch.empty((n, 12), dtype=torch.float32, device=dev)
        ws = torch.empty(capi.knn_workspace_size(n, n, k), dtype=torch.uint8, device=dev)
        err = torch.zeros(1, dtype=torch.int32, device=dev)

        def step():
            capi.knn_features_device(x, k, nn, ptr, f11, workspace=ws, stream=stream)
            capi.compute_features_optimal_device(x, nn, ptr, feat, 10, 1, 10, err=err, stream=stream)

        kernel = "csr_features_optimal"
        bytes_per_point = 16 + 4 * k + 48
        units = n
        metric_unit = "Mpoints/s"
    else:
        raise SystemExit(f"unknown workload {wl}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    capi.profile_reset()
    capi.profile_enable(True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    capi.profile_enable(False)
    prof = capi.profile_report()

    total = t1 - t0
    ms_per_step = total / args.steps * 1e3
    value = units * args.steps / total / 1e6
    launches, kms = prof.get(kernel, (0, 0.0))
    avg_kernel_ms = kms / max(launches, 1)
    roofline = None
    if bytes_per_point is not None and launches:
        achieved = bytes_per_point * units / (avg_kernel_ms * 1e-3) / 1e9
        traffic = _pmc_traffic(kernel)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": kernel, "kernel_ms": round(avg_kernel_ms, 4),
                    "algorithmic_bytes_per_point": bytes_per_point}
    extra["kernels_ms_per_step"] = {kk: round(v[1] / max(v[0], 1), 4) for kk, v in sorted(prof.items())}
    lat_ms = np.array(lat) * 1e3
    extra["latency_ms"] = {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                           "p99": round(float(np.percentile(lat_ms, 99)), 3)}

    cpu = None
    if not args.no_cpu_baseline and wl in ("lidar10m_k20", "lidar10m_k50"):
        cpu = _cpu_baseline(xyz, k, args.cpu_sample)

    out = {
        "metric": "Mpoints/s knn+features (k=20, 10M pts, % HBM roofline); p50/p99 per-cloud latency",
        "value": round(value, 3),
        "unit": metric_unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (distances), f64 (moments, eigensolve)",
        "data": "synthetic",
        "config": cfg,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
