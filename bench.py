"""Benchmark of pgeof's hot path on one MI355X.

Headline (default): knn_search(k=20) -> CSR -> compute_features on a 10 M-point
synthetic LiDAR scene, inputs resident in HBM, through the fused device entry
point ``pgeof_knn_features_device`` (results equal the two-call path bit for
bit).  One step = one pass over the whole cloud (one "cloud" of the per-cloud
latency).  Prints ONE JSON line.

Other workloads (--workload): lidar10m_k50 (k=50), radius_multiscale_2m
(radius r=0.5 max 64 -> CSR -> multiscale [10,20,50,100]), optimal_1m
(precision scene, knn 100 -> optimal k 10..100), features_lidar10m_k20 (the
CSR feature kernel alone on a precomputed kNN CSR), host_api_lidar10m_k20
(the Python API on host numpy arrays, PCIe copies included), serving (a
stream of clouds of 5k..2M points through the serving loop).
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
    ap.add_argument("--requests", type=int, default=200, help="serving workload: number of requests")
    return ap.parse_args()


def _pmc_traffic(kernel: str, workload: str):
    """HBM bytes per step of `kernel` on `workload`, measured with rocprofv3
    --pmc (FETCH_SIZE, WRITE_SIZE) and calibrated on a copy kernel of known
    size; read from the committed summary, or None when not measured."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    e = d.get(f"{kernel}@{workload}")
    return None if e is None else float(e["bytes_per_step"])


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


def _cpu_scipy(xyz: np.ndarray, k: int, sample: int):
    """scipy's cKDTree, the reference's own comparator (tests/bench_knn.py:13-21):
    tree over all points + kNN of `sample` queries on every host core, scaled
    to the whole cloud (kNN only, no features)."""
    from scipy.spatial import cKDTree

    cores = min(16, len(os.sched_getaffinity(0)))
    rng = np.random.default_rng(0)
    q = xyz[rng.choice(len(xyz), size=min(sample, len(xyz)), replace=False)]
    t0 = time.perf_counter()
    tree = cKDTree(xyz)
    t1 = time.perf_counter()
    tree.query(q, k, workers=cores)
    t2 = time.perf_counter()
    total = (t1 - t0) + (t2 - t1) * len(xyz) / len(q)
    return {"value": len(xyz) / total / 1e6, "unit": "Mpoints/s", "cores": cores,
            "sample": f"cKDTree build over all points ({t1 - t0:.2f} s) + k={k} query of {len(q)} points "
                      f"({t2 - t1:.2f} s), scaled; kNN only"}


def _copy_gbs(torch, dev) -> float:
    """Achievable HBM bandwidth of a plain 1 GiB device copy (read + write)."""
    a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        b.copy_(a)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / 10
    del a, b
    return 2.0 * (1 << 30) / (ms * 1e-3) / 1e9


def _crop(base: np.ndarray, extent: float, size: int, rng) -> np.ndarray:
    """A square window of the base scene holding about `size` points."""
    side = extent * min(1.0, float(np.sqrt(size / len(base))))
    cx, cy = rng.uniform(side / 2, extent - side / 2, size=2) if side < extent else (extent / 2, extent / 2)
    m = (np.abs(base[:, 0] - cx) <= side / 2) & (np.abs(base[:, 1] - cy) <= side / 2)
    pts = base[m][:size]
    return np.ascontiguousarray(pts)


def _serving(args, capi, scenes):
    """Config 5: a stream of requests (log-uniform sizes 5k..2M, Poisson
    arrivals at 0.8 of the measured capacity), knn k=20 + features each."""
    import torch  # noqa: F401  (device already initialised by the caller)

    k = 20
    n_req = args.requests
    extent = 200.0 * float(np.sqrt(2_000_000 / 10_000_000))
    base = scenes.lidar_scene(2_000_000, seed=4, extent=extent)
    sizes = scenes.request_sizes(n_req, seed=4)
    rng = np.random.default_rng(4)
    clouds = [_crop(base, extent, int(sz), rng) for sz in sizes]
    clouds = [c for c in clouds if len(c) >= k]
    n_pts = sum(len(c) for c in clouds)
    budget = 64 << 30

    def run(n_slots, arrivals=None, sequential=False):
        """sequential: each request is submitted when the previous one has
        completed (the single-request path); otherwise all are submitted at
        once (closed loop) or at `arrivals` (open loop)."""
        lat = {}
        with capi.Server(n_slots, budget, k) as srv:
            outs = {}
            t0 = time.perf_counter()
            pending = 0
            for i, c in enumerate(clouds):
                if sequential and pending:
                    for rid, st, ms in srv.wait(1, timeout_ms=60_000):
                        assert st == 0, st
                        lat[rid] = ms
                        pending -= 1
                        outs.pop(rid, None)
                if arrivals is not None:
                    dt = arrivals[i] - (time.perf_counter() - t0)
                    if dt > 0:
                        # drain completions while waiting for the next arrival
                        for rid, st, ms in srv.wait(64, timeout_ms=dt * 1e3):
                            assert st == 0, st
                            lat[rid] = ms
                            pending -= 1
                            outs.pop(rid, None)
                outs[i] = np.empty((len(c), 11), np.float32)
                srv.submit(i, c, outs[i])
                pending += 1
            while pending:
                for rid, st, ms in srv.wait(64, timeout_ms=60_000):
                    assert st == 0, st
                    lat[rid] = ms
                    pending -= 1
                    outs.pop(rid, None)
            wall = time.perf_counter() - t0
        return wall, np.array([lat[i] for i in range(len(clouds))])

    run(4)  # warm-up (allocations, code objects)
    wall_seq, lat_seq = run(1, sequential=True)
    wall_cap, _ = run(4)
    cap_rate = n_pts / wall_cap  # points/s, closed loop
    rate_req = 0.8 * cap_rate / (n_pts / len(clouds))
    gaps = np.random.default_rng(5).exponential(1.0 / rate_req, size=len(clouds))
    arrivals = np.cumsum(gaps) - gaps[0]
    wall_open, lat_open = run(4, arrivals)
    return {
        "requests": len(clouds), "points": int(n_pts), "slots": 4, "hbm_budget_gb": budget >> 30,
        "capacity_mpts_s": round(cap_rate / 1e6, 2),
        "one_at_a_time_mpts_s": round(n_pts / wall_seq / 1e6, 2),
        # service time of one request on an idle GPU (host copies included)
        "one_at_a_time_latency_ms": {"p50": round(float(np.percentile(lat_seq, 50)), 2),
                                     "p99": round(float(np.percentile(lat_seq, 99)), 2)},
        "open_loop_rate_req_s": round(rate_req, 2),
        "open_loop_mpts_s": round(n_pts / wall_open / 1e6, 2),
        "open_loop_latency_ms": {"p50": round(float(np.percentile(lat_open, 50)), 2),
                                 "p99": round(float(np.percentile(lat_open, 99)), 2)},
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

    if wl == "serving":
        res = _serving(args, capi, scenes)
        out = {"metric": "Mpoints/s knn+features (k=20) over a request stream; p50/p99 per-cloud latency",
               "value": res["open_loop_mpts_s"], "unit": "Mpoints/s", "n_gpus": 1, "steps": res["requests"],
               "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
               "config": {"workload": wl, **res}, "roofline": None, "cpu_baseline": None}
        print(json.dumps(out))
        return

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

        # one search kernel, launched as the budgeted pass over every point
        # and the unbudgeted pass over the deferred ones: its time per step is
        # the sum of both launches
        kernel = ("knn_features_fused_dense", "knn_features_fused")
        bytes_per_point = 60 + 4 * k  # xyz 12 + nn_ptr 4 + nn 4k + features 44
        units = n
        metric_unit = "Mpoints/s"
    elif wl == "host_api_lidar10m_k20":
        # the Python API on host numpy arrays: PCIe copies in both directions
        # are inside the timed region (knn_search returns 1.6 GB of indices
        # and distances; compute_features takes the 0.8 GB CSR back)
        k, n = 20, 10_000_000
        xyz = scenes.lidar_scene(n, seed=1)
        cfg.update({"points": n, "k": k, "path": "pgeof.knn_search + pgeof.compute_features on host arrays"})

        def step():
            idx, _ = pgeof.knn_search(xyz, xyz, k)
            nn, ptr = scenes.knn_csr(idx)
            pgeof.compute_features(xyz, nn, ptr)

        kernel = None
        bytes_per_point = None
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
    names = kernel if isinstance(kernel, tuple) else ((kernel,) if kernel else ())
    launches = sum(prof.get(kk, (0, 0.0))[0] for kk in names)
    # kernel time per step (all launches of the kernel within one step)
    avg_kernel_ms = sum(prof.get(kk, (0, 0.0))[1] for kk in names) / args.steps
    kernel = "+".join(names)
    roofline = None
    if bytes_per_point is not None and launches:
        achieved = bytes_per_point * units / (avg_kernel_ms * 1e-3) / 1e9
        traffic = _pmc_traffic(kernel, wl)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": kernel, "kernel_ms": round(avg_kernel_ms, 4),
                    "algorithmic_bytes_per_point": bytes_per_point}
    extra["kernels_ms_per_step"] = {kk: round(v[1] / args.steps, 4) for kk, v in sorted(prof.items())}
    extra["kernel_launches_per_step"] = {kk: round(v[0] / args.steps, 2) for kk, v in sorted(prof.items())}
    lat_ms = np.array(lat) * 1e3
    extra["latency_ms"] = {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                           "p99": round(float(np.percentile(lat_ms, 99)), 3)}

    cpu = None
    if not args.no_cpu_baseline and wl in ("lidar10m_k20", "lidar10m_k50"):
        cpu = _cpu_baseline(xyz, k, args.cpu_sample)
        extra["cpu_scipy_knn"] = _cpu_scipy(xyz, k, args.cpu_sample)
    if roofline is not None:
        roofline["copy_gbs"] = round(_copy_gbs(torch, dev), 1)

    if wl == "optimal_1m" and launches:
        # k = 10..100 step 1: 91 eigenvalue solves and one full PCA per point
        extra["eigensolves_per_s"] = round(units * 92 / (avg_kernel_ms * 1e-3), 1)
    metric = {
        "lidar10m_k20": "Mpoints/s knn+features (k=20, 10M pts, % HBM roofline); p50/p99 per-cloud latency",
    }.get(wl, f"Mpoints/s, workload {wl}")
    out = {
        "metric": metric,
        "value": round(value, 3),
        "unit": metric_unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic",
        "config": cfg,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
