#!/usr/bin/env python3
"""Benchmark of the LBVH hot path on one MI355X.

Workload (one "step"): one LBVH build of 2^24 triangles from a Gaussian-cluster
soup (64 clusters, sigma 0.02, edge lengths 1e-3..1e-2, seed 1) with 30-bit
Morton codes (10 bits per axis). Inputs are resident in HBM before timing;
W untimed builds, then K timed builds bracketed by device synchronisation.
`value` is Mtri/s of the build. The same run also traces 2^22 random rays for
closest hit against the 2^24 BVH (Mrays/s), reports per-stage HIP-event
times, the roofline of the dominant kernel, and the CPU oracle's build on the
same soup as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Algorithmic bytes per triangle per build stage (scene box, Morton, radix
# sort, hierarchy = Karras topology + refit), as itemised in DESIGN.md.
STAGE_BYTES = {"bounds": 36.0, "morton": 40.0, "sort": 64.0, "hierarchy": 184.0}
STAGE_LAUNCHES = {"bounds": 2, "morton": 1, "sort": 4, "hierarchy": 1}
STAGE_KERNEL = {"bounds": "k_bounds", "morton": "k_morton", "sort": "k_onesweep",
                "hierarchy": "k_hierarchy"}
# Traversal work model per internal-node visit (V) and triangle test (T).
TRACE_FLOPS_V, TRACE_FLOPS_T = 40.0, 51.0
TRACE_BYTES_V, TRACE_BYTES_T = 56.0, 40.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_traffic():
    """HBM bytes per launch of the dominant kernel, measured by rocprofv3 PMC
    runs and committed under profiles/ (None if absent)."""
    p = ROOT / "profiles" / "lbvh_16m_pmc.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1 << 24)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--bits", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-trace", action="store_true", help="skip the closest-hit leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one plain device copy of this size (PMC calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path's "
              f"multi-GPU sharding is a design note in DESIGN.md, not a measured mode.",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import clear_amd as ca

    n, bits, K, W = args.tris, args.bits, args.steps, args.warmup
    log(f"generating {n} triangles")
    soup = ca.gen_cluster_soup(n, seed=1)
    dev = torch.device("cuda:0")
    tsoup = torch.from_numpy(soup).to(dev)
    inodes, leaves = ca.node_buffers(n, dev)
    ctx = ca.Context(0)
    builder = ca.LBVHBuilder(ctx, n, bits)
    stream = torch.cuda.current_stream()

    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    log(f"warmup {W}")
    for _ in range(W):
        builder.perform_soup(tsoup, n, bits, inodes, leaves, stream)
    torch.cuda.synchronize()

    # Timed region; stage boundaries carry HIP events recorded on the build's
    # own stream by the library (negligible cost).
    ctx.set_profiling(True)
    log(f"timing {K} builds")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        builder.perform_soup(tsoup, n, bits, inodes, leaves, stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stages_ms = ctx.last_profile()
    ctx.set_profiling(False)
    ms_per_step = dt / K * 1e3
    mtri_s = n * K / dt / 1e6

    names = ["bounds", "morton", "sort", "hierarchy"]
    stage_info = {}
    for name, ms in zip(names, stages_ms):
        gbs = STAGE_BYTES[name] * n / (ms * 1e-3) / 1e9
        stage_info[name] = {"ms": round(ms, 4), "GB_s": round(gbs, 1),
                            "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                            "launches": STAGE_LAUNCHES[name]}
    dom = max(stage_info, key=lambda k: stage_info[k]["ms"])
    dom_ms_launch = stage_info[dom]["ms"] / STAGE_LAUNCHES[dom]
    dom_bytes_launch = STAGE_BYTES[dom] * n / STAGE_LAUNCHES[dom]
    achieved = dom_bytes_launch / (dom_ms_launch * 1e-3) / 1e9
    traffic = None
    tr = load_traffic()
    if tr and tr.get("kernel") == STAGE_KERNEL[dom] and tr.get("tris") == n:
        traffic = tr.get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": STAGE_KERNEL[dom], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "algorithmic_bytes_per_launch": dom_bytes_launch,
                "ms_per_launch": round(dom_ms_launch, 4)}
    whole_gbs = sum(STAGE_BYTES.values()) * n / (ms_per_step * 1e-3) / 1e9

    # Closest-hit leg: 2^22 random rays against the 2^24 BVH.
    trace = None
    if not args.no_trace:
        R = args.rays
        rays = torch.from_numpy(ca.gen_random_rays(R, seed=2)).to(dev)
        hit = torch.empty(R, dtype=torch.int32, device=dev)
        ht = torch.empty(R, dtype=torch.float32, device=dev)
        st = torch.empty(2 * R, dtype=torch.int32, device=dev)
        verts, faces = ca.soup_to_mesh(soup)
        tv = torch.from_numpy(verts).to(dev)
        tf = torch.from_numpy(faces.view(np.uint8)).to(dev)
        ca.trace_closest(ctx, inodes, leaves, n, tv, tf, rays, R, hit, ht, st)
        torch.cuda.synchronize()
        stv = st.view(R, 2).to(torch.int64)
        V = int(stv[:, 0].sum().item())
        T = int(stv[:, 1].sum().item())
        hits = int((hit != -1).sum().item())
        for _ in range(2):
            ca.trace_closest(ctx, inodes, leaves, n, tv, tf, rays, R, hit, ht, None)
        Kt = max(3, K // 4)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(Kt):
            ca.trace_closest(ctx, inodes, leaves, n, tv, tf, rays, R, hit, ht, None)
        e1.record(stream)
        torch.cuda.synchronize()
        tms = e0.elapsed_time(e1) / Kt
        flops = TRACE_FLOPS_V * V + TRACE_FLOPS_T * T
        byts = TRACE_BYTES_V * V + TRACE_BYTES_T * T
        ai = flops / byts
        roof = min(FP32_PEAK_TFLOPS, ai * HBM_PEAK_GBS / 1e3)
        achieved_tf = flops / (tms * 1e-3) / 1e12
        trace = {"rays": R, "ms": round(tms, 3), "Mrays_s": round(R / (tms * 1e-3) / 1e6, 1),
                 "hit_fraction": round(hits / R, 4), "visits_per_ray": round(V / R, 2),
                 "tests_per_ray": round(T / R, 2), "TFLOP_s": round(achieved_tf, 3),
                 "roof_TFLOP_s": round(roof, 2), "frac": round(achieved_tf / roof, 4)}
        del rays, hit, ht, st, tv, tf
        log(f"trace: {trace}")

    # CPU baseline: the oracle's serial build (1 thread) on the same soup.
    cpu = None
    if not args.no_cpu:
        import oracle as orc

        m = min(n, 1 << 24)
        verts, faces = ca.soup_to_mesh(soup[:m])
        log(f"cpu baseline on {m} triangles")
        c0 = time.perf_counter()
        orc.lbvh_build(verts, faces, bits)
        cdt = time.perf_counter() - c0
        cpu = {"value": round(m / cdt / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
               "sample": f"oracle LBVH build (serial C++) of the same 2^{m.bit_length() - 1} "
                         f"triangle soup, {cdt:.1f} s, host cores available: {os.cpu_count()}"}

    line = {
        "metric": "LBVH Mtri/s; path-trace Msamples/s wavefront vs megakernel; adjoint/fwd time",
        "value": round(mtri_s, 2),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "LBVH build of a 2^24-triangle Gaussian-cluster soup "
                               "(64 clusters, sigma 0.02, edges 1e-3..1e-2, seed 1), "
                               "30-bit Morton (10 bits/axis), 48-B node output",
                   "tris": n, "bits_per_dim": bits, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stage_info,
        "build_effective_GB_s": round(whole_gbs, 1),
        "trace": trace,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
