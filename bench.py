#!/usr/bin/env python3
"""Benchmark of the CLeaR hot path on one MI355X.

Headline ("value", one step = one LBVH build): 2^24 triangles of a
Gaussian-cluster soup (64 clusters, sigma 0.02, edge lengths 1e-3..1e-2,
seed 1), 30-bit Morton codes (10 bits per axis), 48-B node output; inputs
resident in HBM; W untimed builds, then K timed builds bracketed by device
synchronisation; value = Mtri/s. Per-stage times come from HIP events the
library records on the build's stream; the roofline is that of the dominant
kernel (k_hier_local) with the algorithmic bytes per triangle stated below.

The roofline also reports the whole build's fraction of HBM peak on the
build's algorithmic 324 B per triangle; its measured traffic comes from a
committed PMC profile only when that profile was taken on the current kernel
source (its sha256 is recorded with it).

Secondary legs of the same metric family, each on its stated workload:
  dropin   clr_lbvh_build (the LBVHBuilder::perform drop-in) from 16-B
           vertices + 56-B Face records of the same 2^24 triangles; its nodes
           must equal the soup build's;
  trace    2^22 random rays (origins in the unit cube, directions uniform,
           tmax = inf) against the 2^24 BVH: Mrays/s, reference-order walk,
           ordered walk and its 4-wide form (the renderers' walk);
  render   the 3,117,058-triangle outdoor scene (fBm terrain 1024^2, 200
           spheres, coat BSDF, normal map, 1024x512 HDR sky, thin lens 0.02) at
           1920x1080, 64 spp, max depth 8: Msamples/s of the plain megakernel
           (reference traversal), the megakernel (ordered traversal) and the
           wavefront renderer (autotuned); the timed wavefront image must equal
           the megakernel's bit for bit; the roofline uses the oracle's
           traversal counts per camera sample (every 16th row, 2 spp);
  adjoint  the 200,000-triangle, 64-material differentiation scene, 512x512,
           32 spp, depth 5, Russian roulette off: forward and adjoint time;
  batch    256 cluster scenes: one segmented build vs 256 builds, batched
           (wavefront, and megakernel waves sharing one tile counter; the two
           images must be equal bit for bit) vs scene-by-scene rendering;
  cpu      the CPU oracle on the host: LBVH build of the same soup on 1 thread
           (cpu_baseline) and on all of the GPU's host threads, the reference
           closest-hit walk on 1 thread and on all threads, and the renderer
           with the reference walk on 1 thread and on all threads.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-trace]
                    [--no-render] [--no-adjoint] [--no-batch] [--no-dropin]
                    [--wavefront-config key=value,...]
                    [--batch-wavefront-config key=value,...]
                    [--batch-per-scene N] [--no-batch-build-compare]
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
# Per kernel: bytes per triangle it reads and writes by design. k_bounds reads
# the soup (36) and writes the 12-B box centre (the hierarchy pass gathers the
# triangle itself for the leaf box); k_morton reads the centre and writes the
# key (the hierarchy pass writes the arrival words of the splits it leaves
# open); the radix sort moves 4-B keys and payloads in 4 passes of 8-bit
# digits (64 B); the hierarchy is the Karras topology (20 B) + refit (164 B),
# all attributed to the block-local kernel, which builds every node except
# the few above block boundaries.
STAGES = [("bounds", "k_bounds", 48.0, 1), ("morton", "k_morton", 16.0, 1),
          ("sort", "k_onesweep", 64.0, 4), ("hier_local", "k_hier_local", 184.0, 1),
          ("hier_global", "k_hier_global", 0.0, 1)]
# Algorithmic bytes per triangle of the whole build (bounds 36, Morton 40,
# sort 64, Karras 20, refit 164: the build's contract model, 8-bit digits).
BUILD_BYTES_PER_TRI = 324.0


def lbvh_model(bits: int):
    """(stages, whole-build bytes per triangle) for `bits` per axis: the model
    above, with the sort's ceil(3 bits / 8) passes of 8-bit digits. Up to 10
    bits the keys are 32-bit (16 B per pass: 4-B keys and payloads, read and
    written). Above, keys are 64-bit (24 B per pass), and the Morton pass
    writes, and the hierarchy reads, 8-B keys (4 B more each)."""
    passes = (3 * bits + 7) // 8
    wide = bits > 10
    per_pass = 24.0 if wide else 16.0
    extra = 4.0 if wide else 0.0
    stages = [("bounds", "k_bounds", 48.0, 1), ("morton", "k_morton", 16.0 + extra, 1),
              ("sort", "k_onesweep", per_pass * passes, passes),
              ("hier_local", "k_hier_local", 184.0 + extra, 1), ("hier_global", "k_hier_global", 0.0, 1)]
    return stages, BUILD_BYTES_PER_TRI - 64.0 + per_pass * passes + 2.0 * extra
# Traversal work model per internal-node visit (V) and triangle test (T), and
# per visit of a 4-wide record (W: four slab tests, 4 x 24 B of boxes and
# 4 x 4 B of references).
TRACE_FLOPS_V, TRACE_FLOPS_T = 40.0, 51.0
TRACE_BYTES_V, TRACE_BYTES_T = 56.0, 40.0
TRACE_FLOPS_W, TRACE_BYTES_W = 80.0, 112.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_source_sha() -> str:
    """sha256 of the LBVH kernels' source (the traffic profile records it)."""
    import hashlib

    return hashlib.sha256((ROOT / "clear_amd" / "csrc" / "lbvh.hip").read_bytes()).hexdigest()


def load_traffic():
    """HBM bytes per launch of k_hier_local from rocprofv3 PMC runs
    (FETCH_SIZE, WRITE_SIZE) scaled by a copy-kernel calibration, committed
    under profiles/ together with the sha256 of the kernel source it was
    measured on. None if absent or measured on other kernel source."""
    p = ROOT / "profiles" / "lbvh_16m_pmc.json"
    if not p.exists():
        return None
    try:
        tr = json.loads(p.read_text())
    except Exception:
        return None
    if tr.get("lbvh_hip_sha256") != kernel_source_sha():
        log("traffic profile was measured on other kernel source: traffic not reported")
        return None
    return tr


def lbvh_leg(ca, np, torch, args, dev):
    n, bits, K, W = args.tris, args.bits, args.steps, args.warmup
    log(f"generating {n} triangles")
    soup = ca.gen_cluster_soup(n, seed=1)
    tsoup = torch.from_numpy(soup).to(dev)
    inodes, leaves = ca.node_buffers(n, dev)
    ctx = ca.Context(0)
    builder = ca.LBVHBuilder(ctx, n, bits)
    stream = torch.cuda.current_stream()
    log(f"warmup {W}")
    for _ in range(W):
        builder.perform_soup(tsoup, n, bits, inodes, leaves, stream)
    torch.cuda.synchronize()
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
    assert ctx.device_status() == 0
    stage_info = {}
    stages, build_bpt = lbvh_model(bits)
    for (name, kernel, bpt, launches), ms in zip(stages, stages_ms):
        e = {"kernel": kernel, "ms": round(ms, 4), "launches": launches, "bytes_per_tri": bpt}
        if bpt:
            gbs = bpt * n / (ms * 1e-3) / 1e9
            e.update({"GB_s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)})
        stage_info[name] = e
    loc = stage_info["hier_local"]
    loc_bpt = loc["bytes_per_tri"]
    achieved = loc_bpt * n / (loc["ms"] * 1e-3) / 1e9
    tr = load_traffic()
    # The committed counter profile is of the 10-bit build.
    traffic = tr.get("hbm_bytes_per_launch") if tr and tr.get("tris") == n and bits == 10 else None
    roofline = {"bound": "hbm", "kernel": "k_hier_local", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes_per_launch": loc_bpt * n,
                "ms_per_launch": loc["ms"]}
    whole = build_bpt * n / (dt / K) / 1e9
    roofline["whole_build"] = {"achieved": round(whole, 1), "frac": round(whole / HBM_PEAK_GBS, 4),
                               "algorithmic_bytes_per_tri": build_bpt}
    return ctx, soup, tsoup, inodes, leaves, {
        "mtri_s": n * K / dt / 1e6, "ms_per_step": dt / K * 1e3, "stages": stage_info,
        "roofline": roofline, "build_effective_GB_s": round(whole, 1)}


def dropin_leg(ca, np, torch, ctx, soup, ref_inodes, ref_leaves, args, dev):
    """clr_lbvh_build, the drop-in for LBVHBuilder::perform, from 16-B
    vertices + 56-B Face records of the same soup (vertex 3i + k = corner k
    of triangle i): its bounds pass gathers each face's three vertex indices
    and vertices (56 + 48 B) where the soup build reads 36 B. Its nodes must
    equal the soup build's."""
    n, bits, K, W = args.tris, args.bits, args.steps, args.warmup
    verts, faces = ca.soup_to_mesh(soup)
    tv = torch.from_numpy(verts).to(dev)
    tf = torch.from_numpy(faces.view(np.uint8)).to(dev)
    del verts, faces
    inodes, leaves = ca.node_buffers(n, dev)
    builder = ca.LBVHBuilder(ctx, n, bits)
    for _ in range(W):
        builder.perform(tv, tf, n, bits, inodes, leaves)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(K):
        builder.perform(tv, tf, n, bits, inodes, leaves)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / K
    same = torch.equal(inodes, ref_inodes) and torch.equal(leaves, ref_leaves)
    assert same, "the drop-in build's nodes differ from the soup build's"
    assert ctx.device_status() == 0
    bpt = lbvh_model(bits)[1] - 36.0 + 104.0
    out = {"workload": f"clr_lbvh_build of the same {n} triangles as clr_float3 vertices + Face records",
           "ms": round(ms, 4), "Mtri_s": round(n / (ms * 1e-3) / 1e6, 1), "algorithmic_bytes_per_tri": bpt,
           "frac_hbm": round(bpt * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "nodes_equal_soup_build": same}
    log(f"drop-in build: {out}")
    return out


def trace_leg(ca, np, torch, ctx, soup, inodes, leaves, args, dev):
    n, R = args.tris, args.rays
    rays = torch.from_numpy(ca.gen_random_rays(R, seed=2)).to(dev)
    hit = torch.empty(R, dtype=torch.int32, device=dev)
    ht = torch.empty(R, dtype=torch.float32, device=dev)
    st = torch.empty(2 * R, dtype=torch.int32, device=dev)
    verts, faces = ca.soup_to_mesh(soup)
    tv = torch.from_numpy(verts).to(dev)
    tf = torch.from_numpy(faces.view(np.uint8)).to(dev)
    out = {"rays": R, "workload": f"{R} random rays vs the {n}-triangle BVH"}
    for name, ordered, wide in [("reference_walk", False, False), ("ordered_walk", True, False),
                                ("wide_walk", True, True)]:
        fv, bv = (TRACE_FLOPS_W, TRACE_BYTES_W) if wide else (TRACE_FLOPS_V, TRACE_BYTES_V)
        ca.trace_closest(ctx, inodes, leaves, n, tv, tf, rays, R, hit, ht, st, ordered=ordered, wide=wide)
        torch.cuda.synchronize()
        stv = st.view(R, 2).to(torch.int64)
        V, T = int(stv[:, 0].sum()), int(stv[:, 1].sum())
        hits = int((hit != -1).sum())
        ca.trace_closest(ctx, inodes, leaves, n, tv, tf, rays, R, hit, ht, None, ordered=ordered, wide=wide)
        reps = 3
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ca.trace_closest(ctx, inodes, leaves, n, tv, tf, rays, R, hit, ht, None, ordered=ordered, wide=wide)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        flops = fv * V + TRACE_FLOPS_T * T
        byts = bv * V + TRACE_BYTES_T * T
        roof = min(FP32_PEAK_TFLOPS, flops / byts * HBM_PEAK_GBS / 1e3)
        tf_s = flops / (ms * 1e-3) / 1e12
        out[name] = {"ms": round(ms, 3), "Mrays_s": round(R / (ms * 1e-3) / 1e6, 1),
                     "hit_fraction": round(hits / R, 4), "visits_per_ray": round(V / R, 2),
                     "tests_per_ray": round(T / R, 2), "TFLOP_s": round(tf_s, 3),
                     "roof_TFLOP_s": round(roof, 2), "frac": round(tf_s / roof, 4)}
        log(f"trace {name}: {out[name]}")
    assert ctx.device_status() == 0
    return out


def fixed_wavefront_config(ca, ctx, spec: str):
    """Sets the wavefront configuration from "key=value,..." and returns it."""
    cfg = ca.wavefront_config(ctx)
    for kv in spec.split(","):
        k, v = kv.split("=")
        setattr(cfg, k.strip(), int(v))
    ca.set_wavefront_config(ctx, cfg)
    return cfg


def render_leg(ca, np, torch, ctx, args):
    W, H, spp, depth = 1920, 1080, args.render_spp, 8
    log("building the outdoor scene")
    host = ca.HostScene(ca.SCENE_TERRAIN, W, H)
    ds = ca.DeviceScene(ctx, host)
    states, pixels = ca.new_image(W, H)
    out = {"workload": f"outdoor scene, {host.num_faces} triangles, {W}x{H}, {spp} spp, "
                       f"max depth {depth}, thin lens 0.02", "triangles": host.num_faces,
           "spp": spp}
    tune_t = time.perf_counter()
    # Tuned on 4-spp renders (about 30 ms each at 1080p): 1-spp timings are
    # noisy enough to settle on a different configuration from run to run.
    if args.wavefront_config:  # a fixed configuration (profiling runs), no search
        cfg = fixed_wavefront_config(ca, ctx, args.wavefront_config)
    else:
        cfg = ca.wavefront_config(ctx) if args.no_autotune else ca.wavefront_autotune(ctx, ds, 4, depth)
    out["wavefront_config"] = cfg.as_dict()
    out["autotune_s"] = round(time.perf_counter() - tune_t, 2)
    legs = [("megakernel_plain", ca.RENDER_MEGAKERNEL, ca.RENDER_REFERENCE_TRAVERSAL),
            ("megakernel", ca.RENDER_MEGAKERNEL, 0), ("wavefront", ca.RENDER_WAVEFRONT, 0)]
    images = {}
    for name, mode, flags in legs:
        # Untimed warm-up render into scratch copies; the timed render starts
        # from freshly seeded states and a zero image, so the legs' images can
        # be compared.
        ca.seed_pixel_states(ctx, W, H, states)
        ca.render(ctx, ds, 1, depth, states, pixels, mode=mode, flags=flags)
        ca.seed_pixel_states(ctx, W, H, states)
        pixels.zero_()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ca.render(ctx, ds, spp, depth, states, pixels, mode=mode, flags=flags)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        images[name] = (pixels.clone(), states.clone())
        out[name] = {"ms": round(ms, 2), "Msamples_s": round(W * H * spp / (ms * 1e-3) / 1e6, 1)}
        log(f"render {name}: {out[name]}")
    assert ctx.device_status() == 0
    # Both renderers consume every pixel's random stream in the same order and
    # add the same terms in the same order: the timed images are identical.
    same = all(torch.equal(a, b) for a, b in zip(images["megakernel"], images["wavefront"]))
    out["wavefront_equals_megakernel"] = bool(same)
    assert same, "the timed wavefront image differs from the megakernel image"
    del images
    out["wavefront_vs_megakernel"] = round(out["wavefront"]["Msamples_s"] / out["megakernel"]["Msamples_s"], 3)
    out["wavefront_vs_plain_megakernel"] = round(
        out["wavefront"]["Msamples_s"] / out["megakernel_plain"]["Msamples_s"], 3)
    return out, host


def render_roofline(render, work, W, H, spp):
    """Roofline of the render legs from the oracle's traversal counts per
    camera sample (the ordered walk's 4-wide form that both renderers use,
    all rays cast: camera, bounce and shadow rays; the wavefront renderer
    walks environment shadow rays any-hit), with the traversal work model per
    record visit and triangle test. Shading flops are not counted, so the
    fractions are lower bounds."""
    samples = W * H * spp
    per = work["samples"]
    legs = {"megakernel": (work["visits"], work["tests"]),
            "wavefront": (work["visits"] - work["env_shadow_visits"] + work["env_shadow_anyhit_visits"],
                          work["tests"] - work["env_shadow_tests"] + work["env_shadow_anyhit_tests"])}
    out = {"work_sample": work["sample"], "rays_per_sample": round(work["rays"] / per, 3)}
    for name, (v, t) in legs.items():
        V, T = v / per * samples, t / per * samples
        flops = TRACE_FLOPS_W * V + TRACE_FLOPS_T * T
        byts = TRACE_BYTES_W * V + TRACE_BYTES_T * T
        roof = min(FP32_PEAK_TFLOPS, flops / byts * HBM_PEAK_GBS / 1e3)
        tf_s = flops / (render[name]["ms"] * 1e-3) / 1e12
        out[name] = {"bound": "latency (traversal)", "visits_per_sample": round(v / per, 2),
                     "tests_per_sample": round(t / per, 2), "achieved": round(tf_s, 3), "peak": round(roof, 2),
                     "unit": "TFLOP/s", "frac": round(tf_s / roof, 4)}
    return out


def adjoint_leg(ca, np, torch, ctx):
    from clear_amd.diff import DiffRenderer

    host = ca.HostScene(ca.SCENE_ADJOINT, 512, 512)
    r = DiffRenderer(ctx, host, 32, 5, flags=ca.RENDER_NO_RR)
    states = r.seed_states()
    g = torch.ones(512, 512, 3, device="cuda")
    r.render(states)
    r.adjoint(states, g)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        r.render(states)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        r.adjoint(states, g)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, adj = (t1 - t0) / reps * 1e3, (t2 - t1) / reps * 1e3
    assert ctx.device_status() == 0
    out = {"workload": f"differentiation scene, {host.num_faces} triangles, 64 materials, 512x512, "
                       "32 spp, depth 5, no Russian roulette",
           "forward_ms": round(fwd, 2), "adjoint_ms": round(adj, 2), "adjoint_over_forward": round(adj / fwd, 3)}
    log(f"adjoint: {out}")
    return out


def batch_leg(ca, np, torch, ctx, args):
    """256 cluster-soup scenes (log-uniform 1k..500k triangles, seeds
    100..355): one segmented LBVH build against 256 single-scene builds, then
    one batched wavefront render (128x128, 32 spp, depth 5, launch shapes
    autotuned on the batch) and one batched megakernel render against
    rendering the scenes one by one."""
    n_scenes, W, H, spp, depth = args.batch_scenes, 128, 128, 32, 5
    rng = np.random.default_rng(3)
    counts = np.exp(rng.uniform(np.log(1000), np.log(500000), n_scenes)).astype(np.int64)
    log(f"building {n_scenes} cluster scenes ({int(counts.sum())} soup triangles)")
    hosts = [ca.HostScene(ca.SCENE_CLUSTERS, W, H, seed=100 + i, num_triangles=int(n))
             for i, n in enumerate(counts)]
    batch = ca.SceneBatch(ctx, hosts)
    total = int(batch.offsets[-1])
    out = {"workload": f"{n_scenes} Gaussian-cluster scenes, {int(counts.min())}..{int(counts.max())} soup "
                       f"triangles (log-uniform 1k..500k, seeds 100..{99 + n_scenes}) + light and floor quads, "
                       f"{total} triangles in all; {W}x{H}, {spp} spp, depth 5",
           "scenes": n_scenes, "triangles": total}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    seg_ms = timed(lambda: batch.build(), 5)
    b = ca.LBVHBuilder(ctx, int(counts.max()) + 4)
    tp, ip, lp = batch.tris.data_ptr(), batch.inodes.data_ptr(), batch.leaves.data_ptr()

    def per_scene_builds():
        for s in range(n_scenes):
            o, n = int(batch.offsets[s]), int(batch.offsets[s + 1] - batch.offsets[s])
            b.perform_soup(tp + 36 * o, n, 10, ip + ca.INODE_BYTES * (o - s), lp + ca.LEAF_BYTES * o)

    out["segmented_build"] = {"ms": round(seg_ms, 3), "Mtri_s": round(total / seg_ms / 1e3, 1)}
    if args.batch_build_compare:
        per_ms = timed(per_scene_builds, 3)
        out["per_scene_builds"] = {"ms": round(per_ms, 3), "Mtri_s": round(total / per_ms / 1e3, 1)}
        out["segmented_speedup"] = round(per_ms / seg_ms, 2)
    log(f"batch builds: segmented {out['segmented_build']}, one by one {out.get('per_scene_builds')}")
    assert ctx.device_status() == 0

    npix = W * H
    states = torch.empty(n_scenes * npix * 4, dtype=torch.int32, device="cuda")
    pixels = torch.zeros(n_scenes * npix * 4, dtype=torch.float32, device="cuda")

    def seed_all():
        for s in range(n_scenes):
            ca.seed_pixel_states(ctx, W, H, states[s * npix * 4:(s + 1) * npix * 4])

    saved = ca.wavefront_config(ctx)
    tune_t = time.perf_counter()
    if args.batch_wavefront_config:  # a fixed configuration (profiling runs), no search
        cfg = fixed_wavefront_config(ca, ctx, args.batch_wavefront_config)
    else:
        cfg = ca.wavefront_config(ctx) if args.no_autotune else ca.wavefront_autotune_batch(ctx, batch, 4, depth)
    out["autotune_s"] = round(time.perf_counter() - tune_t, 2)
    out["wavefront_config"] = cfg.as_dict()
    seed_all()
    ms = timed(lambda: ca.render_batch(ctx, batch, spp, depth, states, pixels), 1)
    samples = n_scenes * npix * spp
    out["batched_wavefront"] = {"ms": round(ms, 2), "Msamples_s": round(samples / ms / 1e3, 1)}
    log(f"batched render: {out['batched_wavefront']}")
    # The megakernel form of the batch (clr_render_batch_mode): persistent
    # waves take 8x8 tiles of all scenes from one counter. Same calls from the
    # same seeds, so its states and image must equal the wavefront's.
    wave_px, wave_st = pixels.clone(), states.clone()
    seed_all()
    pixels.zero_()
    ms = timed(lambda: ca.render_batch(ctx, batch, spp, depth, states, pixels, mode=ca.RENDER_MEGAKERNEL), 1)
    out["batched_megakernel"] = {"ms": round(ms, 2), "Msamples_s": round(samples / ms / 1e3, 1),
                                 "image_equals_batched_wavefront": bool(torch.equal(pixels, wave_px)
                                                                        and torch.equal(states, wave_st))}
    log(f"batched megakernel render: {out['batched_megakernel']}")
    assert out["batched_megakernel"]["image_equals_batched_wavefront"], "batched renderers disagree"
    del wave_px, wave_st
    # Scene-by-scene rendering is timed on every k-th scene (k = scenes /
    # batch_per_scene) and reported as a rate over those scenes' samples.
    step = max(1, n_scenes // max(1, args.batch_per_scene))
    picked = list(range(0, n_scenes, step))
    psamples = len(picked) * npix * spp
    out["per_scene_sample"] = f"every {step}th scene ({len(picked)} scenes)"
    for name, mode, flags in ([] if args.batch_per_scene == 0 else [
                              ("per_scene_wavefront", ca.RENDER_WAVEFRONT, 0),
                              ("per_scene_megakernel", ca.RENDER_MEGAKERNEL, 0),
                              ("per_scene_megakernel_plain", ca.RENDER_MEGAKERNEL,
                               ca.RENDER_REFERENCE_TRAVERSAL)]):
        ca.set_wavefront_config(ctx, saved)
        seed_all()

        def one_by_one():
            for s in picked:
                sl = slice(s * npix * 4, (s + 1) * npix * 4)
                ca.ctx_render_dev(ctx, batch.devs[s], W, H, spp, depth, states[sl], pixels[sl], mode, flags)

        ms = timed(one_by_one, 1)
        out[name] = {"ms": round(ms, 2), "Msamples_s": round(psamples / ms / 1e3, 1)}
        log(f"batch {name}: {out[name]}")
    ca.set_wavefront_config(ctx, cfg)
    assert ctx.device_status() == 0
    if args.batch_per_scene:
        out["batched_vs_per_scene_megakernel"] = round(
            out["batched_wavefront"]["Msamples_s"] / out["per_scene_megakernel"]["Msamples_s"], 2)
        out["batched_megakernel_vs_per_scene_megakernel"] = round(
            out["batched_megakernel"]["Msamples_s"] / out["per_scene_megakernel"]["Msamples_s"], 2)
        out["batched_vs_per_scene_plain_megakernel"] = round(
            out["batched_wavefront"]["Msamples_s"] / out["per_scene_megakernel_plain"]["Msamples_s"], 2)
    del batch, hosts
    return out


def host_threads() -> int:
    """The host threads one GPU's share of the box may use."""
    return min(16, os.cpu_count() or 1)


def cpu_lbvh(ca, soup, bits):
    import oracle as orc

    m = soup.shape[0]
    verts, faces = ca.soup_to_mesh(soup)
    log(f"cpu baseline: oracle LBVH build of {m} triangles")
    c0 = time.perf_counter()
    built = orc.lbvh_build(verts, faces, bits)
    cdt = time.perf_counter() - c0
    # The reference's own CPU build (Scene::build's midpoint BVH, timed at
    # main.cpp:235-237), here the library's host restatement on the same soup.
    log(f"cpu baseline: midpoint BVH build of {m} triangles")
    c0 = time.perf_counter()
    ca.build_midpoint_bvh(verts, faces)
    mdt = time.perf_counter() - c0
    out = {"value": round(m / cdt / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
           "sample": f"oracle LBVH build (serial C++) of the same {m}-triangle soup, "
                     f"{cdt:.1f} s; host threads available: {os.cpu_count()}",
           "midpoint_bvh_build": {"value": round(m / mdt / 1e6, 3), "unit": "Mtri/s", "cores": 1,
                                  "sample": f"clr_bvh_build_midpoint (the reference's CPU BVH::build "
                                            f"split rule) on the same soup, {mdt:.1f} s"}}
    par = getattr(orc, "lbvh_build_parallel", None)
    if par is not None:
        th = host_threads()
        c0 = time.perf_counter()
        par(verts, faces, bits, th)
        pdt = time.perf_counter() - c0
        out["all_cores"] = {"value": round(m / pdt / 1e6, 3), "unit": "Mtri/s", "cores": th,
                            "sample": f"oracle LBVH build on {th} threads, same soup, {pdt:.1f} s"}
    return out, built, verts, faces


def cpu_trace(np, built, verts, faces, rays_np):
    """The reference's closest-hit walk (bvh_traversal.cl:182-215, restated
    by the oracle) over the oracle's build of the same soup: a bounded sample
    of the benchmark's rays on one thread and on all the GPU's host threads
    (chunks of rays on worker threads)."""
    import concurrent.futures as cf

    import oracle as orc

    n = faces.shape[0]
    out = {}
    # About 8 s of CPU work each (0.03 and 0.5 Mrays/s measured on a 16-thread MI355X host).
    for label, threads, count in (("one_core", 1, 1 << 18), ("all_cores", host_threads(), rays_np.shape[1])):
        sample = np.ascontiguousarray(rays_np[:, :count])
        chunks = np.array_split(np.arange(count), max(1, threads * 8))
        c0 = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda idx: orc.trace_closest(built["inodes"], built["leaves"], n, verts, faces,
                                                      np.ascontiguousarray(sample[:, idx])), chunks))
        dt = time.perf_counter() - c0
        out[label] = {"value": round(count / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
                      "sample": f"first {count} of the benchmark rays, reference walk: {dt:.1f} s"}
        log(f"cpu trace {label}: {out[label]}")
    return out


def cpu_render(ca, np, host, depth):
    """The render restatement with the reference's walk (the CPU Equivalent
    path's semantics; the reference loops over pixels on one thread,
    main.cpp:416-443) on one thread and on all the GPU's host threads, each on
    a band of the outdoor scene's middle rows; and the traversal work of the
    ordered walk per camera sample on every 16th row (2 samples per pixel),
    for the render roofline."""
    import oracle as orc

    W, H = host.width, host.height
    o = orc.lbvh_build(host.vertices, host.faces, 10)
    out = {}
    # About 8 s of CPU work each (0.09 and 1.4 Msamples/s measured on a 16-thread MI355X host).
    for label, threads, rows, spp in (("one_core", 1, 10, 40), ("all_cores", host_threads(), 96, 64)):
        st = orc.seed_states(W, H)
        px = np.zeros(W * H * 4, np.float32)
        b = (H // 2) * W
        c0 = time.perf_counter()
        orc.render(host.arrays, host.num_faces, o["inodes"], o["leaves"], W, H, spp, depth, st, px,
                   pix_begin=b, pix_end=b + rows * W, threads=threads, flags=ca.RENDER_REFERENCE_TRAVERSAL)
        dt = time.perf_counter() - c0
        out[label] = {"value": round(rows * W * spp / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads,
                      "kind": "port", "sample": f"render restatement, reference walk, {rows} middle rows of the "
                                                f"{W}x{H} outdoor scene, {spp} spp, depth {depth}: {dt:.1f} s"}
        log(f"cpu render {label}: {out[label]}")
    st = orc.seed_states(W, H)
    px = np.zeros(W * H * 4, np.float32)
    orc.work_counters(reset=True, count_any_hit=True)
    rows = list(range(8, H, 16))
    for y in rows:
        orc.render(host.arrays, host.num_faces, o["inodes"], o["leaves"], W, H, 2, depth, st, px,
                   pix_begin=y * W, pix_end=(y + 1) * W, threads=host_threads())
    work = orc.work_counters(reset=True, count_any_hit=False)
    work["samples"] = len(rows) * W * 2
    work["sample"] = (f"oracle renders (ordered walk, 4-wide form) of every 16th row ({len(rows)} rows x {W}), "
                      f"2 spp, depth {depth}")
    return out, work


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1 << 24)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--bits", type=int, default=10)
    ap.add_argument("--render-spp", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-trace", action="store_true", help="skip the closest-hit leg")
    ap.add_argument("--no-render", action="store_true", help="skip the render leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (Face + vertices) build leg")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--wavefront-config", default=None,
                    help="fixed wavefront configuration as key=value pairs (e.g. traceWavesPerCU=24,ldsDepth=16), "
                         "used instead of the autotuner")
    ap.add_argument("--batch-wavefront-config", default=None,
                    help="fixed wavefront configuration of the batch leg, as --wavefront-config")
    ap.add_argument("--no-autotune", action="store_true",
                    help="render legs use the default wavefront configuration (for kernel profiles)")
    ap.add_argument("--batch-per-scene", type=int, default=32,
                    help="scenes rendered one by one for the batch comparison (0: skip)")
    ap.add_argument("--no-batch-build-compare", dest="batch_build_compare", action="store_false",
                    help="skip timing the 256 one-by-one builds of the batch leg (counter runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one plain device-to-device copy of this size (calibrates "
                         "HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding "
              f"of this path is a design, not a measured mode.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import clear_amd as ca

    dev = torch.device("cuda:0")
    ctx, soup, tsoup, inodes, leaves, lb = lbvh_leg(ca, np, torch, args, dev)
    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    dropin = None if args.no_dropin else dropin_leg(ca, np, torch, ctx, soup, inodes, leaves, args, dev)
    trace = None if args.no_trace else trace_leg(ca, np, torch, ctx, soup, inodes, leaves, args, dev)
    del tsoup
    render, host = (None, None) if args.no_render else render_leg(ca, np, torch, ctx, args)
    adjoint = None if args.no_adjoint else adjoint_leg(ca, np, torch, ctx)
    batch = None if args.no_batch else batch_leg(ca, np, torch, ctx, args)
    cpu = cpu_r = cpu_t = None
    if not args.no_cpu:
        cpu, built, verts, faces = cpu_lbvh(ca, soup, args.bits)
        if trace is not None:
            cpu_t = cpu_trace(np, built, verts, faces, ca.gen_random_rays(args.rays, seed=2))
        del built, verts, faces
        if host is not None:
            cpu_r, work = cpu_render(ca, np, host, 8)
            render["roofline"] = render_roofline(render, work, host.width, host.height, args.render_spp)
            log(f"render roofline: {render['roofline']}")

    line = {
        "metric": "LBVH Mtri/s; path-trace Msamples/s wavefront vs megakernel; adjoint/fwd time",
        "value": round(lb["mtri_s"], 2),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(lb["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"LBVH build of a {args.tris}-triangle Gaussian-cluster soup "
                               "(64 clusters, sigma 0.02, edges 1e-3..1e-2, seed 1), "
                               f"{3 * args.bits}-bit Morton ({args.bits} bits/axis), 48-B node output",
                   "tris": args.tris, "bits_per_dim": args.bits, "parallelism": "single GPU"},
        "roofline": lb["roofline"],
        "cpu_baseline": cpu,
        "stages": lb["stages"],
        "build_effective_GB_s": lb["build_effective_GB_s"],
        "dropin_build": dropin,
        "trace": trace,
        "cpu_trace_baseline": cpu_t,
        "render": render,
        "adjoint": adjoint,
        "batch": batch,
        "cpu_render_baseline": cpu_r,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
