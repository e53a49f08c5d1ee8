"""Batched rendering (clr_render_batch, clr_render_batch_mode) of many
cluster-soup scenes whose LBVHs come from one segmented build: every scene's
image must equal, bit for bit, the single-scene wavefront render (and the
megakernel) of that scene with its own clr_lbvh_build, for both batched
renderers, in graph and direct-launch modes and when path-slot
chunks straddle scene boundaries; one small scene is also checked against the
CPU oracle at the renderer's image tolerance."""
import os

import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

pytestmark = pytest.mark.gpu

W, H = 60, 44  # edge tiles in both directions
THREADS = min(16, os.cpu_count() or 1)
SIZES = [1000, 2500, 7000, 20000, 3000, 60000]


@pytest.fixture(scope="module")
def scenes():
    return [ca.HostScene(ca.SCENE_CLUSTERS, W, H, seed=100 + i, num_triangles=n) for i, n in enumerate(SIZES)]


def _single(ctx, host, spp, depth, mode):
    ds = ca.DeviceScene(ctx, host)
    st, px = ca.new_image(W, H)
    ca.seed_pixel_states(ctx, W, H, st)
    ca.render(ctx, ds, spp, depth, st, px, mode=mode)
    return st, px


def _batch(ctx, batch, spp, depth, mode=None):
    import torch

    n, w, h = len(batch), batch.width, batch.height
    st = torch.empty(n * w * h * 4, dtype=torch.int32, device="cuda")
    for s in range(n):
        ca.seed_pixel_states(ctx, w, h, st[s * w * h * 4:(s + 1) * w * h * 4])
    px = torch.zeros(n * w * h * 4, dtype=torch.float32, device="cuda")
    ca.render_batch(ctx, batch, spp, depth, st, px, mode=mode)
    return st, px


def _check_equal(ctx, scenes, batch, spp, depth, label, mode=None):
    import torch

    st, px = _batch(ctx, batch, spp, depth, mode)
    torch.cuda.synchronize()
    assert ctx.device_status() == 0
    for s, host in enumerate(scenes):
        sst, spx = _single(ctx, host, spp, depth, ca.RENDER_WAVEFRONT)
        a = px[s * W * H * 4:(s + 1) * W * H * 4]
        print(f"[batch] {label} scene {s} ({host.num_faces} faces): "
              f"max |batch - single| = {float((a - spx).abs().max()):.3g}", flush=True)
        assert torch.equal(a, spx)
        assert torch.equal(st[s * W * H * 4:(s + 1) * W * H * 4], sst)
    return px


def test_segmented_nodes_equal_per_scene_builds(gpu_ctx, scenes):
    import torch

    print("[batch] segmented nodes vs per-scene builds", flush=True)
    batch = ca.SceneBatch(gpu_ctx, scenes)
    for s, host in enumerate(scenes):
        ds = ca.DeviceScene(gpu_ctx, host)
        o = int(batch.offsets[s])
        n = host.num_faces
        li = batch.leaves[o * ca.LEAF_BYTES:(o + n) * ca.LEAF_BYTES]
        ii = batch.inodes[(o - s) * ca.INODE_BYTES:(o - s + n - 1) * ca.INODE_BYTES]
        assert torch.equal(li, ds.leaves)
        assert torch.equal(ii, ds.inodes)


def test_batch_equals_single_scene_renders(gpu_ctx, scenes):
    import torch

    print(f"[batch] {len(scenes)} scenes {W}x{H} spp=4 depth=5 vs single wavefront renders", flush=True)
    batch = ca.SceneBatch(gpu_ctx, scenes)
    px = _check_equal(gpu_ctx, scenes, batch, 4, 5, "graph")
    # and the megakernel (the wavefront equals it bit for bit)
    for s in (0, 5):
        _, mpx = _single(gpu_ctx, scenes[s], 4, 5, ca.RENDER_MEGAKERNEL)
        assert torch.equal(px[s * W * H * 4:(s + 1) * W * H * 4], mpx)


def test_batch_with_large_scenes_equals_single_scene_renders(gpu_ctx):
    """Scenes of the batch workload's largest sizes (up to 500k triangles)
    render in a batch exactly as they render alone."""
    big = [ca.HostScene(ca.SCENE_CLUSTERS, W, H, seed=300 + i, num_triangles=n)
           for i, n in enumerate([450000, 1200, 320000])]
    print(f"[batch] scenes of {[h.num_faces for h in big]} faces {W}x{H} spp=2 depth=5", flush=True)
    batch = ca.SceneBatch(gpu_ctx, big)
    _check_equal(gpu_ctx, big, batch, 2, 5, "large")
    _check_equal(gpu_ctx, big, batch, 2, 5, "large, megakernel", mode=ca.RENDER_MEGAKERNEL)


def test_megakernel_batch_equals_single_scene_renders(gpu_ctx, scenes):
    """clr_render_batch_mode(MEGAKERNEL) on a batch of at most four tiles
    per CU (the instantiation with the whole stack in LDS)."""
    print(f"[batch] {len(scenes)} scenes {W}x{H} spp=3 depth=5, batched megakernel", flush=True)
    batch = ca.SceneBatch(gpu_ctx, scenes)
    _check_equal(gpu_ctx, scenes, batch, 3, 5, "megakernel", mode=ca.RENDER_MEGAKERNEL)


def test_megakernel_batch_of_many_tiles_equals_wavefront_batch(gpu_ctx):
    """A batch of more than four tiles per CU (the persistent instantiation
    with the split stack, 24 scenes x 256 tiles): both batched renderers and
    single-scene megakernel renders give the same images and states."""
    import torch

    sizes = [1000 + 1700 * i for i in range(24)]
    hosts = [ca.HostScene(ca.SCENE_CLUSTERS, 128, 128, seed=500 + i, num_triangles=n)
             for i, n in enumerate(sizes)]
    print(f"[batch] {len(hosts)} scenes 128x128 spp=2 depth=5, batched megakernel vs wavefront", flush=True)
    batch = ca.SceneBatch(gpu_ctx, hosts)
    mst, mpx = _batch(gpu_ctx, batch, 2, 5, ca.RENDER_MEGAKERNEL)
    wst, wpx = _batch(gpu_ctx, batch, 2, 5, ca.RENDER_WAVEFRONT)
    torch.cuda.synchronize()
    assert gpu_ctx.device_status() == 0
    print(f"[batch] max |megakernel - wavefront| = {float((mpx - wpx).abs().max()):.3g}, "
          f"mean {float(mpx.mean()):.4g}", flush=True)
    assert float(mpx.mean()) > 0.0
    assert torch.equal(mpx, wpx)
    assert torch.equal(mst, wst)
    npix = 128 * 128
    for s in (0, 23):
        ds = ca.DeviceScene(gpu_ctx, hosts[s])
        st, px = ca.new_image(128, 128)
        ca.seed_pixel_states(gpu_ctx, 128, 128, st)
        ca.render(gpu_ctx, ds, 2, 5, st, px, mode=ca.RENDER_MEGAKERNEL)
        assert torch.equal(mpx[s * npix * 4:(s + 1) * npix * 4], px)
        assert torch.equal(mst[s * npix * 4:(s + 1) * npix * 4], st)


def test_batch_chunks_straddle_scenes_and_direct_launches(gpu_ctx, scenes):
    saved = ca.wavefront_config(gpu_ctx)
    try:
        cfg = ca.wavefront_config(gpu_ctx)
        cfg.maxSlots = 5000  # not a multiple of the per-scene positions (64 * 8 * 6 tiles)
        cfg.useGraph = 0
        cfg.checkEvery = 1
        cfg.raysPerLane = 2  # two walkers per lane, rebinding scenes independently
        ca.set_wavefront_config(gpu_ctx, cfg)
        print("[batch] chunks of 5000 slots, direct launches", flush=True)
        batch = ca.SceneBatch(gpu_ctx, scenes[:4])
        _check_equal(gpu_ctx, scenes[:4], batch, 3, 5, "chunked")
    finally:
        ca.set_wavefront_config(gpu_ctx, saved)


def test_batch_small_scene_against_oracle(gpu_ctx, scenes):
    """Bar of the render suite: |gpu - oracle| <= 1e-3 |oracle| + 1e-6 spp per
    channel on >= 99.9 % of pixels, identical final states on >= 99.9 %."""
    host = scenes[1]
    spp, depth = 8, 5
    print(f"[batch] scene of {host.num_faces} faces vs oracle spp={spp}", flush=True)
    batch = ca.SceneBatch(gpu_ctx, [scenes[0], host])
    st, px = _batch(gpu_ctx, batch, spp, depth)
    gi = px[W * H * 4:].cpu().numpy().reshape(-1, 4)[:, :3]
    gst = st[W * H * 4:].cpu().numpy().view(np.uint32).reshape(-1, 4)
    o = orc.lbvh_build(host.vertices, host.faces, 10)
    ost = orc.seed_states(W, H)
    opx = np.zeros(W * H * 4, np.float32)
    orc.render(host.arrays, host.num_faces, o["inodes"], o["leaves"], W, H, spp, depth, ost, opx,
               threads=THREADS)
    oi = opx.reshape(-1, 4)[:, :3]
    ok = (np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-6 * spp).all(axis=1)
    same = (gst == ost.reshape(-1, 4)).all(axis=1)
    print(f"[batch] pixels within tol {ok.mean():.5f}, identical states {same.mean():.5f}, "
          f"mean {oi.mean():.5f}", flush=True)
    assert oi.mean() > 1e-4
    assert ok.mean() >= 0.999
    assert same.mean() >= 0.999


def test_batch_rejects_reference_traversal_and_mixed_sizes(gpu_ctx, scenes):
    batch = ca.SceneBatch(gpu_ctx, scenes[:2])
    import torch

    st = torch.zeros(2 * W * H * 4, dtype=torch.int32, device="cuda")
    px = torch.zeros(2 * W * H * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(ca.ClrError):
        ca.render_batch(gpu_ctx, batch, 1, 5, st, px, flags=ca.RENDER_REFERENCE_TRAVERSAL)
    with pytest.raises(ca.ClrError):
        ca.render_batch(gpu_ctx, batch, 1, 5, st, px, flags=ca.RENDER_REFERENCE_TRAVERSAL,
                        mode=ca.RENDER_MEGAKERNEL)
    with pytest.raises(ca.ClrError):
        ca.render_batch(gpu_ctx, batch, 1, 5, st, px, mode=7)
    other = ca.HostScene(ca.SCENE_CLUSTERS, 32, 32, seed=7, num_triangles=1000)
    with pytest.raises(ValueError):
        ca.SceneBatch(gpu_ctx, [scenes[0], other])


def test_batch_autotune_keeps_images(gpu_ctx, scenes):
    saved = ca.wavefront_config(gpu_ctx)
    try:
        batch = ca.SceneBatch(gpu_ctx, scenes[:3])
        cfg = ca.wavefront_autotune_batch(gpu_ctx, batch, 1, 5)
        print(f"[batch] autotuned: {cfg.as_dict()}", flush=True)
        _check_equal(gpu_ctx, scenes[:3], batch, 2, 5, "autotuned")
    finally:
        ca.set_wavefront_config(gpu_ctx, saved)
