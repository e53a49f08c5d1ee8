"""Wavefront renderer parity: same image and final RNG states as the
megakernel (bit for bit: per pixel the random draws and the order of radiance
additions are the same), and the oracle bar of test_gpu_render.py, across
graph / direct launches, chunked path pools, early-stop settings and the
autotuner."""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc
from test_gpu_render import compare, oracle_render

pytestmark = pytest.mark.gpu


def run(ctx, ds, spp, depth, mode, calls=1, flags=0):
    import torch

    states, pixels = ca.new_image(ds.width, ds.height)
    ca.seed_pixel_states(ctx, ds.width, ds.height, states)
    for _ in range(calls):
        ca.render(ctx, ds, spp, depth, states, pixels, mode=mode, flags=flags)
    torch.cuda.synchronize()
    assert ctx.device_status() == 0
    return pixels.cpu().numpy(), states.cpu().numpy().view(np.uint32)


def same_as_megakernel(ctx, host, spp, depth, calls=1, cfg=None, flags=0):
    ds = ca.DeviceScene(ctx, host)
    if cfg is not None:
        ca.set_wavefront_config(ctx, cfg)
    wp, ws = run(ctx, ds, spp, depth, ca.RENDER_WAVEFRONT, calls, flags)
    mp, ms = run(ctx, ds, spp, depth, ca.RENDER_MEGAKERNEL, calls, flags)
    ok = (wp.view(np.uint32) == mp.view(np.uint32)).reshape(-1, 4)[:, :3].all(axis=1)
    print(f"[wavefront] {host.width}x{host.height} spp={spp}x{calls}: identical pixels {ok.mean():.6f}",
          flush=True)
    np.testing.assert_array_equal(ws, ms)
    np.testing.assert_array_equal(wp, mp)
    return wp, ws


def default_cfg(ctx, **kw):
    c = ca.wavefront_config(ctx)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_cornell_matches_megakernel_and_oracle(gpu_ctx):
    host = ca.HostScene(ca.SCENE_CORNELL, 256, 256)
    wp, ws = same_as_megakernel(gpu_ctx, host, 16, 5, cfg=default_cfg(gpu_ctx))
    o = orc.lbvh_build(host.vertices, host.faces, 10)
    compare("wavefront cornell", host, 16, 5, (wp, ws, o["inodes"], o["leaves"]),
            oracle_render(host, 16, 5))


@pytest.mark.parametrize("lens", [0.0, 0.05])
def test_showcase(gpu_ctx, lens):
    host = ca.HostScene(ca.SCENE_SHOWCASE, 128, 128, lens_radius=lens)
    same_as_megakernel(gpu_ctx, host, 8, 8, cfg=default_cfg(gpu_ctx))


def test_terrain_reduced(gpu_ctx):
    host = ca.HostScene(ca.SCENE_TERRAIN, 192, 108, detail=256)
    same_as_megakernel(gpu_ctx, host, 4, 8, cfg=default_cfg(gpu_ctx))


@pytest.mark.parametrize("flags", [0, ca.RENDER_NO_RR], ids=["roulette", "no_roulette"])
def test_terrain_many_tiles(gpu_ctx, flags):
    """More than four 8x8 tiles per CU (48 x 36 tiles), so the megakernel runs
    its persistent instantiation with the split stack: with Russian roulette
    the merged sample/bounce loop, without it the per-sample loop."""
    host = ca.HostScene(ca.SCENE_TERRAIN, 384, 288, detail=256)
    same_as_megakernel(gpu_ctx, host, 2, 8, cfg=default_cfg(gpu_ctx), flags=flags)


def test_reference_traversal_showcase(gpu_ctx):
    host = ca.HostScene(ca.SCENE_SHOWCASE, 96, 96)
    same_as_megakernel(gpu_ctx, host, 4, 8, cfg=default_cfg(gpu_ctx), flags=ca.RENDER_REFERENCE_TRAVERSAL)


def test_furnace_deep(gpu_ctx):
    host = ca.HostScene(ca.SCENE_FURNACE, 64, 64, albedo=0.8, fov_y=40.0)
    same_as_megakernel(gpu_ctx, host, 16, 64, cfg=default_cfg(gpu_ctx))


def test_chunks_direct_launches_and_calls(gpu_ctx):
    """maxSlots below the pixel count (3 chunks), plain launches, no early
    stop, three calls carrying the states."""
    host = ca.HostScene(ca.SCENE_SHOWCASE, 72, 40)
    cfg = default_cfg(gpu_ctx, maxSlots=1024, useGraph=0, checkEvery=0, itersPerGraph=4)
    same_as_megakernel(gpu_ctx, host, 2, 8, calls=3, cfg=cfg)


def test_graph_shapes(gpu_ctx):
    host = ca.HostScene(ca.SCENE_CORNELL, 61, 13, detail=2)
    for tb, sb, g, chk, ld, rf, k in [(64, 256, 2, 1, 16, 1, 1), (256, 64, 16, 0, 32, 64, 2),
                                      (128, 128, 64, 3, 16, 24, 2), (64, 64, 4, 0, 32, 8, 2)]:
        cfg = default_cfg(gpu_ctx, traceBlock=tb, shadeBlock=sb, itersPerGraph=g, checkEvery=chk,
                          traceWavesPerCU=4, shadeWavesPerCU=4, ldsDepth=ld, refillMin=rf, raysPerLane=k)
        same_as_megakernel(gpu_ctx, host, 4, 5, cfg=cfg)


@pytest.mark.parametrize("ld", [16, 32])
def test_two_rays_per_lane(gpu_ctx, ld):
    """Two walkers per lane in the extend / shadow kernels: same image as the
    megakernel, with textured alpha and both environment and area-light shadow
    rays (showcase scene), graph and direct launches."""
    host = ca.HostScene(ca.SCENE_SHOWCASE, 96, 64)
    for graph in (1, 0):
        cfg = default_cfg(gpu_ctx, raysPerLane=2, ldsDepth=ld, useGraph=graph, maxSlots=4096)
        same_as_megakernel(gpu_ctx, host, 4, 6, cfg=cfg)


def test_autotune_leaves_caller_buffers_alone(gpu_ctx):
    import torch

    host = ca.HostScene(ca.SCENE_CORNELL, 64, 64, detail=3)
    ds = ca.DeviceScene(gpu_ctx, host)
    states, pixels = ca.new_image(64, 64)
    ca.seed_pixel_states(gpu_ctx, 64, 64, states)
    before = states.clone()
    cfg = ca.wavefront_autotune(gpu_ctx, ds, 1, 5)
    torch.cuda.synchronize()
    print(f"[wavefront] autotuned: {cfg.as_dict()}", flush=True)
    assert torch.equal(states, before) and float(pixels.abs().sum()) == 0.0
    assert cfg.traceBlock in (64, 128, 256) and cfg.itersPerGraph % 2 == 0
    got = ca.wavefront_config(gpu_ctx).as_dict()
    assert got == cfg.as_dict()
    same_as_megakernel(gpu_ctx, host, 4, 5)


def test_rejects_bad_config(gpu_ctx):
    cfg = default_cfg(gpu_ctx, itersPerGraph=3)
    with pytest.raises(ca.ClrError):
        ca.set_wavefront_config(gpu_ctx, cfg)
    cfg = default_cfg(gpu_ctx, traceBlock=96)
    with pytest.raises(ca.ClrError):
        ca.set_wavefront_config(gpu_ctx, cfg)
    cfg = default_cfg(gpu_ctx, ldsDepth=24)
    with pytest.raises(ca.ClrError):
        ca.set_wavefront_config(gpu_ctx, cfg)
    cfg = default_cfg(gpu_ctx, raysPerLane=3)
    with pytest.raises(ca.ClrError):
        ca.set_wavefront_config(gpu_ctx, cfg)
