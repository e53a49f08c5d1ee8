"""Parity at 20 bits per axis, the reference main's setting (main.cpp:365),
where keys have 60 bits and trees can be deeper than the 64-entry traversal
stack (bvh_traversal.cl:186).

* A cluster soup built at 20 bits: nodes byte-identical to the oracle's,
  closest hits of both walks bit-equal, and Cornell-box renders (megakernel
  and wavefront) within the render bar of the oracle's on 20-bit trees.
* A scene built to overflow the reference walk's stack: point pairs whose
  keys each clear one of the 60 key bits (so every level of the tree has two
  internal children, the pair on the left), then 4097 points at the all-ones
  key (12 more levels). A ray down the diagonal enters every node of the
  right spine, so the reference walk (which pushes both internal children of
  a node whose box passes, the right one on top) holds more than 64 pending
  nodes. Both the GPU walk and the oracle drop the push that does not fit;
  the GPU raises device status bit 2, and its visit counts equal the
  oracle's. The ordered walk tests children's boxes before pushing and stays
  within its stack.
"""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc
from test_gpu_render import THREADS
from test_gpu_traverse import run_gpu

pytestmark = pytest.mark.gpu

STATUS_STACK_OVERFLOW = 4


def test_cluster_soup_build_and_trace_at_20_bits(gpu_ctx):
    import torch

    n = 1 << 18
    soup = ca.gen_cluster_soup(n, seed=9)
    print(f"[bits 20] cluster soup n={n}", flush=True)
    inodes, leaves = ca.node_buffers(n)
    ca.LBVHBuilder(gpu_ctx, n, 20).perform_soup(torch.from_numpy(soup).cuda(), n, 20, inodes, leaves)
    torch.cuda.synchronize()
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 20)
    np.testing.assert_array_equal(leaves.cpu().numpy().reshape(-1, 48), o["leaves"])
    np.testing.assert_array_equal(inodes.cpu().numpy().reshape(-1, 48)[: n - 1], o["inodes"])
    rays = ca.gen_random_rays(1 << 16, seed=3)
    for ordered in (False, True):
        gh, gt, gst = run_gpu(gpu_ctx, o["inodes"], o["leaves"], n, v, f, rays, ordered=ordered)
        if ordered:
            h, t, st = orc.trace_ordered(o["inodes"], o["leaves"], n, v, f, rays)
        else:
            h, t, st = orc.trace_closest(o["inodes"], o["leaves"], n, v, f, rays, stats=True)
        np.testing.assert_array_equal(gh, h)
        np.testing.assert_array_equal(gt.view(np.uint32), t.view(np.uint32))
        np.testing.assert_array_equal(gst, st)
    assert gpu_ctx.device_status() == 0


@pytest.mark.parametrize("mode", [ca.RENDER_MEGAKERNEL, ca.RENDER_WAVEFRONT])
def test_cornell_render_at_20_bits(gpu_ctx, mode):
    import torch

    W = H = 64
    spp, depth = 8, 5
    host = ca.HostScene(ca.SCENE_CORNELL, W, H, detail=4)
    print(f"[bits 20] cornell {W}x{H} spp={spp} mode={mode}", flush=True)
    ds = ca.DeviceScene(gpu_ctx, host, bits=20)
    states, pixels = ca.new_image(W, H)
    ca.seed_pixel_states(gpu_ctx, W, H, states)
    ca.render(gpu_ctx, ds, spp, depth, states, pixels, mode=mode)
    torch.cuda.synchronize()
    assert gpu_ctx.device_status() == 0
    o = orc.lbvh_build(host.vertices, host.faces, 20)
    np.testing.assert_array_equal(ds.leaves.cpu().numpy().reshape(-1, 48), o["leaves"])
    st = orc.seed_states(W, H)
    px = np.zeros(W * H * 4, np.float32)
    orc.render(host.arrays, host.num_faces, o["inodes"], o["leaves"], W, H, spp, depth, st, px, threads=THREADS)
    g = pixels.cpu().numpy().reshape(-1, 4)[:, :3]
    c = px.reshape(-1, 4)[:, :3]
    ok = (np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-6 * spp).all(axis=1)
    assert c.mean() > 1e-4 and ok.mean() >= 0.999


def deep_comb_soup():
    """Degenerate point triangles in [0, 1]^3 whose quantised 20-bit
    coordinates are q (cell centres (q + 0.5) / 2^20): the origin, (1, 1, 1)
    (which fixes the scene box), a pair at each (all ones except bit m of
    axis a), and 4096 more points in the all-ones cell."""
    top = (1 << 20) - 1
    cells = [(0, 0, 0)]
    for a in range(3):
        for m in range(20):
            q = [top, top, top]
            q[a] = top - (1 << m)
            cells += [tuple(q), tuple(q)]
    cells += [(top, top, top)] * 4096
    pts = (np.array(cells, np.float64) + 0.5) / float(1 << 20)
    pts = np.vstack([pts, [[1.0, 1.0, 1.0]]]).astype(np.float32)
    return np.repeat(pts, 3, axis=0).reshape(-1, 9)


def test_reference_walk_stack_overflow_is_flagged_and_matches_oracle(gpu_ctx):
    import torch

    soup = deep_comb_soup()
    n = soup.shape[0]
    print(f"[bits 20] deep comb n={n}", flush=True)
    inodes, leaves = ca.node_buffers(n)
    ca.LBVHBuilder(gpu_ctx, n, 20).perform_soup(torch.from_numpy(soup).cuda(), n, 20, inodes, leaves)
    torch.cuda.synchronize()
    assert gpu_ctx.device_status() == 0
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 20)
    np.testing.assert_array_equal(inodes.cpu().numpy().reshape(-1, 48)[: n - 1], o["inodes"])
    # Rays down the diagonal towards the origin, from just outside the corner.
    R = 256
    rays = np.zeros((7, R), np.float32)
    rays[0:3] = 1.5
    rng = np.random.default_rng(1)
    d = -np.ones((3, R)) + rng.normal(0.0, 1e-9, (3, R))
    rays[3:6] = d / np.linalg.norm(d, axis=0)
    rays[6] = np.inf
    gh, gt, gst = run_gpu(gpu_ctx, o["inodes"], o["leaves"], n, v, f, rays, ordered=False)
    assert gpu_ctx.device_status() & STATUS_STACK_OVERFLOW, "the reference walk did not report the overflow"
    orc.trace_overflows(reset=True)
    h, t, st = orc.trace_closest(o["inodes"], o["leaves"], n, v, f, rays, stats=True)
    assert orc.trace_overflows(reset=True) == R  # the oracle's walk dropped pushes on every ray too
    np.testing.assert_array_equal(gh, h)
    np.testing.assert_array_equal(gst, st)
    assert (gst[:, 0] > 64).all()  # the rays walk the whole spine
    gh, gt, gst = run_gpu(gpu_ctx, o["inodes"], o["leaves"], n, v, f, rays, ordered=True)
    h, t, st = orc.trace_ordered(o["inodes"], o["leaves"], n, v, f, rays)
    np.testing.assert_array_equal(gh, h)
    np.testing.assert_array_equal(gst, st)
    assert gpu_ctx.device_status() == 0
