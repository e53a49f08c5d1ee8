"""GPU closest-hit parity: face IDs, hit distances and visit counts equal the
CPU oracle's bit for bit (the oracle itself is pinned to brute force)."""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

pytestmark = pytest.mark.gpu


def run_gpu(ctx, o_inodes, o_leaves, n, v, f, rays, stats=True, ordered=False, wide=False):
    import torch

    R = rays.shape[1]
    ti = torch.from_numpy(np.ascontiguousarray(o_inodes).reshape(-1) if n > 1
                          else np.zeros(48, np.uint8)).cuda()
    tl = torch.from_numpy(np.ascontiguousarray(o_leaves).reshape(-1)).cuda()
    tv = torch.from_numpy(v).cuda()
    tf = torch.from_numpy(f.view(np.uint8)).cuda()
    tr = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    hit = torch.empty(R, dtype=torch.int32, device="cuda")
    t = torch.empty(R, dtype=torch.float32, device="cuda")
    st = torch.empty(2 * R, dtype=torch.int32, device="cuda") if stats else None
    ca.trace_closest(ctx, ti, tl, n, tv, tf, tr, R, hit, t, st, ordered=ordered, wide=wide)
    torch.cuda.synchronize()
    out = (hit.cpu().numpy().view(np.uint32), t.cpu().numpy())
    if stats:
        out += (st.cpu().numpy().view(np.uint32).reshape(R, 2),)
    return out


def check(ctx, soup, rays, bits=10):
    n = soup.shape[0]
    print(f"[trace parity] n={n} rays={rays.shape[1]} bits={bits}", flush=True)
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, bits)
    h, t, st = orc.trace_closest(o["inodes"], o["leaves"], n, v, f, rays, stats=True)
    gh, gt, gst = run_gpu(ctx, o["inodes"], o["leaves"], n, v, f, rays)
    np.testing.assert_array_equal(gh, h)
    np.testing.assert_array_equal(gt.view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(gst, st)
    return h


@pytest.mark.parametrize("n", [1, 2, 3, 1000])
def test_small_scenes(gpu_ctx, n):
    soup = ca.gen_cluster_soup(n, seed=n, num_clusters=4, sigma=0.1, e_min=0.02, e_max=0.1)
    check(gpu_ctx, soup, ca.gen_random_rays(4096, seed=n + 1))


def test_axis_aligned_and_finite_tmax(gpu_ctx):
    soup = ca.gen_cluster_soup(20000, seed=3, num_clusters=8, sigma=0.05, e_min=0.005, e_max=0.03)
    rays = ca.gen_random_rays(8192, seed=4)
    rays[3:6, :2048] = 0.0
    rays[3 + (np.arange(2048) % 3), np.arange(2048)] = np.where(np.arange(2048) % 2, 1.0, -1.0)
    rays[6, 4096:] = np.float32(0.3)
    h = check(gpu_ctx, soup, rays)
    assert (h != 0xFFFFFFFF).sum() > 500


def test_cluster_64k_tris_64k_rays(gpu_ctx):
    check(gpu_ctx, ca.gen_cluster_soup(1 << 16, seed=1), ca.gen_random_rays(1 << 16, seed=2))


def test_cluster_1m_tris_256k_rays(gpu_ctx):
    h = check(gpu_ctx, ca.gen_cluster_soup(1 << 20, seed=1), ca.gen_random_rays(1 << 18, seed=2))
    assert (h != 0xFFFFFFFF).mean() > 0.2


def check_ordered(ctx, soup, rays, bits=10):
    n = soup.shape[0]
    print(f"[trace ordered parity] n={n} rays={rays.shape[1]}", flush=True)
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, bits)
    h, t, st = orc.trace_ordered(o["inodes"], o["leaves"], n, v, f, rays)
    gh, gt, gst = run_gpu(ctx, o["inodes"], o["leaves"], n, v, f, rays, ordered=True)
    np.testing.assert_array_equal(gh, h)
    np.testing.assert_array_equal(gt.view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(gst, st)
    assert ctx.device_status() == 0
    return h


@pytest.mark.parametrize("n", [1, 2, 3, 1000])
def test_ordered_small_scenes(gpu_ctx, n):
    soup = ca.gen_cluster_soup(n, seed=n, num_clusters=4, sigma=0.1, e_min=0.02, e_max=0.1)
    check_ordered(gpu_ctx, soup, ca.gen_random_rays(4096, seed=n + 1))


def test_ordered_axis_aligned_and_finite_tmax(gpu_ctx):
    soup = ca.gen_cluster_soup(20000, seed=3, num_clusters=8, sigma=0.05, e_min=0.005, e_max=0.03)
    rays = ca.gen_random_rays(8192, seed=4)
    rays[3:6, :2048] = 0.0
    rays[3 + (np.arange(2048) % 3), np.arange(2048)] = np.where(np.arange(2048) % 2, 1.0, -1.0)
    rays[6, 4096:] = np.float32(0.3)
    check_ordered(gpu_ctx, soup, rays)


def test_ordered_several_rays_per_lane(gpu_ctx):
    """2^20 rays: more rays than the trace grid has lanes (at most 2^18 on
    this device), so every lane walks two rays at once and restarts each walker
    several times; hits, distances and visit / test counts equal the oracle's."""
    soup = ca.gen_cluster_soup(1 << 16, seed=5, num_clusters=16, sigma=0.05, e_min=0.005, e_max=0.03)
    h = check_ordered(gpu_ctx, soup, ca.gen_random_rays(1 << 20, seed=6))
    assert (h != 0xFFFFFFFF).mean() > 0.1


def test_ordered_1m_tris_256k_rays_match_reference_walk(gpu_ctx):
    """Ordered and reference walks agree on every ray here (no exact ties)."""
    soup = ca.gen_cluster_soup(1 << 20, seed=1)
    rays = ca.gen_random_rays(1 << 18, seed=2)
    h = check_ordered(gpu_ctx, soup, rays)
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 10)
    gh, gt, _ = run_gpu(gpu_ctx, o["inodes"], o["leaves"], 1 << 20, v, f, rays)
    assert (gh == h).mean() > 0.99999


# ---- 4-wide ordered walk (clr_trace_closest_wide) ----------------------------

def check_wide(ctx, soup, rays, bits=10):
    """The wide walk (the renderers' traversal) against the oracle's
    restatement: hits, distances and per-ray record visits and tests; and its
    hits against the binary ordered walk's."""
    n = soup.shape[0]
    print(f"[trace wide parity] n={n} rays={rays.shape[1]} bits={bits}", flush=True)
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, bits)
    h, t, st = orc.trace_wide(o["inodes"], o["leaves"], n, v, f, rays)
    gh, gt, gst = run_gpu(ctx, o["inodes"], o["leaves"], n, v, f, rays, wide=True)
    np.testing.assert_array_equal(gh, h)
    np.testing.assert_array_equal(gt.view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(gst, st)
    oh, ot, _ = run_gpu(ctx, o["inodes"], o["leaves"], n, v, f, rays, ordered=True)
    np.testing.assert_array_equal(gh, oh)
    np.testing.assert_array_equal(gt.view(np.uint32), ot.view(np.uint32))
    assert ctx.device_status() == 0
    return h


@pytest.mark.parametrize("n", [1, 2, 3, 5, 1000])
def test_wide_small_scenes(gpu_ctx, n):
    soup = ca.gen_cluster_soup(n, seed=n, num_clusters=4, sigma=0.1, e_min=0.02, e_max=0.1)
    check_wide(gpu_ctx, soup, ca.gen_random_rays(4096, seed=n + 1))


def test_wide_axis_aligned_and_finite_tmax(gpu_ctx):
    soup = ca.gen_cluster_soup(20000, seed=3, num_clusters=8, sigma=0.05, e_min=0.005, e_max=0.03)
    rays = ca.gen_random_rays(8192, seed=4)
    rays[3:6, :2048] = 0.0
    rays[3 + (np.arange(2048) % 3), np.arange(2048)] = np.where(np.arange(2048) % 2, 1.0, -1.0)
    rays[6, 4096:] = np.float32(0.3)
    check_wide(gpu_ctx, soup, rays)


def test_wide_1m_tris_several_rays_per_lane(gpu_ctx):
    soup = ca.gen_cluster_soup(1 << 20, seed=1)
    h = check_wide(gpu_ctx, soup, ca.gen_random_rays(1 << 19, seed=2))
    assert (h != 0xFFFFFFFF).mean() > 0.1


def test_wide_20_bits_per_axis(gpu_ctx):
    soup = ca.gen_cluster_soup(1 << 16, seed=9)
    check_wide(gpu_ctx, soup, ca.gen_random_rays(1 << 15, seed=3), bits=20)


# ---- CPU midpoint BVH (clr_trace_closest_bvh) --------------------------------

def run_gpu_bvh(ctx, nodes, v, f, rays, ordered):
    import torch

    R = rays.shape[1]
    tn = torch.from_numpy(np.ascontiguousarray(nodes).view(np.uint8).reshape(-1)).cuda()
    tv = torch.from_numpy(v).cuda()
    tf = torch.from_numpy(f.view(np.uint8)).cuda()
    tr = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    hit = torch.empty(R, dtype=torch.int32, device="cuda")
    t = torch.empty(R, dtype=torch.float32, device="cuda")
    st = torch.empty(2 * R, dtype=torch.int32, device="cuda")
    ca.trace_closest_bvh(ctx, tn, len(f), tv, tf, tr, R, hit, t, st, ordered=ordered)
    torch.cuda.synchronize()
    return (hit.cpu().numpy().view(np.uint32), t.cpu().numpy(), st.cpu().numpy().view(np.uint32).reshape(R, 2))


@pytest.mark.parametrize("n", [1, 2, 3, 65, 4097, 20000])
def test_midpoint_bvh_walks(gpu_ctx, n):
    """Reference-order walk over the midpoint BVH: face, t and (visits, tests)
    equal the oracle's walk bit for bit; ordered traversal over the same nodes:
    equal to the brute-force lexicographic minimum of (t, face)."""
    soup = ca.gen_cluster_soup(n, seed=n + 30, num_clusters=8, sigma=0.05, e_min=0.005, e_max=0.03)
    v, f = ca.soup_to_mesh(soup)
    nodes = ca.build_midpoint_bvh(v, f)
    rays = ca.gen_random_rays(8192, seed=n + 31)
    rays[3:6, :2048] = 0.0
    rays[3 + (np.arange(2048) % 3), np.arange(2048)] = np.where(np.arange(2048) % 2, 1.0, -1.0)
    rays[6, 6144:] = np.float32(0.3)
    print(f"[trace midpoint] n={n} rays={rays.shape[1]}", flush=True)
    h, t, st = orc.trace_bvh(nodes, v, f, rays)
    gh, gt, gst = run_gpu_bvh(gpu_ctx, nodes, v, f, rays, ordered=False)
    np.testing.assert_array_equal(gh, h)
    np.testing.assert_array_equal(gt.view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(gst, st)
    bh, bt = orc.trace_bruteforce_minid(v, f, rays)
    oh, ot, _ = run_gpu_bvh(gpu_ctx, nodes, v, f, rays, ordered=True)
    np.testing.assert_array_equal(oh, bh)
    np.testing.assert_array_equal(ot.view(np.uint32), bt.view(np.uint32))
    assert gpu_ctx.device_status() == 0
    if n >= 4097:
        assert (oh != 0xFFFFFFFF).sum() > 200


def test_midpoint_and_lbvh_ordered_hits_agree_1m(gpu_ctx):
    """The ordered traversal returns the same (face, t) over the LBVH and the
    midpoint BVH of a 1M-triangle soup (both equal the lexicographic minimum)."""
    soup = ca.gen_cluster_soup(1 << 20, seed=1)
    v, f = ca.soup_to_mesh(soup)
    nodes = ca.build_midpoint_bvh(v, f)
    rays = ca.gen_random_rays(1 << 18, seed=2)
    o = orc.lbvh_build(v, f, 10)
    print("[trace midpoint] 1M soup, ordered over both structures", flush=True)
    lh, lt, _ = run_gpu(gpu_ctx, o["inodes"], o["leaves"], 1 << 20, v, f, rays, ordered=True)
    mh, mt, _ = run_gpu_bvh(gpu_ctx, nodes, v, f, rays, ordered=True)
    np.testing.assert_array_equal(mh, lh)
    np.testing.assert_array_equal(mt.view(np.uint32), lt.view(np.uint32))
    assert gpu_ctx.device_status() == 0
