"""Pins the CPU traversal oracle (oracle/traverse_oracle.cpp).

The LBVH traversal must return exactly what a brute-force loop of the same
triangle test over every face returns (its boxes only cull); closed-form
single-triangle cases fix the triangle test itself.
"""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc


def scene(n, seed):
    soup = ca.gen_cluster_soup(n, seed=seed, num_clusters=8, sigma=0.05, e_min=0.01, e_max=0.05)
    v, f = ca.soup_to_mesh(soup)
    return soup, v, f


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 4), (50, 5), (3000, 6)])
def test_traversal_equals_bruteforce(n, seed):
    soup, v, f = scene(n, seed)
    o = orc.lbvh_build(v, f, 10)
    rays = ca.gen_random_rays(2000, seed=seed + 100)
    h, t = orc.trace_closest(o["inodes"], o["leaves"], n, v, f, rays)
    hb, tb, ties = orc.trace_bruteforce(v, f, rays)
    np.testing.assert_array_equal(t, tb)
    untied = ties <= 1
    np.testing.assert_array_equal(h[untied], hb[untied])
    if n >= 3000:
        assert (h != 0xFFFFFFFF).sum() > 100


def test_triangle_closed_form():
    # Unit right triangle in z = 2; rays along +z.
    soup = np.float32([[0, 0, 2, 1, 0, 2, 0, 1, 2]])
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 10)
    rays = np.zeros((7, 4), np.float32)
    rays[0] = [0.25, 0.9, 0.25, 0.25]
    rays[1] = [0.25, 0.9, 0.25, 0.25]
    rays[2] = [0, 0, 0, 3]
    rays[5] = [1, 1, 1, 1]
    rays[6] = np.inf
    h, t = orc.trace_closest(o["inodes"], o["leaves"], 1, v, f, rays)
    assert h.tolist() == [0, 0xFFFFFFFF, 0, 0xFFFFFFFF]
    assert t[0] == 2.0 and t[2] == 2.0  # exact: (2 - 0) from dot/invDet with unit edges
    # finite tmax below the hit distance: no hit, t stays at tmax
    rays[6] = 1.5
    h, t = orc.trace_closest(o["inodes"], o["leaves"], 1, v, f, rays)
    assert (h == 0xFFFFFFFF).all() and (t == 1.5).all()


def test_det_epsilon_culls_tiny_triangles():
    # |det| = |e1 . (d x e2)| ~ 1e-7 for edges of 1e-3.5: below the 1e-6 cull.
    s = 3e-4
    soup = np.float32([[0, 0, 1, s, 0, 1, 0, s, 1]])
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 10)
    rays = np.zeros((7, 1), np.float32)
    rays[0], rays[1], rays[5], rays[6] = s / 4, s / 4, 1, np.inf
    h, _ = orc.trace_closest(o["inodes"], o["leaves"], 1, v, f, rays)
    assert h[0] == 0xFFFFFFFF


def test_stats_count_visits():
    soup, v, f = scene(2000, 9)
    o = orc.lbvh_build(v, f, 10)
    rays = ca.gen_random_rays(500, seed=10)
    h, t, st = orc.trace_closest(o["inodes"], o["leaves"], 2000, v, f, rays, stats=True)
    assert (st[:, 0] >= 1).any() and (st[:, 1] >= (h != 0xFFFFFFFF)).all()


# ---- ordered traversal (oracle/ordered_traversal.h) -------------------------

def axis_and_bounded_rays(n_rays, seed):
    rays = ca.gen_random_rays(n_rays, seed=seed)
    k = n_rays // 4
    rays[3:6, :k] = 0.0
    rays[3 + (np.arange(k) % 3), np.arange(k)] = np.where(np.arange(k) % 2, 1.0, -1.0)
    rays[6, k:2 * k] = np.float32(0.05)
    return rays


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 4), (3, 9), (50, 5), (3000, 6), (20000, 7)])
def test_ordered_equals_bruteforce_min_t_then_face(n, seed):
    """The ordered walk never culls a box holding an accepted intersection:
    faces and t equal the brute-force lexicographic minimum of (t, face)."""
    soup, v, f = scene(n, seed)
    o = orc.lbvh_build(v, f, 10)
    rays = axis_and_bounded_rays(4000, seed + 100)
    h, t, st = orc.trace_ordered(o["inodes"], o["leaves"], n, v, f, rays)
    hb, tb = orc.trace_bruteforce_minid(v, f, rays)
    np.testing.assert_array_equal(h, hb)
    np.testing.assert_array_equal(t.view(np.uint32), tb.view(np.uint32))
    if n > 1:
        assert (st[:, 0] > 0).any()
    # any-hit mode agrees on hit / miss
    ha, _, _ = orc.trace_ordered(o["inodes"], o["leaves"], n, v, f, rays, any_hit=True)
    np.testing.assert_array_equal(ha != 0xFFFFFFFF, h != 0xFFFFFFFF)


def test_ordered_tie_goes_to_smaller_face():
    """Two coincident triangles (same vertices, faces 0 and 1): the reference
    walk keeps the last one it tests, the ordered walk the smaller index."""
    tri = [0, 0, 1, 1, 0, 1, 0, 1, 1]
    soup = np.float32([tri, tri, [5, 5, 5, 6, 5, 5, 5, 6, 5]])
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 10)
    rays = np.zeros((7, 1), np.float32)
    rays[:, 0] = [0.2, 0.2, 0.0, 0.0, 0.0, 1.0, np.inf]
    h, t, _ = orc.trace_ordered(o["inodes"], o["leaves"], 3, v, f, rays)
    assert h[0] == 0 and t[0] == 1.0
    hr, tr = orc.trace_closest(o["inodes"], o["leaves"], 3, v, f, rays)
    assert tr[0] == 1.0 and hr[0] in (0, 1)


def test_ordered_visits_fewer_nodes_on_clusters():
    soup, v, f = scene(20000, 11)
    o = orc.lbvh_build(v, f, 10)
    rays = ca.gen_random_rays(4000, seed=12)
    _, _, st_ref = orc.trace_closest(o["inodes"], o["leaves"], 20000, v, f, rays, stats=True)
    _, _, st_ord = orc.trace_ordered(o["inodes"], o["leaves"], 20000, v, f, rays)
    assert st_ord[:, 1].sum() < st_ref[:, 1].sum()  # fewer triangle tests


# ---- 4-wide form of the ordered traversal (traverseWide) -----------------------

@pytest.mark.parametrize("n,seed", [(1, 3), (2, 4), (3, 9), (5, 8), (50, 5), (3000, 6), (20000, 7)])
def test_wide_equals_bruteforce_min_t_then_face(n, seed):
    """Testing grandchildren boxes without their parent's never culls an
    accepted intersection either: faces and t equal the brute-force
    lexicographic minimum of (t, face); any-hit agrees on hit / miss."""
    soup, v, f = scene(n, seed)
    o = orc.lbvh_build(v, f, 10)
    rays = axis_and_bounded_rays(4000, seed + 100)
    h, t, st = orc.trace_wide(o["inodes"], o["leaves"], n, v, f, rays)
    hb, tb = orc.trace_bruteforce_minid(v, f, rays)
    np.testing.assert_array_equal(h, hb)
    np.testing.assert_array_equal(t.view(np.uint32), tb.view(np.uint32))
    ha, _, _ = orc.trace_wide(o["inodes"], o["leaves"], n, v, f, rays, any_hit=True)
    np.testing.assert_array_equal(ha != 0xFFFFFFFF, h != 0xFFFFFFFF)


def test_wide_tie_goes_to_smaller_face():
    tri = [0, 0, 1, 1, 0, 1, 0, 1, 1]
    soup = np.float32([tri, tri, [5, 5, 5, 6, 5, 5, 5, 6, 5], tri])
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 10)
    rays = np.zeros((7, 1), np.float32)
    rays[:, 0] = [0.2, 0.2, 0.0, 0.0, 0.0, 1.0, np.inf]
    h, t, _ = orc.trace_wide(o["inodes"], o["leaves"], 4, v, f, rays)
    assert h[0] == 0 and t[0] == 1.0


def test_wide_visits_half_the_records():
    """One record per visit stands for two levels of the binary tree: on a
    cluster soup the wide walk makes at most 60 % of the binary walk's visits
    and the same number of triangle tests within 5 %."""
    soup, v, f = scene(20000, 11)
    o = orc.lbvh_build(v, f, 10)
    rays = ca.gen_random_rays(4000, seed=12)
    _, _, st_ord = orc.trace_ordered(o["inodes"], o["leaves"], 20000, v, f, rays)
    _, _, st_w = orc.trace_wide(o["inodes"], o["leaves"], 20000, v, f, rays)
    assert st_w[:, 0].sum() <= 0.6 * st_ord[:, 0].sum()
    assert abs(int(st_w[:, 1].sum()) - int(st_ord[:, 1].sum())) <= 0.05 * st_ord[:, 1].sum()


# ---- BVH visualizer (orc_bvh_visualizer) ---------------------------------------

@pytest.mark.parametrize("walk", ["lbvh", "midpoint", "ordered"])
def test_bvh_visualizer_counts_and_base_colours(walk):
    """pixel = base colour of the surface hit + 0.025 * box hits, where the box
    hits equal the traversal restatement's counts for the same ray (visits +
    tests of orc_trace_closest / orc_trace_bvh / orc_trace_ordered) and the base
    colour is one of the Cornell box's materials (0 on a miss)."""
    W, H = 40, 32
    host = ca.HostScene(ca.SCENE_CORNELL, W, H, cpu_bvh=True)
    o = orc.lbvh_build(host.vertices, host.faces, 10)
    nodes = host.arrays["bvh_nodes"][: (2 * host.num_faces - 1) * 48].view(ca.BVH_NODE_DTYPE)
    flags = 0 if walk == "ordered" else ca.RENDER_REFERENCE_TRAVERSAL
    px = orc.bvh_visualizer(host.arrays, host.num_faces, o["inodes"], o["leaves"], W, H, flags,
                            bvh=nodes if walk == "midpoint" else None).reshape(-1, 4)[:, :3]
    rays = orc.centre_rays(host.arrays, W, H)
    if walk == "lbvh":
        hit, _, st = orc.trace_closest(o["inodes"], o["leaves"], host.num_faces, host.vertices, host.faces, rays,
                                       stats=True)
    elif walk == "midpoint":
        hit, _, st = orc.trace_bvh(nodes, host.vertices, host.faces, rays)
    else:
        hit, _, st = orc.trace_ordered(o["inodes"], o["leaves"], host.num_faces, host.vertices, host.faces, rays)
    k = st.sum(axis=1).astype(np.float32) * np.float32(0.025)
    assert (st.sum(axis=1) > 0).all()
    base = px - k[:, None]
    palette = np.array([[0.75, 0.75, 0.75], [0.6, 0, 0], [0, 0.6, 0]], np.float32)
    d = np.abs(base[:, None, :] - palette[None]).max(axis=2).min(axis=1)
    assert (d[hit != 0xFFFFFFFF] < 1e-5).all()
    assert (np.abs(base[hit == 0xFFFFFFFF]) < 1e-6).all()
