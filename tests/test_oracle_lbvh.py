"""Pins the CPU LBVH oracle (oracle/lbvh_oracle.cpp) before it is trusted.

The reference ships no tests or golden vectors for this path, so each stage is
checked against something fixed independently of the oracle's code:
closed-form Morton codes, a stable sort of the interleaved key (C++ and
numpy), a top-down binary radix tree, per-range min/max folds, and tiny
hand-checked trees.
"""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc


def point_tris(points):
    """Degenerate triangles (3 equal corners) at the given points."""
    p = np.asarray(points, dtype=np.float32)
    return np.repeat(p, 3, axis=0).reshape(-1, 9)


def build(soup, bits):
    v, f = ca.soup_to_mesh(soup)
    return orc.lbvh_build(v, f, bits), v, f


def soup_for(seed, n, dup=False):
    rng = np.random.default_rng(seed)
    soup = (rng.random((n, 9), dtype=np.float32) * np.float32(0.01)
            + rng.random((n, 1), dtype=np.float32))
    if dup:
        soup[: n // 2] = soup[0]
    return soup.astype(np.float32)


@pytest.mark.parametrize("bits", [1, 3, 10, 20, 21])
def test_morton_closed_form(bits):
    # Scene [0,1]^3: min corner -> 0, max corner -> 2^b - 1 (clamped), centre ->
    # 2^(b-1), quarter -> 2^(b-2) (lbvh_construction.cl:178-179).
    pts = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [0.25, 0.75, 0.5]]
    o, _, _ = build(point_tris(pts), bits)
    m = o["morton"]
    top = (1 << bits) - 1
    assert m[0].tolist() == [0, 0, 0]
    assert m[1].tolist() == [top, top, top]
    assert m[2].tolist() == [1 << (bits - 1)] * 3
    assert m[3].tolist() == [int(0.25 * 2**bits), min(int(0.75 * 2**bits), top), 1 << (bits - 1)]
    np.testing.assert_array_equal(o["scene"], [0, 0, 0, 1, 1, 1])


@pytest.mark.parametrize("bits", [1, 3, 10, 20])
@pytest.mark.parametrize("seed,dup", [(1, False), (2, True), (3, False)])
def test_sort_is_stable_sort_of_interleaved_key(bits, seed, dup):
    o, _, _ = build(soup_for(seed, 3000, dup), bits)
    keys = orc.interleave(o["morton"], bits)
    np.testing.assert_array_equal(o["sorted"], orc.stable_order(keys))
    np.testing.assert_array_equal(o["sorted"], np.argsort(keys, kind="stable"))
    # independent interleave in numpy: digit x + 2y + 4z per level
    m = o["morton"].astype(np.uint64)
    k2 = np.zeros(len(m), dtype=np.uint64)
    for b in range(bits):
        for ax in range(3):
            k2 |= ((m[:, ax] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + ax)
    np.testing.assert_array_equal(keys, k2)


def check_topology_and_boxes(o, soup, bits):
    n = soup.shape[0]
    keys = orc.interleave(o["morton"], bits)[o["sorted"]]
    tree = orc.radix_tree(keys, 3 * bits)
    ino = ca.decode_inodes(o["inodes"], n)
    lv = ca.decode_leaves(o["leaves"], n)
    np.testing.assert_array_equal(lv["objIdx"], o["sorted"])
    # leaf boxes = per-triangle min/max (calcAABBs)
    tri = soup[lv["objIdx"]].reshape(n, 3, 3)
    np.testing.assert_array_equal(lv["min"], tri.min(axis=1))
    np.testing.assert_array_equal(lv["max"], tri.max(axis=1))
    if n < 2:
        return
    np.testing.assert_array_equal(ino["c"][:, 0], tree["c0"])
    np.testing.assert_array_equal(ino["c"][:, 1], tree["c1"])
    np.testing.assert_array_equal(ino["isChild"][:, 0], tree["leaf0"])
    np.testing.assert_array_equal(ino["isChild"][:, 1], tree["leaf1"])
    boxes = orc.range_boxes(np.concatenate([lv["min"], lv["max"]], axis=1), tree["lo"], tree["hi"])
    np.testing.assert_array_equal(ino["min"], boxes[:, :3])
    np.testing.assert_array_equal(ino["max"], boxes[:, 3:])


@pytest.mark.parametrize("bits", [1, 3, 10, 20])
@pytest.mark.parametrize("n", [2, 3, 5, 64, 1000, 4097])
def test_topology_and_refit(bits, n):
    soup = soup_for(n + bits, n, dup=(n % 2 == 0))
    o, _, _ = build(soup, bits)
    check_topology_and_boxes(o, soup, bits)


def test_cluster_scene_topology():
    soup = ca.gen_cluster_soup(20000, seed=1)
    o, _, _ = build(soup, 10)
    check_topology_and_boxes(o, soup, 10)


def test_single_triangle():
    soup = soup_for(7, 1)
    o, _, _ = build(soup, 10)
    assert o["inodes"].shape[0] == 0
    lv = ca.decode_leaves(o["leaves"], 1)
    assert lv["objIdx"][0] == 0


def test_two_and_three_by_hand():
    # Three points on the x axis of [0,1]: codes x = 0, 512, 1023 (bits 10).
    # Keys 0 < 512 < 1023; delta(0,1) < delta(1,2), so node 0 splits at 0
    # (leaf 0 | node 1) and node 1 covers leaves 1, 2.
    o, _, _ = build(point_tris([[0, 0, 0], [0.5, 0, 0], [1, 0, 0]]), 10)
    ino = ca.decode_inodes(o["inodes"], 3)
    assert ino["c"].tolist() == [[0, 1], [1, 2]]
    assert ino["isChild"].tolist() == [[1, 0], [1, 1]]
    # Two coincident points: equal keys, the index tie-break gives [0, 1].
    o, _, _ = build(point_tris([[0.3, 0.3, 0.3], [0.3, 0.3, 0.3]]), 10)
    ino = ca.decode_inodes(o["inodes"], 2)
    assert ino["c"].tolist() == [[0, 1]] and ino["isChild"].tolist() == [[1, 1]]
    np.testing.assert_array_equal(o["sorted"], [0, 1])


def test_all_identical_centroids_use_index_tiebreak():
    n = 1000
    soup = np.tile(np.float32([0.1, 0.2, 0.3, 0.2, 0.3, 0.4, 0.3, 0.1, 0.2]), (n, 1))
    o, _, _ = build(soup, 10)
    np.testing.assert_array_equal(o["sorted"], np.arange(n))
    assert (o["morton"] == o["morton"][0]).all()
    check_topology_and_boxes(o, soup, 10)
    # Degenerate points at one location: zero scene extent, 0/0 -> code 0.
    o, _, _ = build(point_tris([[0.7, 0.1, 0.2]] * 50), 10)
    assert (o["morton"] == 0).all()
    np.testing.assert_array_equal(o["sorted"], np.arange(50))


def test_flat_axis_gets_code_zero():
    soup = soup_for(11, 500)
    soup[:, 2::3] = np.float32(0.25)
    o, _, _ = build(soup, 10)
    assert (o["morton"][:, 2] == 0).all()
    assert o["morton"][:, 0].max() > 1000  # centroids sit inside the box-based bounds


def test_nan_vertex_is_ignored_by_bounds():
    soup = soup_for(12, 200)
    soup[5, 0] = np.nan
    o, _, _ = build(soup, 10)
    assert np.isfinite(o["scene"]).all()
    lv = ca.decode_leaves(o["leaves"], 200)
    k = int(np.where(lv["objIdx"] == 5)[0][0])
    # fmin/fmax skip the NaN corner: x bounds come from the other two corners
    assert lv["min"][k, 0] == min(soup[5, 3], soup[5, 6])


def test_pins_catch_planted_mistakes():
    """A wrong child index or a wrong box in the oracle's output must fail the pins."""
    soup = soup_for(13, 300)
    o, _, _ = build(soup, 10)
    bad = o["inodes"].copy()
    c = bad[7, 36:44].view(np.uint32)
    c[0], c[1] = c[1], c[0]
    with pytest.raises(AssertionError):
        check_topology_and_boxes({**o, "inodes": bad}, soup, 10)
    bad = o["inodes"].copy()
    bad[3, 0:4] = np.frombuffer(np.float32(-5.0).tobytes(), np.uint8)
    with pytest.raises(AssertionError):
        check_topology_and_boxes({**o, "inodes": bad}, soup, 10)
    bad_sorted = o["sorted"].copy()
    bad_sorted[[10, 11]] = bad_sorted[[11, 10]]
    keys = orc.interleave(o["morton"], 10)
    assert not np.array_equal(bad_sorted, orc.stable_order(keys))


@pytest.mark.parametrize("case", ["clusters", "duplicates", "identical", "nan", "flat"])
@pytest.mark.parametrize("bits", [10, 21])
def test_parallel_build_equals_serial_build(case, bits):
    """The multi-threaded oracle build (bench.py's all-cores CPU baseline)
    gives the serial restatement's nodes byte for byte."""
    if case == "clusters":
        soup = ca.gen_cluster_soup(100003, seed=5)
    elif case == "duplicates":
        soup = soup_for(21, 5000, dup=True)
    elif case == "identical":
        soup = np.tile(np.float32([0.1, 0.2, 0.3, 0.2, 0.3, 0.4, 0.3, 0.1, 0.2]), (777, 1))
    elif case == "nan":
        soup = soup_for(22, 3000)
        soup[17, 4] = np.nan
    else:
        soup = soup_for(23, 4000)
        soup[:, 1::3] = np.float32(0.5)
    o, v, f = build(soup, bits)
    for threads in (1, 3, 8):
        p = orc.lbvh_build_parallel(v, f, bits, threads)
        np.testing.assert_array_equal(p["leaves"], o["leaves"])
        np.testing.assert_array_equal(p["inodes"], o["inodes"])
