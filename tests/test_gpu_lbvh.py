"""GPU LBVH build parity: byte-exact against the CPU oracle, and at full size
(2^24 triangles) against size-independent properties computed with numpy."""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

pytestmark = pytest.mark.gpu


def gpu_build(ctx, soup, bits, mesh=False):
    import torch

    n = soup.shape[0]
    inodes, leaves = ca.node_buffers(n)
    b = ca.LBVHBuilder(ctx, 0)
    if mesh:
        v, f = ca.soup_to_mesh(soup)
        tv = torch.from_numpy(v).cuda()
        tf = torch.from_numpy(f.view(np.uint8)).cuda()
        b.perform(tv, tf, n, bits, inodes, leaves)
    else:
        b.perform_soup(torch.from_numpy(np.ascontiguousarray(soup)).cuda(), n, bits, inodes, leaves)
    torch.cuda.synchronize()
    return (inodes.cpu().numpy().reshape(-1, 48)[: n - 1], leaves.cpu().numpy().reshape(-1, 48))


def random_soup(n, seed, dup):
    rng = np.random.default_rng(seed)
    s = (rng.random((n, 9), dtype=np.float32) * np.float32(0.003)
         + rng.random((n, 1), dtype=np.float32)).astype(np.float32)
    if dup:
        s[: n // 3] = s[0]
    return s


def assert_same_as_oracle(ctx, soup, bits, mesh=False):
    n = soup.shape[0]
    print(f"[lbvh parity] n={n} bits={bits} mesh={mesh}", flush=True)
    gi, gl = gpu_build(ctx, soup, bits, mesh)
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, bits)
    np.testing.assert_array_equal(gl, o["leaves"], err_msg=f"leaves n={n} bits={bits}")
    if n > 1:
        np.testing.assert_array_equal(gi, o["inodes"], err_msg=f"inodes n={n} bits={bits}")
    assert ctx.device_status() == 0


@pytest.mark.parametrize("n", [1, 2, 3, 5, 63, 64, 65, 4095, 4096, 4097, 12289, 65537])
@pytest.mark.parametrize("bits", [10, 20])
def test_small_sizes_byte_exact(gpu_ctx, n, bits):
    assert_same_as_oracle(gpu_ctx, random_soup(n, n * 7 + bits, dup=n > 8), bits)


@pytest.mark.parametrize("bits", [1, 2, 3, 7, 11, 16, 21])
def test_bit_widths_byte_exact(gpu_ctx, bits):
    assert_same_as_oracle(gpu_ctx, random_soup(30000, bits, dup=True), bits)


def test_mesh_path_matches_soup_path(gpu_ctx):
    soup = ca.gen_cluster_soup(1 << 16, seed=1)
    assert_same_as_oracle(gpu_ctx, soup, 10, mesh=True)
    a = gpu_build(gpu_ctx, soup, 10, mesh=True)
    b = gpu_build(gpu_ctx, soup, 10, mesh=False)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_cluster_1m_byte_exact(gpu_ctx):
    assert_same_as_oracle(gpu_ctx, ca.gen_cluster_soup(1 << 20, seed=1), 10)


def test_identical_and_degenerate_inputs(gpu_ctx):
    pts = np.tile(np.float32([0.7, 0.1, 0.2]), (5000, 3))
    assert_same_as_oracle(gpu_ctx, pts, 10)  # zero extent: all codes 0, index order
    s = random_soup(3000, 5, dup=False)
    s[:, 2::3] = np.float32(0.5)  # flat in z
    assert_same_as_oracle(gpu_ctx, s, 10)
    s = random_soup(3000, 6, dup=False)
    s[17, 4] = np.nan  # NaN corner
    assert_same_as_oracle(gpu_ctx, s, 10)


def test_repeated_builds_reuse_scratch(gpu_ctx):
    big = random_soup(50000, 8, dup=False)
    small = random_soup(777, 9, dup=True)
    r1 = gpu_build(gpu_ctx, big, 10)
    gpu_build(gpu_ctx, small, 10)
    r3 = gpu_build(gpu_ctx, big, 10)
    np.testing.assert_array_equal(r1[0], r3[0])
    np.testing.assert_array_equal(r1[1], r3[1])
    assert_same_as_oracle(gpu_ctx, small, 10)


def numpy_morton_keys(soup, bits):
    """Independent float32 restatement of the Morton key (numpy IEEE ops)."""
    t = soup.reshape(-1, 3, 3)
    lo, hi = t.min(axis=1), t.max(axis=1)
    smin, smax = lo.min(axis=0), hi.max(axis=0)
    size = (smax - smin).astype(np.float32)
    c = ((lo + hi) * np.float32(0.5)).astype(np.float32)
    npos = ((c - smin) / size).astype(np.float32)
    v = (npos * np.float32(1 << bits)).astype(np.float32)
    q = np.where(v > 0, np.minimum(np.floor(v), (1 << bits) - 1), 0).astype(np.uint64)
    key = np.zeros(len(q), np.uint64)
    for b in range(bits):
        for ax in range(3):
            key |= ((q[:, ax] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + ax)
    return key, lo, hi


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_soup_at_unaligned_address(gpu_ctx, shift):
    """A soup that starts 4, 8 or 12 bytes past a 16-B boundary builds the same
    nodes as the oracle (the bounds pass reads triangles with 4-B loads)."""
    import torch

    n = 70001
    soup = random_soup(n, 11 + shift, dup=False)
    print(f"[lbvh parity] unaligned soup n={n} shift={shift * 4} B", flush=True)
    flat = torch.zeros(n * 9 + shift, dtype=torch.float32, device="cuda")
    flat[shift:] = torch.from_numpy(soup.reshape(-1)).cuda()
    view = flat[shift:]
    assert view.data_ptr() % 16 == 4 * shift
    inodes, leaves = ca.node_buffers(n)
    ca.LBVHBuilder(gpu_ctx, 0).perform_soup(view, n, 10, inodes, leaves)
    torch.cuda.synchronize()
    v, f = ca.soup_to_mesh(soup)
    o = orc.lbvh_build(v, f, 10)
    np.testing.assert_array_equal(leaves.cpu().numpy().reshape(-1, 48), o["leaves"])
    np.testing.assert_array_equal(inodes.cpu().numpy().reshape(-1, 48)[: n - 1], o["inodes"])
    assert gpu_ctx.device_status() == 0


def test_full_size_16m_properties(gpu_ctx):
    """2^24 triangles: permutation, sorted keys with index tie-break, exact leaf
    boxes, parent box == fmin/fmax of children, every leaf reached once."""
    n = 1 << 24
    soup = ca.gen_cluster_soup(n, seed=1)
    gi, gl = gpu_build(gpu_ctx, soup, 10)
    lv = ca.decode_leaves(gl, n)
    obj = lv["objIdx"].astype(np.int64)
    assert np.array_equal(np.sort(obj), np.arange(n))
    key, lo, hi = numpy_morton_keys(soup, 10)
    ks = key[obj]
    assert (ks[1:] >= ks[:-1]).all()
    eq = ks[1:] == ks[:-1]
    assert (obj[1:][eq] > obj[:-1][eq]).all()
    np.testing.assert_array_equal(lv["min"], lo[obj])
    np.testing.assert_array_equal(lv["max"], hi[obj])
    assert gpu_ctx.device_status() == 0
    ino = ca.decode_inodes(gi, n)
    c, leaf = ino["c"].astype(np.int64), ino["isChild"].astype(bool)
    cmin = np.where(leaf[:, :, None], lv["min"][np.minimum(c, n - 1)], ino["min"][np.minimum(c, n - 2)])
    cmax = np.where(leaf[:, :, None], lv["max"][np.minimum(c, n - 1)], ino["max"][np.minimum(c, n - 2)])
    np.testing.assert_array_equal(ino["min"], np.minimum(cmin[:, 0], cmin[:, 1]))
    np.testing.assert_array_equal(ino["max"], np.maximum(cmax[:, 0], cmax[:, 1]))
    # reachability: every internal node except the root is some node's internal
    # child exactly once, every leaf exactly once
    internal_children = c[~leaf]
    leaf_children = c[leaf]
    assert np.array_equal(np.sort(internal_children), np.arange(1, n - 1))
    assert np.array_equal(np.sort(leaf_children), np.arange(n))


# ---- segmented build ---------------------------------------------------------

def segmented_soup(sizes, seed0):
    soups = [ca.gen_cluster_soup(int(m), seed=seed0 + k) for k, m in enumerate(sizes)]
    offs = np.zeros(len(sizes) + 1, np.uint32)
    offs[1:] = np.cumsum(sizes)
    return np.concatenate(soups, axis=0), offs, soups


def run_segmented(ctx, soup, offs, bits=10):
    import torch

    n = int(offs[-1])
    k = len(offs) - 1
    inodes = torch.zeros(max(n - k, 1) * 48, dtype=torch.uint8, device="cuda")
    leaves = torch.zeros(n * 48, dtype=torch.uint8, device="cuda")
    offs_dev = torch.from_numpy(np.asarray(offs, np.uint32).view(np.int32)).cuda()
    ca.LBVHBuilder(ctx, n, bits).perform_segmented(torch.from_numpy(soup).cuda(), offs_dev, n, bits,
                                                   inodes, leaves)
    torch.cuda.synchronize()
    assert ctx.device_status() == 0
    return inodes.cpu().numpy().reshape(-1, 48), leaves.cpu().numpy().reshape(-1, 48)


def test_segmented_matches_oracle_per_scene(gpu_ctx):
    """Scenes of 1, 2, 3 ... 20,000 triangles built in one segmented build equal
    the oracle's build of each scene alone, byte for byte."""
    rng = np.random.default_rng(11)
    sizes = [1, 2, 3, 5, 1000, 1023, 1024, 1025] + list(np.exp(rng.uniform(0, np.log(20000), 29)).astype(int) + 1)
    soup, offs, soups = segmented_soup(sizes, 300)
    print(f"[lbvh segmented] {len(sizes)} scenes, {int(offs[-1])} triangles", flush=True)
    gi, gl = run_segmented(gpu_ctx, soup, offs)
    for s, sp in enumerate(soups):
        n = sp.shape[0]
        v, f = ca.soup_to_mesh(sp)
        o = orc.lbvh_build(v, f, 10)
        b = int(offs[s])
        np.testing.assert_array_equal(gl[b:b + n], o["leaves"], err_msg=f"scene {s} leaves")
        if n > 1:
            np.testing.assert_array_equal(gi[b - s:b - s + n - 1], o["inodes"], err_msg=f"scene {s} inodes")


def test_segmented_256_scenes_match_single_builds(gpu_ctx):
    """256 log-uniform scenes (1k..100k triangles) against one GPU build per scene."""
    import torch

    rng = np.random.default_rng(100)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(100000), 256)).astype(int)
    soup, offs, soups = segmented_soup(sizes, 100)
    print(f"[lbvh segmented] 256 scenes, {int(offs[-1])} triangles", flush=True)
    gi, gl = run_segmented(gpu_ctx, soup, offs)
    for s in range(0, 256, 17):
        n = soups[s].shape[0]
        inodes, leaves = ca.node_buffers(n)
        ca.LBVHBuilder(gpu_ctx, n, 10).perform_soup(torch.from_numpy(soups[s]).cuda(), n, 10, inodes, leaves)
        torch.cuda.synchronize()
        b = int(offs[s])
        np.testing.assert_array_equal(gl[b:b + n], leaves.cpu().numpy().reshape(-1, 48))
        np.testing.assert_array_equal(gi[b - s:b - s + n - 1], inodes.cpu().numpy().reshape(-1, 48)[: n - 1])


def test_segmented_rejects_bad_offsets(gpu_ctx):
    """Offsets live on the device and are checked there: an empty scene, a
    first offset other than 0 or a last one other than numTris raise status
    bit 4 and leave the outputs untouched; more scenes than triangles is
    rejected on the host."""
    import torch

    soup = torch.from_numpy(ca.gen_cluster_soup(100, seed=1)).cuda()
    inodes = torch.zeros(100 * 48, dtype=torch.uint8, device="cuda")
    leaves = torch.zeros(100 * 48, dtype=torch.uint8, device="cuda")
    b = ca.LBVHBuilder(gpu_ctx, 100, 10)
    assert gpu_ctx.device_status() == 0
    for offs in ([0, 50, 50, 100], [1, 100], [0, 60, 40], [0, 40, 30, 100]):
        print(f"[lbvh segmented] rejected offsets {offs}", flush=True)
        dev = torch.tensor(offs, dtype=torch.int32, device="cuda")
        b.perform_segmented(soup, dev, 100, 10, inodes, leaves)
        torch.cuda.synchronize()
        assert gpu_ctx.device_status() == 16
        assert int(inodes.count_nonzero()) == 0 and int(leaves.count_nonzero()) == 0
    with pytest.raises(ca.ClrError):
        b.perform_segmented(soup, torch.tensor([0, 1, 2, 3], dtype=torch.int32, device="cuda"), 2, 10, inodes,
                            leaves)


def test_segmented_large_scenes_match_oracle(gpu_ctx):
    """Scenes of the batched workload's upper range (300k-500k triangles)
    beside small ones, each byte-identical to the oracle's build of it."""
    sizes = [499388, 1007, 312500, 20000, 431000]
    soup, offs, soups = segmented_soup(sizes, 120)
    print(f"[lbvh segmented] large scenes {sizes}", flush=True)
    gi, gl = run_segmented(gpu_ctx, soup, offs)
    for s, sp in enumerate(soups):
        n = sp.shape[0]
        v, f = ca.soup_to_mesh(sp)
        o = orc.lbvh_build(v, f, 10)
        b = int(offs[s])
        np.testing.assert_array_equal(gl[b:b + n], o["leaves"], err_msg=f"scene {s} leaves")
        np.testing.assert_array_equal(gi[b - s:b - s + n - 1], o["inodes"], err_msg=f"scene {s} inodes")
