"""LBVH builds captured into a graph and replayed.

clr_lbvh_build, clr_lbvh_build_soup and clr_lbvh_build_segmented are
stream-ordered with no host synchronisation once their scratch is reserved,
so each can be captured (here through torch.cuda.graph, whose capture stream
is the call's stream) and replayed. Every replay must build from the inputs'
contents at replay time: the triangles change between replays (eager copies
and synchronisations in between), and each replay's nodes must equal the
oracle's build of those triangles byte for byte. The hierarchy's box
hand-off tags each exchange word with a build epoch kept in device memory,
so a replay never accepts a word from the previous replay as fresh.
"""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

pytestmark = pytest.mark.gpu

REPLAYS = 3


def capture(fn):
    import torch

    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, capture_error_mode="thread_local"):
        fn()
    torch.cuda.synchronize()
    return g


def check_nodes(inodes, leaves, n, verts, faces, what):
    o = orc.lbvh_build(verts, faces, 10)
    np.testing.assert_array_equal(leaves.cpu().numpy().reshape(-1, 48)[:n], o["leaves"], err_msg=what)
    np.testing.assert_array_equal(inodes.cpu().numpy().reshape(-1, 48)[: n - 1], o["inodes"], err_msg=what)


@pytest.mark.parametrize("n", [5000, 200000])
def test_captured_soup_build_replays(gpu_ctx, n):
    import torch

    soups = [ca.gen_cluster_soup(n, seed=40 + k) for k in range(REPLAYS + 1)]
    tris = torch.from_numpy(soups[0]).cuda()
    inodes, leaves = ca.node_buffers(n)
    b = ca.LBVHBuilder(gpu_ctx, n, 10)  # reserves the scratch: nothing is allocated while capturing
    g = capture(lambda: b.perform_soup(tris, n, 10, inodes, leaves))
    for k in range(1, REPLAYS + 1):
        print(f"[graph capture] soup n={n} replay {k}", flush=True)
        tris.copy_(torch.from_numpy(soups[k]))
        inodes.zero_()
        leaves.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert gpu_ctx.device_status() == 0
        v, f = ca.soup_to_mesh(soups[k])
        check_nodes(inodes, leaves, n, v, f, f"soup replay {k}")


def test_captured_mesh_build_replays(gpu_ctx):
    """The drop-in build (clr_float3 vertices + Face records): the vertices
    move between replays, the faces stay."""
    import torch

    n = 60000
    soups = [ca.gen_cluster_soup(n, seed=70 + k) for k in range(REPLAYS + 1)]
    v0, faces = ca.soup_to_mesh(soups[0])
    verts = torch.from_numpy(v0).cuda()
    fdev = torch.from_numpy(faces.view(np.uint8)).cuda()
    inodes, leaves = ca.node_buffers(n)
    b = ca.LBVHBuilder(gpu_ctx, n, 10)
    g = capture(lambda: b.perform(verts, fdev, n, 10, inodes, leaves))
    for k in range(1, REPLAYS + 1):
        print(f"[graph capture] mesh replay {k}", flush=True)
        vk, _ = ca.soup_to_mesh(soups[k])
        verts.copy_(torch.from_numpy(vk))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert gpu_ctx.device_status() == 0
        check_nodes(inodes, leaves, n, vk, faces, f"mesh replay {k}")


def test_captured_segmented_build_replays(gpu_ctx):
    import torch

    sizes = [3000, 1, 70000, 2048, 15000]
    offs = np.zeros(len(sizes) + 1, np.uint32)
    offs[1:] = np.cumsum(sizes)
    n, k = int(offs[-1]), len(sizes)
    rounds = [[ca.gen_cluster_soup(m, seed=500 + 10 * r + s) for s, m in enumerate(sizes)]
              for r in range(REPLAYS + 1)]
    tris = torch.from_numpy(np.concatenate(rounds[0])).cuda()
    offs_dev = torch.from_numpy(offs.view(np.int32)).cuda()
    inodes = torch.zeros((n - k) * 48, dtype=torch.uint8, device="cuda")
    leaves = torch.zeros(n * 48, dtype=torch.uint8, device="cuda")
    gpu_ctx.check(ca.lib().clr_lbvh_reserve_segmented(gpu_ctx.handle, n, k, 10))
    b = ca.LBVHBuilder(gpu_ctx, 0, 10)
    g = capture(lambda: b.perform_segmented(tris, offs_dev, n, 10, inodes, leaves))
    for r in range(1, REPLAYS + 1):
        print(f"[graph capture] segmented replay {r}", flush=True)
        tris.copy_(torch.from_numpy(np.concatenate(rounds[r])))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert gpu_ctx.device_status() == 0
        gi = inodes.cpu().numpy().reshape(-1, 48)
        gl = leaves.cpu().numpy().reshape(-1, 48)
        for s, sp in enumerate(rounds[r]):
            m, base = sp.shape[0], int(offs[s])
            v, f = ca.soup_to_mesh(sp)
            o = orc.lbvh_build(v, f, 10)
            np.testing.assert_array_equal(gl[base:base + m], o["leaves"], err_msg=f"replay {r} scene {s}")
            if m > 1:
                np.testing.assert_array_equal(gi[base - s:base - s + m - 1], o["inodes"],
                                              err_msg=f"replay {r} scene {s}")


@pytest.mark.parametrize("mode", [None, ca.RENDER_MEGAKERNEL], ids=["wavefront", "megakernel"])
def test_captured_batch_pipeline_replays(gpu_ctx, mode):
    """One graph holds a batch's whole frame: the segmented build of every
    scene's LBVH (clr_lbvh_build_segmented), the seeding of the pixel states,
    the clearing of the image and the batched render: the wavefront form
    (clr_render_batch), which under capture runs its iteration bound as plain
    launches (no early-stop readback, no nested graph), or the megakernel
    form (clr_render_batch_mode); both upload their scene table through
    kernel arguments. The scenes' vertices (and the triangle
    soup the build reads) move between replays; each replay's nodes must equal
    the oracle's per-scene builds of the moved geometry, and its states and
    image the eager pipeline's on the same geometry, bit for bit."""
    import torch

    W, H, spp, depth = 40, 28, 2, 4
    sizes = [1500, 4000, 900]
    hosts = [ca.HostScene(ca.SCENE_CLUSTERS, W, H, seed=200 + i, num_triangles=m) for i, m in enumerate(sizes)]
    batch = ca.SceneBatch(gpu_ctx, hosts)
    k = len(hosts)
    st = torch.empty(k * W * H * 4, dtype=torch.int32, device="cuda")
    px = torch.zeros(k * W * H * 4, dtype=torch.float32, device="cuda")

    def frame():
        batch.build(10)
        for s in range(k):
            ca.seed_pixel_states(gpu_ctx, W, H, st[s * W * H * 4:(s + 1) * W * H * 4])
        px.zero_()
        ca.render_batch(gpu_ctx, batch, spp, depth, st, px, mode=mode)

    frame()  # eager: sizes the batch buffers, the wavefront pool and the trace spill
    torch.cuda.synchronize()
    assert gpu_ctx.device_status() == 0
    g = capture(frame)
    verts = [h.vertices.copy() for h in hosts]
    for r in range(REPLAYS):
        if r:
            d = np.float32(0.01 * r)
            for s in range(k):
                verts[s][:, 0] += d
                batch.t[s]["vertices"].view(torch.float32).view(-1, 4)[:, 0] += float(d)
            batch.tris.view(-1, 3)[:, 0] += float(d)
        px.fill_(-1.0)
        torch.cuda.synchronize()
        print(f"[graph capture] batch pipeline ({'wavefront' if mode is None else 'megakernel'}) replay {r}",
              flush=True)
        g.replay()
        torch.cuda.synchronize()
        assert gpu_ctx.device_status() == 0
        got = [t.clone() for t in (px, st, batch.inodes, batch.leaves)]
        assert float(px.abs().sum()) > 0.0
        gl = batch.leaves.cpu().numpy().reshape(-1, 48)
        gi = batch.inodes.cpu().numpy().reshape(-1, 48)
        for s, h in enumerate(hosts):
            o, m = int(batch.offsets[s]), h.num_faces
            ob = orc.lbvh_build(verts[s], h.faces, 10)
            np.testing.assert_array_equal(gl[o:o + m], ob["leaves"], err_msg=f"replay {r} scene {s}")
            np.testing.assert_array_equal(gi[o - s:o - s + m - 1], ob["inodes"], err_msg=f"replay {r} scene {s}")
        frame()
        torch.cuda.synchronize()
        for a, b, what in zip(got, (px, st, batch.inodes, batch.leaves), ("image", "states", "inodes", "leaves")):
            assert torch.equal(a, b), f"replay {r}: captured {what} differ from the eager frame's"
