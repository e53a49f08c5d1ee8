"""Parity at the benchmark's own sizes (the LBVH and closest-hit workload:
2^22 and 2^24 Gaussian-cluster triangles, seed 1; 2^22 random rays, seed 2).

The GPU builds are compared with the oracle's build of the same soup byte for
byte (not only through structural properties), and the first 2^18 of the
benchmark's rays are traced against the GPU-built 2^24 BVH with the
reference walk and the ordered walk, each bit-equal to the oracle's walk of
the same nodes (face, t and visit / test counts).
"""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

pytestmark = pytest.mark.gpu

_BUILT = {}


def gpu_and_oracle_build(ctx, n):
    import torch

    if n not in _BUILT:
        _BUILT.clear()  # keep one size in memory at a time
        soup = ca.gen_cluster_soup(n, seed=1)
        inodes, leaves = ca.node_buffers(n)
        ca.LBVHBuilder(ctx, n, 10).perform_soup(torch.from_numpy(soup).cuda(), n, 10, inodes, leaves)
        torch.cuda.synchronize()
        assert ctx.device_status() == 0
        v, f = ca.soup_to_mesh(soup)
        del soup
        _BUILT[n] = (inodes, leaves, v, f, orc.lbvh_build(v, f, 10))
    return _BUILT[n]


@pytest.mark.parametrize("log2n", [22, 24])
def test_build_equals_oracle_at_benchmark_size(gpu_ctx, log2n):
    n = 1 << log2n
    print(f"[full-size parity] LBVH build n=2^{log2n}", flush=True)
    inodes, leaves, _, _, o = gpu_and_oracle_build(gpu_ctx, n)
    np.testing.assert_array_equal(leaves.cpu().numpy().reshape(-1, 48), o["leaves"])
    np.testing.assert_array_equal(inodes.cpu().numpy().reshape(-1, 48)[: n - 1], o["inodes"])


@pytest.mark.parametrize("ordered", [False, True])
def test_benchmark_rays_against_the_16m_bvh(gpu_ctx, ordered):
    import torch

    n, R = 1 << 24, 1 << 18
    inodes, leaves, v, f, o = gpu_and_oracle_build(gpu_ctx, n)
    rays = np.ascontiguousarray(ca.gen_random_rays(1 << 22, seed=2)[:, :R])
    print(f"[full-size parity] {R} of the benchmark rays vs n=2^24, ordered={ordered}", flush=True)
    hit = torch.empty(R, dtype=torch.int32, device="cuda")
    t = torch.empty(R, dtype=torch.float32, device="cuda")
    st = torch.empty(2 * R, dtype=torch.int32, device="cuda")
    ca.trace_closest(gpu_ctx, inodes, leaves, n, torch.from_numpy(v).cuda(),
                     torch.from_numpy(f.view(np.uint8)).cuda(), torch.from_numpy(rays).cuda(), R, hit, t, st,
                     ordered=ordered)
    torch.cuda.synchronize()
    assert gpu_ctx.device_status() == 0
    if ordered:
        h, tt, s = orc.trace_ordered(o["inodes"], o["leaves"], n, v, f, rays)
    else:
        h, tt, s = orc.trace_closest(o["inodes"], o["leaves"], n, v, f, rays, stats=True)
    np.testing.assert_array_equal(hit.cpu().numpy().view(np.uint32), h)
    np.testing.assert_array_equal(t.cpu().numpy().view(np.uint32), tt.view(np.uint32))
    np.testing.assert_array_equal(st.cpu().numpy().view(np.uint32).reshape(R, 2), s)
    assert (h != 0xFFFFFFFF).mean() > 0.3
