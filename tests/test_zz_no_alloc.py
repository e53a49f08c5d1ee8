"""clr_lbvh_build_soup_noalloc against clr_lbvh_build_soup, through the C-ABI.

The no-allocation build works in the library's static per-device array
instead of the context's pool. Its nodes must equal the pooled build's byte
for byte, on two calls in a row, the second with a smaller input (the array
keeps the first call's contents), and it must leave the context's pool
unallocated. Sizes above its limit are rejected.
"""
import ctypes

import numpy as np
import pytest

import clear_amd as ca

pytestmark = pytest.mark.gpu


def _noalloc(ctx, tris, n, bits, inodes, leaves):
    L = ca.lib()
    fn = L.clr_lbvh_build_soup_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p]
    import torch

    return fn(ctx.handle, tris.data_ptr(), n, bits, inodes.data_ptr(), leaves.data_ptr(),
              torch.cuda.current_stream().cuda_stream)


def _max_tris():
    fn = ca.lib().clr_lbvh_noalloc_max_tris
    fn.restype = ctypes.c_uint32
    return int(fn())


def _pooled(ctx, tris, n, bits):
    import torch

    inodes, leaves = ca.node_buffers(n)
    ca.LBVHBuilder(ctx, n, bits).perform_soup(tris, n, bits, inodes, leaves)
    torch.cuda.synchronize()
    return inodes.cpu().numpy()[: (n - 1) * 48], leaves.cpu().numpy()


def test_noalloc_build_matches_pooled_build():
    import torch

    assert _max_tris() == 1 << 20
    fresh = ca.Context(0)  # its pool stays unallocated: the no-allocation build must not grow it
    try:
        for n, seed in ((_max_tris(), 31), (70001, 32)):  # the second call is smaller
            print(f"[no-alloc] n={n} seed={seed}", flush=True)
            tris = torch.from_numpy(ca.gen_cluster_soup(n, seed=seed)).cuda()
            inodes = torch.full((max(n - 1, 1) * 48,), 0xAB, dtype=torch.uint8, device="cuda")
            leaves = torch.full((n * 48,), 0xAB, dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            fresh.check(_noalloc(fresh, tris, n, 10, inodes, leaves))
            torch.cuda.synchronize()
            assert fresh.device_status() == 0
            assert fresh.scratch_bytes() == 0
            ref = ca.Context(0)
            try:
                want_i, want_l = _pooled(ref, tris, n, 10)
            finally:
                ref.close()
            np.testing.assert_array_equal(leaves.cpu().numpy(), want_l)
            np.testing.assert_array_equal(inodes.cpu().numpy()[: (n - 1) * 48], want_i)
    finally:
        fresh.close()


def test_noalloc_rejects_sizes_beyond_its_limit():
    import torch

    ctx = ca.Context(0)
    try:
        n = _max_tris() + 1
        tris = torch.zeros(9, dtype=torch.float32, device="cuda")
        inodes = torch.zeros(48, dtype=torch.uint8, device="cuda")
        leaves = torch.zeros(48, dtype=torch.uint8, device="cuda")
        assert _noalloc(ctx, tris, n, 10, inodes, leaves) == 4  # CLR_ERR_UNSUPPORTED, nothing launched
        assert "exceeds the no-allocation limit" in ca.lib().clr_last_error(ctx.handle).decode()
        assert _noalloc(ctx, tris, 1000, 11, inodes, leaves) == 4
    finally:
        ctx.close()
