"""GPU post-processing (clr_post_*) against the oracle's restatement of
post_processing.cl in its scatter form (itself pinned by closed forms in
test_oracle_post.py): scaling exactly; gaussianScattering and bloom, which add
the same products in a different order, to relative 1e-5 of the image's
largest value; tone-mapped bytes identical on >= 99.9 % of bytes and within 1
everywhere (device and host expf / powf may differ in the last ulp)."""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (7, 1), (3, 5), (9, 17), (48, 64), (67, 130)]


def hdr_image(h, w, spp, seed):
    rng = np.random.default_rng(seed)
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = rng.gamma(0.5, 2.0, (h, w, 3)) * spp
    k = max(1, h * w // 40)  # bright spots: bloom radii from 0 to the cap of 7
    ys, xs = rng.integers(0, h, k), rng.integers(0, w, k)
    img[ys, xs, rng.integers(0, 3, k)] = rng.uniform(30.0, 45.0, k).astype(np.float32) * spp
    img[0, 0, 1] = 200.0 * spp  # corner spot: clamped taps
    return img


def to_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("h,w", SIZES)
def test_post_kernels_match_oracle(gpu_ctx, h, w):
    import torch

    spp = 8
    img = hdr_image(h, w, spp, seed=h * 1000 + w)
    print(f"[post] {w}x{h}", flush=True)
    src = to_dev(img)
    base = np.random.default_rng(7).uniform(0, 1, (h, w, 4)).astype(np.float32)
    base[..., 3] = 0
    # scaling: exact
    out = torch.empty_like(src)
    ca.post_scaling(gpu_ctx, w, h, spp, src, out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy()[..., :3], orc.post_scaling(img, spp)[..., :3])
    for name, fn, ofn in [("gaussian", ca.post_gaussian_scattering, orc.post_gaussian),
                          ("bloom", ca.post_bloom, orc.post_bloom)]:
        out = to_dev(base)  # the kernels add onto the output
        fn(gpu_ctx, w, h, spp, src, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()[..., :3]
        ref = ofn(img, spp, base.copy())[..., :3]
        err = np.abs(got - ref).max()
        print(f"[post] {name} {w}x{h}: max |gpu - oracle| {err:.3g} (max {np.abs(ref).max():.3g})", flush=True)
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * float(np.abs(ref).max()))
    # clear
    ca.post_clear(gpu_ctx, w, h, out)
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0


@pytest.mark.parametrize("h,w", SIZES)
def test_tone_mapping_matches_oracle(gpu_ctx, h, w):
    import torch

    img = hdr_image(h, w, 1, seed=w * 1000 + h) / np.float32(4.0)
    rb = ca.bmp_row_bytes(w)
    dst = torch.full((h * rb,), 0xAB, dtype=torch.uint8, device="cuda")
    ca.post_tone_mapping(gpu_ctx, w, h, rb, to_dev(img), dst)
    torch.cuda.synchronize()
    got = dst.cpu().numpy().astype(int)
    ref = orc.post_tone(img, rb).astype(int)
    same = (got == ref).mean()
    print(f"[post] tone {w}x{h}: identical bytes {same:.5f}", flush=True)
    assert np.abs(got - ref).max() <= 1 and same >= 0.999


def test_render_post_chain_and_bmp(gpu_ctx, tmp_path):
    """render -> clear -> scaling -> bloom -> toneMapping, against the oracle's
    chain on the GPU's own render, then written as a BMP."""
    import torch

    W, H, spp = 96, 72, 16
    host = ca.HostScene(ca.SCENE_CORNELL, W, H)
    ds = ca.DeviceScene(gpu_ctx, host)
    st, px = ca.new_image(W, H)
    ca.seed_pixel_states(gpu_ctx, W, H, st)
    ca.render(gpu_ctx, ds, spp, 5, st, px, mode=ca.RENDER_WAVEFRONT)
    mid = torch.empty_like(px)
    glow = torch.empty_like(px)
    ca.post_clear(gpu_ctx, W, H, glow)
    ca.post_scaling(gpu_ctx, W, H, spp, px, mid)
    ca.post_bloom(gpu_ctx, W, H, 1, mid, glow)
    rb = ca.bmp_row_bytes(W)
    ldr = torch.empty(H * rb, dtype=torch.uint8, device="cuda")
    ca.post_tone_mapping(gpu_ctx, W, H, rb, glow, ldr)
    torch.cuda.synchronize()
    img = px.cpu().numpy().reshape(H, W, 4)
    ref = orc.post_tone(orc.post_bloom(orc.post_scaling(img, spp), 1), rb).astype(int)
    got = ldr.cpu().numpy().astype(int)
    print(f"[post] chain: identical bytes {(got == ref).mean():.5f}, mean byte {got.mean():.1f}", flush=True)
    assert np.abs(got - ref).max() <= 1 and (got == ref).mean() >= 0.999 and got.mean() > 10
    p = tmp_path / "cornell.bmp"
    ca.save_bmp(str(p), ldr.cpu().numpy(), W, H)
    assert p.stat().st_size == 54 + rb * H
