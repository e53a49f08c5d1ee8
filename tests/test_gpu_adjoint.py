"""GPU path-replay adjoint: against the CPU adjoint (oracle, itself pinned by
finite differences in test_oracle_adjoint.py), against central finite
differences of GPU forward renders with common random numbers (relative
error <= 1e-2), through the torch.autograd harness, and its cost relative to
the forward render on the 200k-triangle, 64-material differentiation scene."""
import time

import numpy as np
import pytest

import clear_amd as ca
import oracle as orc
from clear_amd.diff import DiffRenderer, render_image

pytestmark = pytest.mark.gpu

FLAGS = ca.RENDER_NO_RR


def test_matches_cpu_adjoint(gpu_ctx):
    import torch

    W = H = 64
    host = ca.HostScene(ca.SCENE_ADJOINT, W, H)
    print("[adjoint] gpu vs cpu 64x64 spp=4 depth=5", flush=True)
    r = DiffRenderer(gpu_ctx, host, 4, 5, flags=FLAGS)
    states = r.seed_states()
    rng = np.random.default_rng(3)
    d_img = rng.uniform(0.5, 1.5, size=(H, W, 3)).astype(np.float32)
    ga, gs, ge = r.adjoint(states, torch.from_numpy(d_img).cuda())
    torch.cuda.synchronize()
    assert gpu_ctx.device_status() == 0
    o = orc.lbvh_build(host.vertices, host.faces, 10)
    d4 = np.zeros((H * W, 4), np.float32)
    d4[:, :3] = d_img.reshape(-1, 3)
    ca_, cs, ce = orc.render_adjoint(host.arrays, host.num_faces, o["inodes"], o["leaves"], W, H, 4, 5,
                                     FLAGS, host.diff["mat_offsets"], host.diff["light_prop"],
                                     states.cpu().numpy().view(np.uint32), d4)
    for name, g, c in [("albedo", ga, ca_), ("sigma", gs, cs), ("emission", ge, ce)]:
        g = g.cpu().numpy()
        scale = np.abs(c).max()
        err = np.abs(g - c).max()
        print(f"[adjoint] {name}: max |gpu - cpu| = {err:.3g} (max |g| {scale:.3g})", flush=True)
        assert err <= 1e-3 * scale


def test_central_differences_common_random_numbers(gpu_ctx):
    import torch

    W = H = 128
    host = ca.HostScene(ca.SCENE_ADJOINT, W, H)
    print("[adjoint] fd 128x128 spp=8 depth=5", flush=True)
    r = DiffRenderer(gpu_ctx, host, 8, 5, flags=FLAGS)
    states = r.seed_states()
    g_img = torch.from_numpy(np.random.default_rng(9).uniform(0.5, 1.5, size=(H, W, 3)).astype(np.float32)).cuda()
    alb, sig, em = r.params()
    ga, gs, ge = r.adjoint(states, g_img)

    def loss(a, s, e):
        r.set_params(a, s, e)
        return float((r.render(states).double() * g_img.double()).sum())

    checks = []
    for m in torch.argsort(-ga.abs().sum(1))[:3].tolist():
        for c in range(3):
            checks.append(("albedo", m, c, float(ga[m, c])))
    for m in torch.argsort(-gs.abs())[:3].tolist():
        checks.append(("sigma", m, 0, float(gs[m])))
    for c in range(3):
        checks.append(("emission", 0, c, float(ge[c])))
    for kind, m, c, g in checks:
        h = 1e-3 * (15.0 if kind == "emission" else 1.0)
        ap, sp, ep = alb.clone(), sig.clone(), em.clone()
        am, sm, emm = alb.clone(), sig.clone(), em.clone()
        if kind == "albedo":
            ap[m, c] += h
            am[m, c] -= h
        elif kind == "sigma":
            sp[m] += h
            sm[m] -= h
        else:
            ep[c] += h
            emm[c] -= h
        fd = (loss(ap, sp, ep) - loss(am, sm, emm)) / (2 * h)
        rel = abs(g - fd) / abs(fd)
        print(f"[adjoint] fd {kind}[{m},{c}]: adjoint {g:.6g} fd {fd:.6g} rel {rel:.2e}", flush=True)
        assert rel <= 1e-2
    r.set_params(alb, sig, em)


def test_autograd_function_and_descent(gpu_ctx):
    import torch

    W = H = 96
    target_host = ca.HostScene(ca.SCENE_ADJOINT, W, H)  # hidden parameters (seed 4)
    init_host = ca.HostScene(ca.SCENE_ADJOINT, W, H, seed=5)
    print("[adjoint] autograd harness 96x96 spp=8", flush=True)
    rt = DiffRenderer(gpu_ctx, target_host, 8, 5, flags=FLAGS)
    states = rt.seed_states()
    target = rt.render(states)
    r = DiffRenderer(gpu_ctx, init_host, 8, 5, flags=FLAGS)
    a0, s0, e0 = r.params()
    albedo = a0.clone().requires_grad_(True)
    sigma = s0.clone().requires_grad_(True)
    emission = e0.clone().requires_grad_(True)
    losses = []
    for _ in range(4):
        img = render_image(albedo, sigma, emission, r, states)
        loss = ((img - target) ** 2).mean()
        losses.append(float(loss))
        albedo.grad = sigma.grad = emission.grad = None
        loss.backward()
        with torch.no_grad():  # steps of at most 0.05 per albedo component
            albedo -= 0.05 * albedo.grad / albedo.grad.abs().max()
            albedo.clamp_(0.05, 0.95)
    print(f"[adjoint] descent losses {losses}", flush=True)
    assert losses[-1] < 0.9 * losses[0]
    # the autograd gradient equals the adjoint applied to d loss / d image
    img = render_image(albedo, sigma, emission, r, states)
    loss = ((img - target) ** 2).mean()
    albedo.grad = None
    loss.backward()
    d_img = 2.0 * (img.detach() - target) / img.numel()
    ga, _, _ = r.adjoint(states, d_img)
    gref = ga.cpu().numpy()
    np.testing.assert_allclose(albedo.grad.cpu().numpy(), gref, rtol=1e-3, atol=1e-6 * np.abs(gref).max())


def test_adjoint_to_forward_time(gpu_ctx):
    """512x512, 32 spp, depth 5 on the 200,000-triangle sphere (64 materials)."""
    import torch

    host = ca.HostScene(ca.SCENE_ADJOINT, 512, 512)
    r = DiffRenderer(gpu_ctx, host, 32, 5, flags=FLAGS)
    states = r.seed_states()
    g_img = torch.ones(512, 512, 3, device="cuda")
    r.render(states)
    r.adjoint(states, g_img)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        r.render(states)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        r.adjoint(states, g_img)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, adj = (t1 - t0) / 3, (t2 - t1) / 3
    print(f"[adjoint] forward {fwd * 1e3:.1f} ms, adjoint {adj * 1e3:.1f} ms, ratio {adj / fwd:.2f}", flush=True)
    assert gpu_ctx.device_status() == 0
    assert adj / fwd < 4.0
