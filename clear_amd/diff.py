"""torch.autograd harness around the path-replay adjoint (clr_render_adjoint).

The differentiable parameters of a scene (per-material albedo [M, 3] and
Oren-Nayar roughness [M], and one emitter's emission [3]) live as constant
textures inside the device texture blob; DiffRenderer writes torch tensors
there, renders with clr_render from fixed per-pixel random states (so every
call is the same estimator: common random numbers) and back-propagates an
image gradient with clr_render_adjoint. RenderFunction wraps both as a
torch.autograd.Function.
"""
from __future__ import annotations

import numpy as np

from . import (RENDER_MEGAKERNEL, RENDER_NO_RR, DeviceScene, HostScene, new_image, render,
               render_adjoint, seed_pixel_states)


class DiffRenderer:
    """Forward and adjoint renders of one scene with torch parameters.

    `host.diff` (from clr_scene_diff_params) names the parameters' byte
    offsets; `flags` default to CLR_RENDER_NO_RR so the estimate is smooth in
    the parameters."""

    def __init__(self, ctx, host: HostScene, spp: int, max_depth: int, flags: int = RENDER_NO_RR,
                 mode: int = RENDER_MEGAKERNEL, device="cuda"):
        import torch

        self.ctx, self.host, self.spp, self.max_depth = ctx, host, spp, max_depth
        self.flags, self.mode = flags, mode
        self.ds = DeviceScene(ctx, host, device=device)
        d = host.diff
        if len(d["mat_offsets"]) == 0:
            raise ValueError("scene has no differentiable materials")
        self.mat_offsets = np.asarray(d["mat_offsets"], np.uint32)
        self.light_prop = int(d["light_prop"])
        tex = self.ds.t["textures"]
        n4 = tex.numel() // 4
        self.tex32 = tex[: n4 * 4].view(torch.float32)  # shares the blob's storage
        to = lambda a: torch.as_tensor(np.asarray(a, np.int64) // 4, device=device)
        self.alb_idx = to(d["albedo_offsets"])[:, None] + torch.arange(3, device=device)[None, :]
        self.sig_idx = to(d["sigma_offsets"])
        self.em_idx = to([d["emission_offset"]])[0] + torch.arange(3, device=device)
        self.num_materials = len(self.mat_offsets)
        self.width, self.height = host.width, host.height

    def params(self):
        """Current (albedo [M,3], sigma [M], emission [3]) from the blob."""
        return (self.tex32[self.alb_idx].clone(), self.tex32[self.sig_idx].clone(),
                self.tex32[self.em_idx].clone())

    def set_params(self, albedo, sigma, emission) -> None:
        self.tex32[self.alb_idx] = albedo.detach().to(self.tex32.dtype)
        self.tex32[self.sig_idx] = sigma.detach().to(self.tex32.dtype)
        self.tex32[self.em_idx] = emission.detach().to(self.tex32.dtype)

    def seed_states(self):
        import torch

        states = torch.empty(self.width * self.height * 4, dtype=torch.int32, device=self.tex32.device)
        seed_pixel_states(self.ctx, self.width, self.height, states)
        return states

    def render(self, states):
        """Image [H, W, 3] = radiance of spp samples from `states` / spp
        (`states` is not modified)."""
        st, px = new_image(self.width, self.height, device=self.tex32.device)
        st.copy_(states)
        render(self.ctx, self.ds, self.spp, self.max_depth, st, px, mode=self.mode, flags=self.flags)
        return px.view(self.height, self.width, 4)[..., :3] / self.spp

    def adjoint(self, states, d_image):
        """Gradients of sum(d_image * render(states)): (albedo [M,3], sigma [M], emission [3])."""
        import torch

        dev = self.tex32.device
        d4 = torch.zeros(self.height, self.width, 4, dtype=torch.float32, device=dev)
        d4[..., :3] = d_image.to(torch.float32)
        ga = torch.empty(self.num_materials, 3, dtype=torch.float32, device=dev)
        gs = torch.empty(self.num_materials, dtype=torch.float32, device=dev)
        ge = torch.empty(3, dtype=torch.float32, device=dev)
        render_adjoint(self.ctx, self.ds, self.spp, self.max_depth, states, d4, self.mat_offsets,
                       self.light_prop, ga, gs, ge, flags=self.flags)
        return ga, gs, ge


def _function():
    import torch

    class RenderFunction(torch.autograd.Function):
        """image = render(albedo, sigma, emission) with fixed random states;
        backward runs the path-replay adjoint."""

        @staticmethod
        def forward(ctx, albedo, sigma, emission, renderer: DiffRenderer, states):
            renderer.set_params(albedo, sigma, emission)
            ctx.renderer, ctx.states = renderer, states
            ctx.save_for_backward(albedo.detach().clone(), sigma.detach().clone(), emission.detach().clone())
            return renderer.render(states)

        @staticmethod
        def backward(ctx, grad_image):
            albedo, sigma, emission = ctx.saved_tensors
            ctx.renderer.set_params(albedo, sigma, emission)
            ga, gs, ge = ctx.renderer.adjoint(ctx.states, grad_image.contiguous())
            return ga, gs, ge, None, None

    return RenderFunction


_RenderFunction = None


def render_image(albedo, sigma, emission, renderer: DiffRenderer, states):
    """Differentiable render: see RenderFunction."""
    global _RenderFunction
    if _RenderFunction is None:
        _RenderFunction = _function()
    return _RenderFunction.apply(albedo, sigma, emission, renderer, states)
