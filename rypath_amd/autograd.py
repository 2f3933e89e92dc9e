"""PyTorch autograd harness over the C ABI's forward render and path-replay adjoint.

RenderXYZ.apply(values, scene, params, cfg) renders the raw XYZ film (H, W, 3)
on the GPU with the parameter values written into the scene (rp_set_params +
rp_render_frame_device); its backward calls rp_backward_device with the
incoming dLoss/dXYZ image (read in place on the device) and returns
dLoss/dvalues at the values the forward rendered with.  When the scene holds
other values by then (a second forward, or set_params in between), backward
writes the forward's values back first, so the adjoint traces its own primal
at the right parameters.
"""
import numpy as np
import torch

import rypath_amd as rp


class RenderXYZ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, scene, params, cfg):
        vals = values.detach().cpu().numpy().astype(np.float32)
        scene.set_params(params, vals)
        film = torch.empty((scene.height, scene.width, 3), dtype=torch.float32, device="cuda")
        scene.render_frame_device(cfg, film.data_ptr())
        torch.cuda.synchronize()
        ctx.scene, ctx.params, ctx.cfg = scene, params, cfg
        ctx.values, ctx.state = vals, scene.param_state
        ctx.device = values.device
        return film

    @staticmethod
    def backward(ctx, grad_film):
        if ctx.scene.param_state != ctx.state:
            ctx.scene.set_params(ctx.params, ctx.values)
        dl = grad_film.detach().float().contiguous()
        g, _ = ctx.scene.backward_device(ctx.cfg, ctx.params, dl.data_ptr())
        return torch.from_numpy(g).to(ctx.device), None, None, None


def render(values, scene, params, cfg):
    return RenderXYZ.apply(values, scene, params, cfg)
