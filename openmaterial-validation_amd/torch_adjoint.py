"""torch.autograd.Function over the render core's C-ABI: the forward is
omv_renderer_render with the material parameters taken from torch tensors,
the backward is omv_renderer_backward (path replay, detached sampling) into
torch tensors on the same device.  torch is plumbing here (device memory and
streams); the product is the HIP render core."""
from __future__ import annotations

import numpy as np
import torch


class MaterialFit:
    """One differentiable render: scene (committed), renderer, camera and the
    ids of the parameters (fp32 base-colour texture, fp32 metallic-roughness
    texture, OpenMATERIAL material)."""

    def __init__(self, scene, renderer, camera, base_tex: int, mr_tex: int, om_mat: int | None):
        self.scene, self.renderer, self.camera = scene, renderer, camera
        self.base_tex, self.mr_tex, self.om_mat = base_tex, mr_tex, om_mat

    def __call__(self, base: torch.Tensor, metal_rough: torch.Tensor, n: torch.Tensor, k: torch.Tensor) -> torch.Tensor:
        return RenderFunction.apply(self, base, metal_rough, n, k)


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fit: MaterialFit, base, metal_rough, n, k):
        base_c = base.detach().contiguous().float()
        mr_c = metal_rough.detach().contiguous().float()
        if base_c.is_cuda:
            torch.cuda.synchronize(base_c.device)
            fit.scene.update_texture(fit.base_tex, base_c.data_ptr())
            fit.scene.update_texture(fit.mr_tex, mr_c.data_ptr())
        else:
            fit.scene.update_texture(fit.base_tex, base_c.numpy())
            fit.scene.update_texture(fit.mr_tex, mr_c.numpy())
        if fit.om_mat is not None:
            fit.scene.update_om(fit.om_mat, n.detach().cpu().numpy(), k.detach().cpu().numpy())
        ctx.fit = fit
        ctx.shapes = (base.shape, metal_rough.shape)
        ctx.device = base.device
        if base.is_cuda:
            # the film is written straight into a device tensor (no host round trip)
            img = torch.empty((fit.camera.height, fit.camera.width, 4), dtype=torch.float32, device=base.device)
            torch.cuda.synchronize(base.device)
            fit.renderer.render_device(fit.camera, img.data_ptr())
            return img
        img = fit.renderer.render(fit.camera)
        return torch.from_numpy(np.ascontiguousarray(img)).to(base.device)

    @staticmethod
    def backward(ctx, grad_img):
        fit = ctx.fit
        dev = ctx.device if ctx.device.type == "cuda" else torch.device("cuda")
        g = grad_img.detach().to(dev).contiguous().float()
        gb = torch.zeros(ctx.shapes[0], dtype=torch.float32, device=dev)
        gmr = torch.zeros(ctx.shapes[1], dtype=torch.float32, device=dev)
        gom = torch.zeros(6, dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        fit.renderer.backward(fit.camera, g.data_ptr(), textures=[(fit.base_tex, gb.data_ptr()), (fit.mr_tex, gmr.data_ptr())],
                              om_materials=[fit.om_mat] if fit.om_mat is not None else [], om_grads_ptr=gom.data_ptr())
        out = ctx.device
        return None, gb.to(out), gmr.to(out), gom[:3].to(out), gom[3:].to(out)
