"""torch.autograd.Function around the renderer's forward (dmt_render) and its
path-replay adjoint (dmt_adjoint), for the differentiable-rendering tests.

Materials are a (prim_count, 7) float32 tensor: albedo rgb, roughness,
emission rgb per analytic primitive.  The forward renders one frame with
reset_accum set and returns its image (H, W, 3); the backward hands dLoss/dI
to dmt_adjoint as a device array and returns dLoss/dmaterials.
"""
import numpy as np
import torch

import dmt


class RenderMaterials(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mats, scene, frame, device):
        f = frame.copy()
        f.reset_accum = 1
        scene.set_materials(mats.detach().to("cpu", torch.float32).numpy())
        scene.render(f)
        img = scene.readback(f.width, f.height)
        ctx.scene, ctx.frame, ctx.device = scene, f, device
        ctx.mats_device = mats.device
        return torch.from_numpy(np.ascontiguousarray(img[..., :3])).to(mats.device)

    @staticmethod
    def backward(ctx, grad_img):
        f = ctx.frame
        dimg = torch.zeros((f.height, f.width, 4), device="cuda", dtype=torch.float32)
        dimg[..., :3] = grad_img.to("cuda", torch.float32)
        n = ctx.scene.data.prims.shape[0]
        grads = torch.empty((n, dmt.MATERIAL_FLOATS), device="cuda", dtype=torch.float32)
        torch.cuda.synchronize()
        ctx.scene.adjoint(f, dimg.data_ptr(), grads.data_ptr())
        ctx.device.synchronize()
        return grads.to(ctx.mats_device), None, None, None


def render(mats, scene, frame, device):
    return RenderMaterials.apply(mats, scene, frame, device)


def adjoint_numpy(scene, frame, dimage, device):
    """dmt_adjoint with host arrays in and out (through torch device buffers)."""
    f = frame.copy()
    f.reset_accum = 1
    d = torch.zeros((f.height, f.width, 4), device="cuda", dtype=torch.float32)
    d[..., :3] = torch.from_numpy(np.ascontiguousarray(dimage[..., :3], np.float32)).cuda()
    n = scene.data.prims.shape[0]
    g = torch.empty((n, dmt.MATERIAL_FLOATS), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    scene.adjoint(f, d.data_ptr(), g.data_ptr())
    device.synchronize()
    return g.cpu().numpy()


def perturbed_materials(gt, seed, sigma=0.1):
    """The spheres' parameters (rows 1.., row 0 is the ground plane) + N(0, sigma),
    albedo kept in [0.05, 1] and emission >= 0.02 so the perturbed scene stays
    physical and away from the emission switch at 0."""
    rng = np.random.default_rng(seed)
    m = gt.copy()
    m[1:] = gt[1:] + rng.normal(0.0, sigma, size=gt[1:].shape).astype(np.float32)
    m[1:, 0:3] = np.clip(m[1:, 0:3], 0.05, 1.0)
    m[1:, 4:7] = np.maximum(m[1:, 4:7], 0.02)
    return m.astype(np.float32)
