"""Test-only gradient check of the rasterizer's hand-written backward.

The backward under test sits behind a ``torch.autograd.Function``
(``RasterizerLoss``): its forward returns the training loss of a cloud and
keeps the gradients that the backend computes with it, and its backward
hands them to autograd. So the gradients that ``loss.backward()`` leaves on
the parameter tensors are exactly the backend's. A backend is a callable
``cloud -> (loss, [positions, scales, rotations, opacities, sh gradients])``:
the GPU trainer through the C ABI, or Oracle B (the FP32 model of the same
pipeline) on the CPU.

The reference values are FP64 central finite differences of the FP64
restatement of the reference's CPU path (Oracle A: render, then 0.8 L1 +
0.2 D-SSIM, src/train/reference.cpp:110-340), with step h = 1e-6 applied to
the float32 parameter (the step actually taken is the difference of the two
rounded floats). The loss has cutoffs (alpha >= 1/255, the transmittance stop
at 1e-4) where it jumps; a parameter whose forward and backward one-sided
differences disagree sits on such a jump, and is counted and left out rather
than compared. The error measure per parameter group is the relative L2
error ||g - fd|| / ||fd|| over the remaining parameters.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from oracle import oracle as O
from vsplat import fixtures as F

GROUPS = ("positions", "scales", "rotations", "opacities", "sh")


def as_stored(cloud, sh_storage: int = 0):
    """The cloud as the trainer holds it: the SH coefficients rounded to the
    storage type (the rest as halves in the reference's layout, or all bf16),
    so that the FP64 loss is taken of the same function the GPU differentiates."""
    c = cloud.copy()
    c.sh = F.unpack_sh(F.pack_sh(cloud.sh, sh_storage), cloud.coefficients, sh_storage).reshape(cloud.sh.shape).copy()
    return c


def crop_problem(count: int = 200, seed: int = 1004, degree: int = 3, sh_storage: int = 0):
    """The differentiation-stress case at gradcheck size: config 4's
    distributions (log-scales N(-3.5, 1.2), opacity logits N(1, 1.5)), the
    Gaussians packed into the frustum of a 64x64 crop at the centre of its
    1920x1080 view, so that pixels see many splats and most stop at the
    transmittance floor; the SH as the trainer stores them. Returns (cloud,
    camera, target, lambda)."""
    cloud = F.synthetic_cloud(count, degree, seed, sh_sigma=0.3, log_scale=(-3.5, 1.2), opacity_logit=(1.0, 1.5))
    cloud.positions[:, :2] *= 0.1
    cloud.positions[:, 2] *= 0.3
    cloud = as_stored(cloud, sh_storage)
    view = F.look_at((0, 0, -3), (0, 0, 0), (0, -1, 0), 60.0, 1920, 1080)
    crop = dataclasses.replace(view, width=64, height=64, cx=view.cx - (1920 - 64) / 2, cy=view.cy - (1080 - 64) / 2)
    return cloud, crop, F.smooth_target_explicit(64, 64, seed), 0.2


def fp64_loss(cloud, cam, target, lam, threads=None) -> float:
    image = O.render(cloud, cam, threads=threads)
    return O.loss(image, target, lam, threads=threads).loss


class RasterizerLoss(torch.autograd.Function):
    """loss(positions, scales, rotations, opacities, sh) through a backend."""

    @staticmethod
    def forward(ctx, backend, template, positions, scales, rotations, opacities, sh):
        cloud = template.copy()
        tensors = (positions, scales, rotations, opacities, sh)
        for name, t in zip(GROUPS, tensors):
            setattr(cloud, name, np.ascontiguousarray(t.detach().cpu().numpy(), np.float32).reshape(getattr(template, name).shape))
        loss, grads = backend(cloud)
        ctx.save_for_backward(*[torch.from_numpy(np.asarray(g, np.float64).reshape(t.shape).copy()) for g, t in zip(grads, tensors)])
        return torch.tensor(float(loss), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        return (None, None, *[grad * g for g in ctx.saved_tensors])


def analytic_gradients(backend, cloud):
    """The backend's gradients, as autograd leaves them on the parameters."""
    params = [torch.tensor(np.asarray(getattr(cloud, name), np.float64), requires_grad=True) for name in GROUPS]
    loss = RasterizerLoss.apply(backend, cloud, *params)
    loss.backward()
    return float(loss.detach()), [p.grad.numpy() for p in params]


def central_differences(cloud, cam, target, lam, group, indices, h=1e-6, threads=None):
    """FP64 central differences of the FP64 loss for the flat entries `indices`
    of one parameter group, with the two one-sided differences."""
    base = fp64_loss(cloud, cam, target, lam, threads)
    central, forward, backward = (np.zeros(len(indices)) for _ in range(3))
    for n, i in enumerate(indices):
        c = cloud.copy()
        values = getattr(c, group).reshape(-1)
        x0 = np.float32(values[i])
        xp, xm = np.float32(x0 + np.float32(h)), np.float32(x0 - np.float32(h))
        values[i] = xp
        high = fp64_loss(c, cam, target, lam, threads)
        values[i] = xm
        low = fp64_loss(c, cam, target, lam, threads)
        central[n] = (high - low) / (float(xp) - float(xm))
        forward[n] = (high - base) / (float(xp) - float(x0))
        backward[n] = (base - low) / (float(x0) - float(xm))
    return central, forward, backward


@dataclasses.dataclass
class GroupReport:
    group: str
    compared: int
    on_a_jump: int
    rel_err: float
    cosine: float


def compare(analytic, central, forward, backward, group) -> GroupReport:
    """Relative L2 error over the parameters that are not on a jump."""
    scale = float(np.sqrt(np.mean(central ** 2))) + 1e-30
    jump = np.abs(forward - backward) > 1e-3 * scale + 1e-2 * np.abs(central)
    keep = ~jump
    a, f = np.asarray(analytic, np.float64)[keep], central[keep]
    rel = float(np.linalg.norm(a - f) / max(np.linalg.norm(f), 1e-30))
    cos = float((a @ f) / max(np.linalg.norm(a) * np.linalg.norm(f), 1e-30))
    return GroupReport(group, int(keep.sum()), int(jump.sum()), rel, cos)


def check(backend, cloud, cam, target, lam, gaussians, threads=None, h=1e-6):
    """Reports per group for every parameter of the Gaussians listed."""
    _, grads = analytic_gradients(backend, cloud)
    reports = []
    for name, g in zip(GROUPS, grads):
        per = int(np.prod(getattr(cloud, name).shape[1:])) if getattr(cloud, name).ndim > 1 else 1
        idx = np.concatenate([np.arange(k * per, (k + 1) * per) for k in gaussians])
        central, fwd, bwd = central_differences(cloud, cam, target, lam, name, idx, h, threads)
        reports.append(compare(g.reshape(-1)[idx], central, fwd, bwd, name))
    return reports


class OracleBBackend:
    """Oracle B (the FP32 model of the GPU pipeline) as a backend, on the CPU:
    float or fixed-point sums, the loss from its per-tile totals."""

    def __init__(self, cam, target, lam, float_sums=True):
        self.cam, self.target, self.lam, self.float_sums = cam, target, lam, float_sums

    def __call__(self, cloud):
        m = O.Model(cloud, self.cam)
        r = m.forward()
        m.loss(r, self.target, self.lam)
        (m.backward_float if self.float_sums else m.backward_fixed)(r)
        g, _ = m.gradients(r)
        totals = r.tile_totals.astype(np.float64).sum(0) / (3.0 * self.cam.width * self.cam.height)
        return (1.0 - self.lam) * totals[1] + self.lam * (1.0 - totals[0]), g
