"""The SAD render as a differentiable function of the sites, on the HIP kernels.

    field = CandidateField.seeded(sites, width, height)   # initial field + JFA seed, on the device
    field.refine(sites, passes=16)                        # VPT passes (the field carries no gradient)
    image = render(sites, field)                          # [H, W, 3]
    (image * weights).sum().backward()                    # sites.grad: [N, 10] in Site field order

The forward is torch.ops.sad.render_sad.  The backward is
torch.ops.sad.render_sad_backward, the tiled gradient kernel driven by the
upstream gradient; it recomputes each pixel's softmax weights from the sites
and the candidate field, so autograd keeps nothing per pixel beyond the field
the forward needed (32 B per pixel) and the sites (40 B per site).  Its sums
are the library's exact fixed-point sums, so a backward repeats bit for bit.

Reference counterparts: the SADRenderFunction / SADRenderer / CandidateUpdater
classes of backends/pytorch/modules.py:11-129; those run unchanged on the ops
registered by sad_amd.ops as well.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import ops  # noqa: F401  (registers torch.ops.sad)

_sad = torch.ops.sad


def inv_scale_sq(width: int, height: int) -> float:
    """1 / max(W, H)^2 in float32, the metric normalisation of every kernel (sad_cuda.cu:2117)"""
    m = np.float32(max(int(width), int(height)))
    return float(np.float32(1.0) / (m * m))


def jump_schedule(width: int, height: int, passes: int, first: int = 0) -> list[int]:
    """VPT jump of passes first .. first + passes - 1: pow2/2, pow2/4, ... then 1, pow2 the
    power of two at or above max(W, H) (jumpStepForIndex, sad_cuda.cu:1157-1178)"""
    pow2 = 1
    while pow2 < max(width, height):
        pow2 <<= 1
    stages = max(pow2.bit_length() - 1, 1)
    return [max(pow2 >> (min(i, stages - 1) + 1), 1) for i in range(first, first + passes)]


def _on_hip(t: torch.Tensor, name: str, dtype: torch.dtype) -> None:
    if t.device.type != "cuda":
        raise ValueError(f"{name}: expected a tensor on a HIP device, got {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")


@dataclass
class CandidateField:
    """A candidate field: per pixel 8 site ids in two int32 planes [W*H, 4] (0xFFFFFFFF = empty),
    and the index of its next VPT pass."""

    width: int
    height: int
    cand0: torch.Tensor
    cand1: torch.Tensor
    passes_done: int = 0

    def __post_init__(self):
        for name in ("cand0", "cand1"):
            t = getattr(self, name)
            _on_hip(t, name, torch.int32)
            if tuple(t.shape) != (self.width * self.height, 4):
                raise ValueError(f"{name}: expected shape ({self.width * self.height}, 4), got {tuple(t.shape)}")
        if self.cand1.device != self.cand0.device:
            raise ValueError("cand0 and cand1 are on different devices")

    @classmethod
    def seeded(cls, sites: torch.Tensor, width: int, height: int, seed: int = 0) -> "CandidateField":
        """xorshift-initialised field (initCandidatesKernel, render.cu:95-136) with every active site
        written into its home cell (jfaSeedKernel, jfa.cu:66-92), on the sites' device"""
        _on_hip(sites, "sites", torch.float32)
        with torch.cuda.device(sites.device):
            c0, c1 = _sad.init_candidates(int(width), int(height), int(sites.shape[0]), int(seed))
            _sad.jfa_seed(c0, sites.contiguous(), int(sites.shape[0]), 1, int(width), int(height))
        return cls(int(width), int(height), c0, c1)

    @torch.no_grad()
    def refine(self, sites: torch.Tensor, passes: int = 1, inject_count: int = 4, radius_probes: int = 0,
               radius_scale: float = 64.0) -> "CandidateField":
        """`passes` VPT passes (updateCandidatesKernel, jfa.cu:271-386) on the jump schedule, in place"""
        _on_hip(sites, "sites", torch.float32)
        if sites.device != self.cand0.device:
            raise ValueError("sites and the candidate field are on different devices")
        iss = inv_scale_sq(self.width, self.height)
        for jump in jump_schedule(self.width, self.height, passes, self.passes_done):
            self.cand0, self.cand1 = _sad.vpt_pass(self.cand0, self.cand1, sites, iss, self.width, self.height,
                                                   int(jump), int(radius_probes), float(radius_scale),
                                                   int(inject_count), 0)
        self.passes_done += int(passes)
        return self


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sites, cand0, cand1, iss, width, height):
        sites = sites.contiguous()
        ctx.save_for_backward(sites, cand0, cand1)
        ctx.dims = (float(iss), int(width), int(height))
        return _sad.render_sad(cand0, cand1, sites, float(iss), int(width), int(height))

    @staticmethod
    def backward(ctx, grad_image):
        sites, cand0, cand1 = ctx.saved_tensors
        iss, width, height = ctx.dims
        if grad_image.device != sites.device:
            raise ValueError("the image gradient is on another device than the sites")
        grad = _sad.render_sad_backward(cand0, cand1, sites, grad_image.contiguous(), iss, width, height)
        return grad, None, None, None, None, None


def render(sites: torch.Tensor, field: CandidateField, iss: float | None = None) -> torch.Tensor:
    """The softmax render (renderVoronoiKernel, render.cu:5-92) of `sites` [N, 10] through `field`,
    differentiable in the sites.  iss defaults to 1 / max(W, H)^2."""
    _on_hip(sites, "sites", torch.float32)
    if sites.dim() != 2 or sites.shape[1] != 10:
        raise ValueError(f"sites: expected shape (N, 10), got {tuple(sites.shape)}")
    if sites.device != field.cand0.device:
        raise ValueError("sites and the candidate field are on different devices")
    if iss is None:
        iss = inv_scale_sq(field.width, field.height)
    return _Render.apply(sites, field.cand0.contiguous(), field.cand1.contiguous(), iss, field.width, field.height)


__all__ = ["CandidateField", "inv_scale_sq", "jump_schedule", "render"]
