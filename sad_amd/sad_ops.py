"""Module-level names of the reference's `sad_ops` package
(backends/pytorch/torch_ext/sad_ops/__init__.py), bound to the HIP operators.

`import sad_amd.sad_ops as sad_ops` stands where the reference imports
`torch_ext.sad_ops`; `sad_ops.ops` is `torch.ops.sad`."""
import torch

from . import ops as _register  # noqa: F401  (defines torch.ops.sad)

ops = torch.ops.sad

render_sad = ops.render_sad
render_sad_padded = ops.render_sad_padded
render_sad_backward = ops.render_sad_backward
init_candidates = ops.init_candidates
vpt_pass = ops.vpt_pass

__all__ = ["ops", "render_sad", "render_sad_padded", "render_sad_backward", "init_candidates", "vpt_pass"]
