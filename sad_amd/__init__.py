"""sad_amd: MI355X (gfx950) HIP backend of the SAD fit/render path.

libsad_hip.so exports the C ABI of include/sad_hip.h; this package binds it
with ctypes (sad_amd._lib), drives the native fit/render loops
(sad_amd.trainer), registers the torch.ops.sad operator set on HIP tensors
(sad_amd.ops) and provides the differentiable render (sad_amd.differentiable).
"""
from . import _lib
from ._lib import SadHipError, load

__all__ = ["SadHipError", "load", "_lib"]
