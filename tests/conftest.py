import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def gpu():
    """Device helper for GPU tests: numpy <-> torch HIP tensors, the current
    stream handle, and the C ABI binding."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    from sad_amd import _lib
    _lib.load()
    return GpuHelper(torch)


class GpuHelper:
    def __init__(self, torch):
        self.torch = torch
        self.dev = torch.device("cuda:0")
        from sad_amd import _lib
        self.lib = _lib

    def put(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        elif a.dtype == np.uint16:
            a = a.view(np.int16)
        return self.torch.from_numpy(a.copy()).to(self.dev)

    def zeros(self, shape, dtype=np.float32):
        tdt = {np.float32: self.torch.float32, np.uint32: self.torch.int32, np.int32: self.torch.int32,
               np.uint16: self.torch.int16}[dtype]
        return self.torch.zeros(shape, dtype=tdt, device=self.dev)

    def get(self, t, dtype=None):
        self.torch.cuda.synchronize()
        a = t.detach().cpu().numpy()
        if dtype is not None:
            a = a.view(dtype)
        return a

    @property
    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def call(self, name, *args):
        conv = [a.data_ptr() if hasattr(a, "data_ptr") else a for a in args]
        self.lib.call(name, *conv, self.stream)
