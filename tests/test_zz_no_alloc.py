"""sadh_compute_gradients_tiled_noalloc: the tiled gradient with its
fixed-point lines in the library's device array instead of an allocated
workspace.  Two calls in a row, the second on a smaller scene, each equal bit
for bit to sadh_compute_gradients_tiled on the same inputs; a site count above
the array's capacity is refused."""
import ctypes as C

import numpy as np
import pytest

from sad_amd import _lib
from tests.scenes import converged_field, iss_of, random_sites

pytestmark = pytest.mark.gpu

NOALLOC_MAX_SITES = 131072


def _case(orc, w, h, n, seed):
    sites = random_sites(w, h, n, seed)
    c0, c1 = converged_field(orc, sites, w, h, passes=8)
    rng = np.random.default_rng(seed + 1)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    mask[rng.random((h, w)) < 0.1] = 0.0
    return dict(w=w, h=h, n=n, sites=sites, c0=c0, c1=c1, target=target, mask=mask)


def _run(gpu, name, c, removal):
    w, h, n = c["w"], c["h"], c["n"]
    grads = [gpu.zeros((n,)) for _ in range(10)]
    rd = gpu.zeros((n,))
    t = gpu.put(c["target"])
    gpu.call(name, gpu.put(c["c0"]), gpu.put(c["c1"]), t, t, gpu.put(c["mask"]), *grads, gpu.put(c["sites"]),
             iss_of(w, h), n, rd, int(removal), 0.0, w, h, w, h)
    return np.stack([gpu.get(g) for g in grads], 1), gpu.get(rd)


def test_noalloc_gradients_equal_the_allocating_entry_point(gpu, orc):
    big = _case(orc, 256, 192, 3000, 5)
    small = _case(orc, 96, 80, 400, 6)
    got = [_run(gpu, "sadh_compute_gradients_tiled_noalloc", big, True),
           _run(gpu, "sadh_compute_gradients_tiled_noalloc", small, True)]
    ref = [_run(gpu, "sadh_compute_gradients_tiled", big, True),
           _run(gpu, "sadh_compute_gradients_tiled", small, True)]
    for (g, r), (gr, rr) in zip(got, ref):
        assert np.array_equal(g.view(np.uint32), gr.view(np.uint32))
        assert np.array_equal(r.view(np.uint32), rr.view(np.uint32))
    assert np.any(got[1][0] != 0)
    # a larger site count than the device array holds: refused before any launch
    lib = _lib.load()
    z = C.c_void_p(16)
    rc = lib.sadh_compute_gradients_tiled_noalloc(*([z] * 15), z, 1.0, NOALLOC_MAX_SITES + 1, z, 0, 0.0, 2, 2, 2, 2,
                                                  None)
    assert rc == -3 and "exceed" in lib.sadh_last_error().decode()
