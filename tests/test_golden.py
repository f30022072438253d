"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py):
the oracle must keep reproducing them (CPU), and the HIP kernels must match
them bit for bit (GPU, through the C ABI)."""
import os

import numpy as np
import pytest

from tests.golden import make_golden as mg
from tests.scenes import hilbert_tables, iss_of

G = os.path.dirname(mg.__file__)


def _load(name):
    return np.load(os.path.join(G, name))


# ------------------------------------------------------------------ oracle (CPU)

def test_oracle_candidate_fixtures(orc):
    _, sites = mg.fit256_scene()
    ref = _load("cand_fit256.npz")
    c0, c1 = orc.init_candidates(len(sites), 0, False, 256, 256)
    c0 = orc.jfa_seed(c0, sites, 256, 256)
    assert [mg.digest(c0), mg.digest(c1)] == list(ref["init_sha"])
    f0, f1 = mg.vpt_passes(orc, c0, c1, sites, 256, 256, 16)
    assert [mg.digest(f0), mg.digest(f1)] == list(ref["vpt16_sha"])
    assert np.array_equal(f0[:32 * 256], ref["vpt16_rows"][0])
    s = mg.small_scene()
    c0, c1 = orc.init_candidates(len(s), 7, False, 64, 48)
    c0 = orc.jfa_seed(c0, s, 64, 48)
    f0, f1 = mg.vpt_passes(orc, c0, c1, s, 64, 48, 10)
    small = _load("cand_small.npz")
    assert np.array_equal(f0, small["cand0"]) and np.array_equal(f1, small["cand1"])


def test_oracle_sort_and_adam_fixtures(orc):
    sites, mass, energy, keys16 = mg.score_pairs_inputs()
    ref = _load("sort_pairs.npz")
    pairs = orc.densify_pairs(sites, mass, energy, 1.0, 0.7, len(sites))
    assert np.array_equal(pairs, ref["pairs"])
    assert np.array_equal(orc.radix_sort(pairs), ref["sorted"])
    assert np.array_equal(orc.radix_sort(keys16, 1 << 16), ref["sorted16"])
    # the sorted fixture is a stable sort by key (independent check)
    assert np.array_equal(ref["sorted"], ref["pairs"][np.argsort(ref["pairs"][:, 0], kind="stable")])
    s, st, g = mg.adam_inputs()
    ns, nst = orc.adam(s, st, g, mg.ADAM_LRS, 0.9, 0.999, 1e-8, 7, 200, 160)
    ad = _load("adam_step.npz")
    assert np.array_equal(ns, ad["sites"]) and np.array_equal(nst, ad["state"])


# ------------------------------------------------------------------ HIP kernels

@pytest.mark.gpu
def test_gpu_candidate_fixtures(gpu):
    _, sites = mg.fit256_scene()
    w = h = 256
    n = len(sites)
    ref = _load("cand_fit256.npz")
    c0, c1 = gpu.zeros((w * h, 4), np.uint32), gpu.zeros((w * h, 4), np.uint32)
    gpu.call("sadh_init_candidates", c0, c1, n, 0, 0, w, h)
    gs = gpu.put(sites)
    gpu.call("sadh_jfa_seed", c0, gs, n, w, h, 1)
    a0, a1 = gpu.get(c0, np.uint32), gpu.get(c1, np.uint32)
    assert [mg.digest(a0), mg.digest(a1)] == list(ref["init_sha"])
    pk = gpu.zeros((n, 8), np.uint16)
    gpu.call("sadh_pack_candidate_sites", gs, pk, n)
    from oracle import orc
    ho, hp = hilbert_tables(orc, sites, w, h)
    gho, ghp = gpu.put(ho), gpu.put(hp)
    b0, b1 = gpu.zeros((w * h, 4), np.uint32), gpu.zeros((w * h, 4), np.uint32)
    for i in range(16):
        step = (orc.jump_step(i, w, h) << 16) | (i & 0xFFFF)
        gpu.call("sadh_update_candidates", c0, c1, b0, b1, pk, n, step, iss_of(w, h), 0, 64.0, 0, 4, gho, ghp, 1,
                 32, 1, w, h, w, h)
        c0, b0, c1, b1 = b0, c0, b1, c1
    f0, f1 = gpu.get(c0, np.uint32), gpu.get(c1, np.uint32)
    assert np.array_equal(f0[:32 * w], ref["vpt16_rows"][0])
    assert [mg.digest(f0), mg.digest(f1)] == list(ref["vpt16_sha"])


@pytest.mark.gpu
def test_gpu_sort_and_adam_fixtures(gpu):
    ref = _load("sort_pairs.npz")
    keys16 = mg.score_pairs_inputs()[3]
    for src, key, mk in ((ref["pairs"], "sorted", 0xFFFFFFFF), (keys16, "sorted16", 1 << 16)):
        n = len(src)
        d = gpu.put(src)
        scratch = gpu.zeros((n, 2), np.uint32)
        hist = gpu.zeros((256 * ((n + 1023) // 1024),), np.uint32)
        bs = gpu.zeros((16,), np.uint32)
        gpu.call("sadh_radix_sort_pairs", d, scratch, hist, bs, n, mk)
        assert np.array_equal(gpu.get(d, np.uint32), ref[key])
    s, st, g = mg.adam_inputs()
    gs, gst = gpu.put(s), gpu.put(st)
    gg = [gpu.put(g[j]) for j in range(10)]
    gpu.call("sadh_adam_update", gs, gst, *gg, *mg.ADAM_LRS, 0.9, 0.999, 1e-8, 7, len(s), 200, 160)
    ad = _load("adam_step.npz")
    assert np.array_equal(gpu.get(gs), ad["sites"]) and np.array_equal(gpu.get(gst), ad["state"])
