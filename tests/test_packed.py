"""Quantised 16-byte inference sites (makePackedInferenceSites,
backends/metal/sources/utils/packed_inference.swift:86-199; decodePackedSite,
shaders/sad_common.metal:71-120; renderPackedInference, rendering/packed_renderer.swift;
--packed-psnr, training/training.swift:676-703).

CPU tests pin the oracle: the bit layout on known answers, the ranges against
numpy, round-trip error within half a quantisation step per field, the atan2
used for the direction against numpy.  GPU tests: device packing and decoding
bit-exact against the oracle, the packed render job equal to the render job on
the decoded sites, and the fit's packed PSNR against a host PSNR of its render."""
import numpy as np
import pytest

from sad_amd import config, synth

W, H = 300, 200


def _sites(n, seed, inactive=0.1):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 10), np.float32)
    s[:, 0] = rng.uniform(0, W - 1, n)
    s[:, 1] = rng.uniform(0, H - 1, n)
    s[:, 2] = rng.uniform(2.0, 8.0, n)
    s[:, 3] = rng.uniform(1.0, 30.0, n)
    s[:, 4:7] = rng.uniform(0.05, 0.95, (n, 3))
    a = rng.uniform(-np.pi, np.pi, n)
    s[:, 7], s[:, 8] = np.cos(a), np.sin(a)
    s[:, 9] = rng.uniform(-3, 3, n)
    k = int(n * inactive)
    s[:k, 0:2] = -1.0
    s[:k, 2] = 50.0                       # out of range, but inactive: must not widen the ranges
    return s


def test_atan2_against_numpy(orc):
    rng = np.random.default_rng(0)
    a = rng.uniform(-np.pi, np.pi, 200000)
    r = rng.uniform(0.01, 10, 200000)
    y, x = (r * np.sin(a)).astype(np.float32), (r * np.cos(a)).astype(np.float32)
    got = orc.atan2(y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - ref)) <= 1e-6
    sp = orc.atan2(np.array([0, 0, 1, -1, 0, 1, -1], np.float32), np.array([1, -1, 0, 0, 0, 1, -1], np.float32))
    assert np.allclose(sp, [0, np.pi, np.pi / 2, -np.pi / 2, 0, np.pi / 4, -3 * np.pi / 4], atol=3e-7)


def test_known_answers(orc):
    s = np.zeros((6, 10), np.float32)
    s[:, 7] = 1.0
    s[0, 0:2] = (0, 0)
    s[1, 0:2] = (W - 1, H - 1)
    s[2, 0:2] = (-1, -1)                                  # inactive
    s[3, 0:2] = (10, 10); s[3, 7:9] = (-1.0, 0.0)         # angle pi -> 65535
    s[4, 0:2] = (10, 10); s[4, 7:9] = (0.0, 1.0)          # pi/2 -> 0.75 * 65535 + 0.5 = 49151.75
    s[5, 0:2] = (10, 10); s[5, 7:9] = (0.0, 0.0)          # degenerate -> (1, 0)
    s[:, 2] = [1, 2, 99, 3, 3, 3]
    s[:, 3] = [4, 8, 99, 5, 5, 5]
    s[:, 4] = [0.0, 1.0, 9, 0.5, 0.5, 0.5]
    s[:, 5] = [0.2, 0.2, 9, 0.2, 0.2, 0.2]                # one value: scale = 1e-6
    s[:, 6] = [0.0, 1.0, 9, 0.5, 0.5, 0.5]
    s[:, 9] = [0.0, 0.5, 0, -1.0, 2.0, 0.0]
    p, q = orc.pack_inference(s, W, H)
    assert q.tolist() == [np.float32(1), np.float32(2), np.float32(4), np.float32(4), 0.0, 1.0, np.float32(0.2),
                          np.float32(1e-6), 0.0, 1.0]
    assert p[0, 0] == 0 and p[1, 0] == 32767 | (32767 << 15)
    assert (p[2, 0] >> 30) & 1 == 1 and all(((p[i, 0] >> 30) & 1) == 0 for i in (0, 1, 3, 4, 5))
    assert p[0, 1] & 0x7ff == 0 and p[1, 1] & 0x7ff == 2047 and (p[1, 1] >> 22) & 0x3ff == 1023
    assert p[0, 2] & 0xffff == 0 and p[1, 2] & 0xffff == 32768 and p[1, 2] >> 16 == 65535
    assert p[0, 3] & 0xffff == 32768 and p[3, 3] & 0xffff == 65535 and p[4, 3] & 0xffff == 49151
    assert p[5, 3] & 0xffff == 32768
    assert p[1, 3] >> 16 == np.float16(0.5).view(np.uint16) and p[3, 3] >> 16 == np.float16(-1.0).view(np.uint16)
    d = orc.unpack_inference(p, q, W, H)
    assert d[2, 0] == -1.0 and d[2, 1] == -1.0
    assert d[1, 0] == W - 1 and d[1, 1] == H - 1 and d[0, 0] == 0.0
    assert abs(d[4, 7]) < 1e-4 and abs(d[4, 8] - 1.0) < 1e-6


def test_round_trip_and_ranges(orc):
    s = _sites(5000, 1)
    p, q = orc.pack_inference(s, W, H)
    act = s[:, 0] >= 0
    for k, col in enumerate([2, 3, 4, 5, 6]):
        v = s[act, col].astype(np.float64)
        assert q[2 * k] == np.float32(v.min())
        assert q[2 * k + 1] == np.float32(np.float32(v.max()) - np.float32(v.min()))
    d = orc.unpack_inference(p, q, W, H)
    assert np.all(d[~act, 0] == -1.0)
    e = np.abs(d[act].astype(np.float64) - s[act])
    assert e[:, 0].max() <= (W - 1) / 32767 * 0.5 + 1e-4 and e[:, 1].max() <= (H - 1) / 32767 * 0.5 + 1e-4
    assert e[:, 2].max() <= q[1] / 65535 * 0.5 + 1e-5 and e[:, 3].max() <= q[3] / 65535 * 0.5 + 1e-5
    assert e[:, 4].max() <= q[5] / 2047 * 0.5 + 1e-6 and e[:, 5].max() <= q[7] / 2047 * 0.5 + 1e-6
    assert e[:, 6].max() <= q[9] / 1023 * 0.5 + 1e-6
    ang = np.arctan2(d[act, 8], d[act, 7]) - np.arctan2(s[act, 8], s[act, 7])
    ang = np.abs((ang + np.pi) % (2 * np.pi) - np.pi)
    assert ang.max() <= np.pi / 65535 + 2e-6
    assert np.all(d[act, 9] == s[act, 9].astype(np.float16).astype(np.float32))


# ---------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_pack_unpack_bitexact(gpu, orc):
    s = _sites(20000, 2)
    s[5, 2] = np.nan                                      # non-finite values keep out of the ranges
    s[6, 4] = np.inf
    packed = gpu.zeros((20000, 4), np.uint32)
    quant = gpu.zeros((10,))
    gpu.call("sadh_pack_inference_sites", gpu.put(s), 20000, W, H, packed, quant)
    p, q = orc.pack_inference(s, W, H)
    assert np.array_equal(gpu.get(quant).view(np.uint32), q.view(np.uint32))
    assert np.array_equal(gpu.get(packed, np.uint32), p)
    dec = gpu.zeros((20000, 10))
    gpu.call("sadh_unpack_inference_sites", packed, quant, 20000, W, H, dec)
    assert np.array_equal(gpu.get(dec).view(np.uint32), orc.unpack_inference(p, q, W, H).view(np.uint32))


@pytest.mark.gpu
def test_gpu_render_packed_equals_render_of_decoded(orc):
    from sad_amd import trainer
    s = _sites(3000, 3)
    p, q = trainer.pack_inference(s, W, H)
    a = trainer.render_packed_sites(p, q, W, H, cand_passes=16)
    b = trainer.render_sites(orc.unpack_inference(p, q, W, H), W, H, cand_passes=16)
    assert np.array_equal(a["image"], b["image"])
    c = trainer.render_sites(s, W, H, cand_passes=16)
    assert trainer.psnr(a["image"], c["image"]) > 35.0     # quantisation moves the render only slightly


@pytest.mark.gpu
def test_gpu_fit_packed_psnr():
    from sad_amd import trainer
    t = synth.make_target(160, 120, 51, n_ellipses=20)
    s0 = synth.init_sites_uniform(t, 900, 51)
    opts = config.fit_options(iterations=100, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, max_sites=900, log_freq=50, packed_psnr=1)
    r = trainer.fit(t, s0, opts, want_render=True)
    print("final", r["final_psnr"], "packed", r["packed_psnr"])
    assert abs(trainer.psnr(r["packed_render"], t) - r["packed_psnr"]) <= 0.01
    assert r["final_psnr"] - 1.0 <= r["packed_psnr"] <= r["final_psnr"] + 0.05
