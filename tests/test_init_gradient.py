"""Gradient-weighted site initialisation (initGradientWeightedKernel,
backends/cuda/src/densify_prune.cu:5-83; sad_cuda.cu:2078-2095).

CPU tests pin the oracle:
  * seed positions against the published R2 sequence (Roberts 2018: alpha =
    (1/g, 1/g^2), g the plastic number) in float64, and bit-exactly against
    the float32 fused multiply-add route computed in numpy;
  * the Sobel gradient on a luma ramp against its closed form (2s in the
    interior, s in the clamped first column), with the acceptance and log_tau
    rules that follow from it, and mask weighting;
  * the serial seed walk against a walk written here over the per-seed
    acceptance, including sites that run out of attempts.
GPU tests: the HIP kernel bit-exact against the oracle (sites and the final
seed counter), the torch op, and a whole fit started from gradient init.
"""
import numpy as np
import pytest

from sad_amd import config, synth

G_PLASTIC = 1.32471795724474602596


def ramp(w, h, s):
    x = np.arange(w, dtype=np.float32) * np.float32(s)
    t = np.repeat(np.repeat(x[None, :, None], h, 0), 3, 2)
    return np.ascontiguousarray(t, np.float32)


def patch_image(w, h, seed):
    """flat grey with one textured square: long runs of rejected seeds"""
    t = np.full((h, w, 3), 0.5, np.float32)
    rng = np.random.default_rng(seed)
    t[h // 3: h // 3 + 6, w // 2: w // 2 + 6] = rng.uniform(0, 1, (6, 6, 3)).astype(np.float32)
    return t


def walk(accept, n, attempts, seed0=0):
    """site i draws from the seed after site i-1's last draw; -> (seed or -1 per site, end)"""
    out, s = [], seed0
    for _ in range(n):
        got = -1
        for _ in range(attempts):
            if accept(s):
                got = s
                s += 1
                break
            s += 1
        out.append(got)
    return out, s


def r2_float32(seed, w, h):
    """position of a seed by the float32 fused multiply-add route, in numpy"""
    g = np.float32(G_PLASTIC)
    a1 = np.float32(1.0) / g
    a2 = np.float32(1.0) / (g * g)
    fs = np.float32(seed)
    # a float32 product is exact in float64, and so is adding 0.5 to it at these
    # magnitudes, so one rounding to float32 reproduces fmaf
    u = np.float32(np.float64(fs) * np.float64(a1) + 0.5)
    v = np.float32(np.float64(fs) * np.float64(a2) + 0.5)
    u = u - np.floor(u)
    v = v - np.floor(v)
    return np.float32(u * np.float32(w - 1)), np.float32(v * np.float32(h - 1))


def test_r2_positions(orc):
    w, h = 513, 301
    t = np.zeros((h, w, 3), np.float32)
    a1, a2 = 1.0 / G_PLASTIC, 1.0 / G_PLASTIC ** 2
    assert abs(a1 - 0.7548776662466927) < 1e-15 and abs(a2 - 0.5698402909980532) < 1e-15
    for s in [0, 1, 2, 3, 17, 1000, 65535, 123457, 3000001]:
        ok, x, y, pix, g = orc.init_grad_eval(t, s)
        ex, ey = r2_float32(s, w, h)
        assert (np.float32(x), np.float32(y)) == (ex, ey), s
        assert pix == int(ey) * w + int(ex)
        # float64 R2 point; float32 keeps ~24 bits of s * alpha
        u64, v64 = (0.5 + s * a1) % 1.0, (0.5 + s * a2) % 1.0
        tolu = (s + 1) * 2.0 ** -23
        assert abs(x - u64 * (w - 1)) <= tolu * (w - 1) + 1e-3 and abs(y - v64 * (h - 1)) <= tolu * (h - 1) + 1e-3
        assert not ok and g == 0.0       # flat image: gradient 0, rejected
    assert orc.init_grad_eval(t, 0)[1:3] == (256.0, 150.0)   # seed 0 is the centre (0.5, 0.5)


def test_sobel_ramp_closed_form(orc):
    w, h, s = 300, 200, 0.002
    t = ramp(w, h, s)
    luma = np.float32(0.299) + np.float32(0.587) + np.float32(0.114)
    for seed in range(2000):
        ok, x, y, pix, g = orc.init_grad_eval(t, seed, threshold=1.5 * s)
        ix = pix % w
        expect = (s if ix == 0 else 2 * s) * luma
        assert abs(g - expect) <= 1e-6, (seed, ix, g, expect)
        assert ok == (ix >= 1)


def test_init_ramp_sites(orc):
    w, h, s, n = 300, 200, 0.002, 500
    t = ramp(w, h, s)
    sites, end = orc.init_gradient_weighted(t, n, threshold=0.001, log_tau0=5.0, radius0=4.0)
    assert end == n                                   # every seed accepted (gradient >= s > threshold)
    for i in range(n):
        ex, ey = r2_float32(i, w, h)
        assert sites[i, 0] == ex and sites[i, 1] == ey
        ix, iy = int(ex), int(ey)
        g = (s if ix == 0 else 2 * s)
        assert abs(sites[i, 2] - (5.0 + 2.0 * g ** 0.25)) <= 5e-5
        assert sites[i, 3] == 4.0
        assert np.array_equal(sites[i, 4:7], t[iy, ix])
        assert sites[i, 7:10].tolist() == [1.0, 0.0, 0.0]
    # threshold above the gradient everywhere: every site runs out of attempts
    sites, end = orc.init_gradient_weighted(t, 50, threshold=0.01, attempts=7)
    assert end == 50 * 7
    assert np.all(sites[:, 0:2] == -1.0) and np.all(sites[:, 4:7] == 0.5) and np.all(sites[:, 2] == 5.0)


def test_mask_weighting(orc):
    w, h, s = 200, 120, 0.004
    t = ramp(w, h, s)
    m = np.ones((h, w), np.float32)
    m[:, : w // 2] = 0.0
    m[: h // 2, w // 2:] = 0.25
    for seed in range(1500):
        ok, x, y, pix, g = orc.init_grad_eval(t, seed, mask=m, threshold=0.0025)
        mv = m.flat[pix]
        if mv <= 0.0:
            assert not ok and g == 0.0
        else:
            _, _, _, _, g1 = orc.init_grad_eval(t, seed, threshold=0.0025)
            assert g == np.float32(g1) * np.float32(mv)
            assert ok == (g > np.float32(0.0025))


@pytest.mark.parametrize("attempts,n,seed0", [(256, 400, 0), (5, 300, 0), (3, 200, 1234)])
def test_walk_matches_serial_rule(orc, attempts, n, seed0):
    w, h = 160, 120
    t = patch_image(w, h, 3) if attempts < 256 else synth.make_target(w, h, 5, n_ellipses=12)
    m = None
    acc = lambda sd: orc.init_grad_eval(t, sd, mask=m, threshold=0.01)[0]   # noqa: E731
    seeds, end = walk(acc, n, attempts, seed0)
    sites, got_end = orc.init_gradient_weighted(t, n, threshold=0.01, attempts=attempts, seed0=seed0)
    assert got_end == end
    if attempts < 256:
        assert seeds.count(-1) > 0          # some sites ran out of attempts
    for i, sd in enumerate(seeds):
        if sd < 0:
            assert sites[i, 0] == -1.0 and sites[i, 4] == 0.5
        else:
            _, x, y, pix, g = orc.init_grad_eval(t, sd, threshold=0.01)
            assert (sites[i, 0], sites[i, 1]) == (np.float32(x), np.float32(y))
            assert np.array_equal(sites[i, 4:7], t.reshape(-1, 3)[pix])


# ---------------------------------------------------------------------- GPU

def _gpu_init(gpu, t, n, mask=None, thr=0.01, attempts=256, seed0=0, log_tau0=5.0, radius0=4.0):
    h, w, _ = t.shape
    sites = gpu.zeros((n, 10))
    cnt = gpu.put(np.array([seed0], np.uint32))
    gpu.call("sadh_init_gradient_weighted", sites, n, cnt, gpu.put(t), None if mask is None else gpu.put(mask),
             thr, attempts, log_tau0, radius0, w, h)
    return gpu.get(sites), int(gpu.get(cnt, np.uint32)[0])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["natural", "natural_many_rounds", "masked", "patch_failures", "patch_open_gap",
                                  "ramp_none", "seed0"])
def test_gpu_init_bitexact(gpu, orc, case):
    w, h, n, thr, att, seed0, mask = 257, 190, 3000, 0.01, 256, 0, None
    t = synth.make_target(w, h, 11, n_ellipses=25)
    if case == "natural_many_rounds":
        n = 40000                                       # > 32768 seeds: several rounds of the walk
    elif case == "masked":
        mask = np.zeros((h, w), np.float32)
        mask[20:150, 30:200] = np.linspace(0.1, 1.0, 170, dtype=np.float32)[None, :]
    elif case == "patch_failures":
        t, att, n = patch_image(w, h, 4), 6, 2500
    elif case == "patch_open_gap":
        t, att, n = patch_image(w, h, 5), 40, 5000      # the walk ends in a run of failures
    elif case == "ramp_none":
        t, thr, att, n = ramp(w, h, 0.001), 0.5, 9, 1000
    elif case == "seed0":
        seed0 = 987654
    print(f"init case={case} n={n} attempts={att}", flush=True)
    got, end = _gpu_init(gpu, t, n, mask, thr, att, seed0)
    ref, rend = orc.init_gradient_weighted(t, n, mask=mask, threshold=thr, attempts=att, seed0=seed0)
    assert end == rend
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_gpu_init_op(gpu, orc):
    import torch
    from sad_amd import sad_ops
    w, h, n = 180, 140, 2000
    t = synth.make_target(w, h, 12, n_ellipses=20)
    m = np.ones((h, w), np.float32)
    s = torch.zeros((n + 10, 10), dtype=torch.float32, device="cuda")
    out = sad_ops.ops.init_gradient_weighted(s, gpu.put(t), gpu.put(m), n, 4.5, 3.0, 2.0)
    ref, _ = orc.init_gradient_weighted(t, n, threshold=np.float32(0.01) * np.float32(2.0), log_tau0=4.5,
                                        radius0=3.0)
    got = gpu.get(out)
    assert np.array_equal(got[:n].view(np.uint32), ref.view(np.uint32))
    assert np.all(got[n:] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fixed_sites", "densify_prune"])
def test_gpu_fit_gradient_init(orc, case):
    """sadh_fit with init_mode 2 against the oracle's fit from the same init.

    Without densify/prune the bar is the fit bar, 0.05 dB.  With them the
    pruned set follows score ranks that round-off can swap, and from a
    gradient-weighted init the schedule is chaotic: on this image a 1e-6
    relative change of the learning rate moves the oracle's own final PSNR by
    up to 0.06 dB (0.6 dB when pruning runs to the last iteration).  The GPU's
    final PSNR must then lie within 0.05 dB of the range the oracle spans
    under that perturbation."""
    from sad_amd import trainer
    if case == "fixed_sites":
        w, h = 160, 128
        over = dict(iterations=120, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                    prune_percentile=0.0, n_sites=1500, max_sites=1500, log_freq=40)
    else:
        w, h = 192, 160
        over = dict(iterations=200, target_bpp=2.0, densify_freq=20, prune_freq=20, densify_start=20,
                    densify_end=100, prune_start=40, prune_end=140, n_sites=800, max_sites=2400, log_freq=50)
    t = synth.make_target(w, h, 31, n_ellipses=30)
    opts = config.fit_options(init_mode=2, **over)
    g = trainer.fit(t, None, opts)
    r = orc.fit(t, None, opts)
    print(case, "gpu", g["log_psnr"], g["actual_sites"], "oracle", r["log_psnr"], r["actual_sites"])
    assert g["log_psnr"][0] == pytest.approx(r["log_psnr"][0], abs=1e-3)
    assert g["plan"] == r["plan"]
    assert (g["n_densify"], g["n_prune"], g["actual_sites"], g["active_estimate"]) == \
           (r["n_densify"], r["n_prune"], r["actual_sites"], r["active_estimate"])
    lo = hi = r["final_psnr"]
    if case == "densify_prune":
        for sc in (1.0 + 1e-6, 1.0 - 1e-6):
            rp = orc.fit(t, None, dict(opts, lr_scale=float(np.float32(sc))))
            print("oracle lr_scale", np.float32(sc), rp["log_psnr"])
            lo, hi = min(lo, rp["final_psnr"]), max(hi, rp["final_psnr"])
    assert lo - 0.05 <= g["final_psnr"] <= hi + 0.05
