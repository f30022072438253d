"""SSIM metric and loss term (computeSSIMKernel, backends/cuda/src/utils.cu:29-87;
the SSIM branch of computeGradientsTiledKernel, gradients.cu:236-266; the
fit's render-before-gradient and metric logs, sad_cuda.cu:2307-2330,2432-2447).

CPU tests pin the oracle:
  * the window SSIM against a float64 numpy restatement of the 3x3 box-window
    SSIM of Wang et al. (2004) on luma, borders clamped (edge padding);
  * the loss factor dSSIM/dx_k against float64 central differences of the
    pixel's own window SSIM in its own luma;
  * the gradient with the SSIM term against the L2 gradient plus the autograd
    backward of grad_output = -w dSSIM/dx_k (0.299, 0.587, 0.114): the site
    gradient is linear in dL/dpred.
GPU tests: the metric and the loss-term gradient against the oracle, a fit
with --ssim-weight and --ssim against the oracle's fit, and the batched engine
with both (and the packed PSNR) against single fits."""
import numpy as np
import pytest

from sad_amd import config, synth
from tests.scenes import converged_field, iss_of, random_sites
from tests.tolerance import Rel

LUMA = np.array([0.299, 0.587, 0.114])
C1, C2 = 0.0001, 0.0009


def ssim64(r, t):
    """per-pixel SSIM of the 3x3 box window of the luma, float64, edges clamped"""
    x, y = r.astype(np.float64) @ LUMA, t.astype(np.float64) @ LUMA
    h, w = x.shape

    def box(a):
        p = np.pad(a, 1, mode="edge")
        return sum(p[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)) / 9.0

    mx, my = box(x), box(y)
    sx = np.maximum(box(x * x) - mx * mx, 0.0)
    sy = np.maximum(box(y * y) - my * my, 0.0)
    sxy = box(x * y) - mx * my
    return ((2 * mx * my + C1) * (2 * sxy + C2)) / np.maximum((mx * mx + my * my + C1) * (sx + sy + C2), 1e-8)


def _images(w, h, seed):
    t = synth.make_target(w, h, seed, n_ellipses=15)
    rng = np.random.default_rng(seed)
    r = np.clip(t + rng.normal(0, 0.05, t.shape), 0, 1).astype(np.float32)
    return r, t


def test_ssim_metric_matches_float64(orc):
    r, t = _images(70, 45, 1)
    s, per = orc.ssim_sum(r, t)
    ref = ssim64(r, t)
    # fp32 E[x^2] - mu^2 cancels: ~3e-8 absolute in a variance against C2 = 9e-4
    assert np.max(np.abs(per - ref)) <= 2e-4 and np.mean(np.abs(per - ref)) <= 3e-5
    assert abs(s - ref.sum()) <= 3e-5 * ref.size
    s1, per1 = orc.ssim_sum(t, t)
    assert np.max(np.abs(per1 - 1.0)) <= 1e-4                 # identical images
    m = np.ones((45, 70), np.float32)
    m[:, :20] = 0.0
    sm, _ = orc.ssim_sum(r, t, m)
    assert abs(sm - ref[:, 20:].sum()) <= 3e-5 * ref.size


def test_ssim_dcentre_matches_central_differences(orc):
    """dSSIM(window of k)/dx_k at every interior pixel (a border pixel repeats in
    its own clamped window, which the reference's formula does not count)."""
    r, t = _images(24, 20, 2)
    d = orc.ssim_dcentre(r, t)
    eps = 1e-4
    worst = 0.0
    for y in range(1, 19):
        for x in range(1, 23):
            sl = (slice(y - 1, y + 2), slice(x - 1, x + 2))
            rp, rm = r[sl].astype(np.float64), r[sl].astype(np.float64)
            rp[1, 1] += eps            # luma moves by eps * (0.299 + 0.587 + 0.114) = eps
            rm[1, 1] -= eps
            fd = (ssim64(rp, t[sl])[1, 1] - ssim64(rm, t[sl])[1, 1]) / (2 * eps)
            worst = max(worst, abs(d[y, x] - fd) / max(1.0, abs(fd)))
    assert worst <= 2e-3, worst


def test_ssim_gradient_is_l2_plus_backward(orc):
    """The site gradient is linear in dL/dpred: with the SSIM term it equals the
    L2 gradient plus the autograd backward of -w dSSIM/dx_k * luma weights."""
    w, h, n = 64, 48, 300
    sites = random_sites(w, h, n, 3)
    c0, c1 = converged_field(orc, sites, w, h, passes=8)
    iss = iss_of(w, h)
    t = synth.make_target(w, h, 3, n_ellipses=10)
    rendered = orc.render(c0, c1, sites, iss, w, h)
    mask = np.ones((h, w), np.float32)
    wgt = 0.7
    g_l2, ab_l2, _, _ = orc.gradients(c0, c1, t, mask, sites, iss, w, h)
    g_s, _, _, _ = orc.gradients(c0, c1, t, mask, sites, iss, w, h, rendered=rendered, ssim_weight=wgt)
    go = (-wgt * orc.ssim_dcentre(rendered, t).astype(np.float64))[..., None] * LUMA[None, None, :]
    g_b, ab_b = orc.render_backward(c0, c1, go.astype(np.float32), sites, iss, w, h)
    assert np.abs(g_b).max() > 1.0                       # the SSIM term is not negligible here
    assert Rel(g_s, g_l2 + g_b, ab_l2 + ab_b) <= 1.0
    assert Rel(g_s, g_l2 - g_b, ab_l2 + ab_b) > 100.0    # a flipped sign fails


# ---------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_ssim_metric(gpu, orc):
    r, t = _images(333, 211, 4)
    m = np.ones((211, 333), np.float32)
    m[50:90, 100:300] = 0.0
    acc = gpu.zeros((1,))
    gpu.call("sadh_compute_ssim", gpu.put(r), gpu.put(t), gpu.put(m), acc, 333, 211)
    s, _ = orc.ssim_sum(r, t, m)
    assert abs(float(gpu.get(acc)[0]) - s) <= 1e-5 * abs(s) + 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("removal", [False, True])
def test_gpu_gradients_with_ssim(gpu, orc, removal):
    w, h, n = 160, 128, 900
    sites = random_sites(w, h, n, 10)
    c0, c1 = converged_field(orc, sites, w, h, passes=12)
    iss = iss_of(w, h)
    rng = np.random.default_rng(11)
    t = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    mask[rng.random((h, w)) < 0.1] = 0.0
    rendered = orc.render(c0, c1, sites, iss, w, h)
    grads = [gpu.zeros((n,)) for _ in range(10)]
    rd = gpu.zeros((n,))
    gpu.call("sadh_compute_gradients_tiled", gpu.put(c0), gpu.put(c1), gpu.put(t), gpu.put(rendered), gpu.put(mask),
             *grads, gpu.put(sites), iss, n, rd, int(removal), 0.6, w, h, w, h)
    acc, ab, racc, rab = orc.gradients(c0, c1, t, mask, sites, iss, w, h, compute_removal=removal,
                                       rendered=rendered, ssim_weight=0.6)
    got = np.stack([gpu.get(g) for g in grads], 1)
    assert Rel(got, acc, ab) <= 1.0
    if removal:
        assert Rel(gpu.get(rd), racc, rab) <= 1.0


@pytest.mark.gpu
def test_gpu_fit_ssim(orc):
    from sad_amd import trainer
    w, h = 128, 96
    t = synth.make_target(w, h, 41, n_ellipses=20)
    s0 = synth.init_sites_uniform(t, 700, 41)
    opts = config.fit_options(iterations=120, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, max_sites=700, log_freq=40, ssim_weight=0.5, ssim_metric=1)
    g = trainer.fit(t, s0, opts)
    r = orc.fit(t, s0, opts)
    print("gpu", g["log_psnr"], g["log_ssim"], "oracle", r["log_psnr"], r["log_ssim"])
    assert abs(g["final_psnr"] - r["final_psnr"]) <= 0.05
    assert np.max(np.abs(g["log_ssim"] - r["log_ssim"])) <= 1e-3
    assert g["best_ssim"] == pytest.approx(float(np.max(g["log_ssim"])), abs=1e-6)
    # the SSIM term changes the fit: the same fit without it ends elsewhere
    plain = trainer.fit(t, s0, dict(opts, ssim_weight=0.0))
    assert abs(plain["final_psnr"] - g["final_psnr"]) > 1e-3


@pytest.mark.gpu
def test_gpu_batch_ssim_and_packed_equal_single_fits():
    """The batched engine with the SSIM loss, the SSIM metric and the packed
    PSNR: each job ends exactly where its own sadh_fit ends (the fits repeat
    bit for bit), with the same logged SSIMs, packed PSNR and renders."""
    from sad_amd import trainer
    specs = [(128, 96, 41, 700), (90, 140, 42, 500)]
    tg = [synth.make_target(w, h, sd, n_ellipses=20) for w, h, sd, _ in specs]
    st = [synth.init_sites_uniform(t, n, sd) for t, (_, _, sd, n) in zip(tg, specs)]
    opts = config.fit_options(iterations=90, target_bpp=1.5, densify_freq=30, prune_freq=30, densify_start=30,
                              prune_start=30, max_sites=1500, log_freq=30, ssim_weight=0.5, ssim_metric=1,
                              packed_psnr=1)
    b = trainer.fit_batch(tg, st, opts, want_render=True)["jobs"]
    for i, (t, s0) in enumerate(zip(tg, st)):
        g = trainer.fit(t, s0, dict(opts, n_sites=s0.shape[0]), want_render=True)
        print(i, "batch", b[i]["log_ssim"], b[i]["packed_psnr"], "single", g["log_ssim"], g["packed_psnr"])
        assert np.array_equal(b[i]["sites"].view(np.uint32), g["sites"].view(np.uint32))
        assert np.array_equal(b[i]["log_psnr"], g["log_psnr"]) and np.array_equal(b[i]["log_ssim"], g["log_ssim"])
        assert len(g["log_ssim"]) == 4 and 0.0 < g["final_ssim"] <= 1.0
        assert b[i]["packed_psnr"] == g["packed_psnr"] and b[i]["best_ssim"] == g["best_ssim"]
        assert np.array_equal(b[i]["render"], g["render"])
        assert np.array_equal(b[i]["packed_render"], g["packed_render"])
