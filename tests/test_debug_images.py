"""The Metal CLI's debug pictures: the hashed-colour cell render
(renderVoronoiColoring, backends/metal/shaders/encoders/render_encoder.metal:73-134,
hashColor sad_common.metal:542-565) and the tau heatmap (renderCentroidsTauHeatmap,
render_encoder.metal:136-190), written by --render-cells and --include-debug-mask.

CPU tests pin the oracle: the id hash against a numpy restatement of the Metal
source in uint32 arithmetic; the cell picture as the render of the same sites
with their colours replaced by the hash colours; the heatmap on closed forms
(background grey, the three ramp anchors, the half-alpha dot edge, front-to-back
compositing).  GPU tests: both pictures bit-identical to the oracle, a fit's
pictures, and the command-line files."""
import numpy as np
import pytest

from sad_amd import cli, config, sites_io, synth
from tests.scenes import converged_field, iss_of, random_sites


def hash_np(ids):
    """hashColor of sad_common.metal:542-565, uint32 arithmetic"""
    with np.errstate(over="ignore"):
        i = np.asarray(ids, np.uint32)

        def mix(h, final_shift):
            h = (h ^ np.uint32(61)) ^ (h >> np.uint32(16))
            h = h + (h << np.uint32(3))
            if final_shift:
                h = h ^ (h >> np.uint32(4))
            return h * np.uint32(0x27D4EB2D)

        r = mix(i.copy(), True)
        g = mix(i * np.uint32(2654435761), False)
        h = i * np.uint32(1103515245)
        h = (h ^ np.uint32(61)) ^ (h >> np.uint32(16))
        h = h ^ (h >> np.uint32(4))
        b = h * np.uint32(0x27D4EB2D)
    return np.stack([(c & np.uint32(0xFFFFFF)).astype(np.float32) / np.float32(16777215.0) for c in (r, g, b)], 1)


def test_hash_colors_match_numpy(orc):
    ids = np.concatenate([np.arange(5000), np.array([0xFFFFFFFE, 1 << 31, 123456789])]).astype(np.uint32)
    got = orc.hash_colors(ids)
    assert np.array_equal(got, hash_np(ids))
    assert got.min() >= 0.0 and got.max() <= 1.0
    assert len({tuple(c) for c in got[:5000]}) == 5000            # distinct ids, distinct colours


def test_cells_are_the_render_with_hash_colours(orc):
    w, h, n = 120, 90, 400
    sites = random_sites(w, h, n, 5)
    c0, c1 = converged_field(orc, sites, w, h, passes=8)
    iss = iss_of(w, h)
    hs = sites.copy()
    hs[:, 4:7] = orc.hash_colors(np.arange(n))
    assert np.array_equal(orc.render_cells(c0, c1, sites, iss, w, h), orc.render(c0, c1, hs, iss, w, h))


def _one_site_field(w, h, ids):
    c0 = np.full((w * h, 4), 0xFFFFFFFF, np.uint32)
    c1 = np.full((w * h, 4), 0xFFFFFFFF, np.uint32)
    for k, i in enumerate(ids):
        c0[:, k] = i
    return c0, c1


def test_tau_heatmap_closed_forms(orc):
    w, h = 40, 30
    s = np.zeros((3, 10), np.float32)
    s[:, 7] = 1.0
    s[0, 0:2] = (10.0, 10.0); s[0, 2] = np.log(2.0)      # tau 2 = min -> blue
    s[1, 0:2] = (25.0, 10.0); s[1, 2] = np.log(5.0)      # tau 5 = mean -> white
    s[2, 0:2] = (20.0, 20.0); s[2, 2] = np.log(11.0)     # tau 11 = max -> red
    for i, want in ((0, (0, 0, 1)), (1, (1, 1, 1)), (2, (1, 0, 0))):
        c0, c1 = _one_site_field(w, h, [i])
        img = orc.tau_heatmap(c0, c1, s, 2.0, 5.0, 11.0, 2.0, w, h)
        x, y = int(s[i, 0]), int(s[i, 1])
        assert np.allclose(img[y, x], want, atol=1e-5)              # dot centre: alpha 1
        assert np.allclose(img[0, 39 if i == 0 else 0], 0.05)       # far away: the grey background
        edge = img[y, x + 2]                                        # distance 2 = the dot radius: alpha 1/2
        assert np.allclose(edge, 0.5 * np.array(want, np.float32) + 0.025, atol=1e-6)
    # the ramp between anchors: tau 3.5 is halfway from min to mean, colour channel 0.5^0.2
    s2 = s.copy(); s2[0, 2] = np.log(3.5)
    c0, c1 = _one_site_field(w, h, [0])
    img = orc.tau_heatmap(c0, c1, s2, 2.0, 5.0, 11.0, 2.0, w, h)
    assert np.allclose(img[10, 10], (0.5 ** 0.2, 0.5 ** 0.2, 1.0), atol=2e-6)
    # front to back: the first candidate's full dot hides the second
    c0, c1 = _one_site_field(w, h, [0, 2])
    s3 = s.copy(); s3[2, 0:2] = (10.0, 10.0)
    img = orc.tau_heatmap(c0, c1, s3, 2.0, 5.0, 11.0, 2.0, w, h)
    assert np.allclose(img[10, 10], (0, 0, 1), atol=1e-5)
    # inactive and out-of-range ids draw nothing
    s4 = s.copy(); s4[0, 0:2] = -1.0
    c0, c1 = _one_site_field(w, h, [0, 7])
    assert np.allclose(orc.tau_heatmap(c0, c1, s4, 2.0, 5.0, 11.0, 2.0, w, h), 0.05)


def test_cli_debug_flags():
    cfg = config.DEFAULT_CONFIG
    assert cli.parse_args(["x.png", "--include-debug-mask"], cfg)["debug_images"]
    assert cli.parse_args(["x.png", "--include-debug-masks"], cfg)["debug_images"]
    assert not cli.parse_args(["x.png"], cfg)["debug_images"]
    assert cli.parse_args(["--render", "s.txt", "--render-cells"], cfg)["render"]["cells"]
    assert cli.parse_args(["--render", "s.txt", "--render-voronoi"], cfg)["render"]["cells"]


# ---------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_cells_and_heatmap_bitexact(gpu, orc):
    w, h, n = 333, 211, 3000
    sites = random_sites(w, h, n, 6)
    c0, c1 = converged_field(orc, sites, w, h, passes=10)
    iss = iss_of(w, h)
    g0, g1, gs = gpu.put(c0), gpu.put(c1), gpu.put(sites)
    out = gpu.zeros((h, w, 3))
    gpu.call("sadh_render_cells", g0, g1, out, gs, iss, n, w, h, w, h)
    assert np.array_equal(gpu.get(out), orc.render_cells(c0, c1, sites, iss, w, h))
    act = sites[:, 0] >= 0
    tau = np.exp(sites[act, 2].astype(np.float64))
    tmin, tmean, tmax = float(tau.min()), float(tau.mean()), float(tau.max())
    gpu.call("sadh_render_tau_heatmap", g0, g1, out, gs, n, tmin, tmean, tmax, 2.0, w, h, w, h)
    ref = orc.tau_heatmap(c0, c1, sites, tmin, tmean, tmax, 2.0, w, h)
    assert np.array_equal(gpu.get(out), ref)
    assert np.mean(ref != np.float32(0.05)) > 0.01                # dots were drawn


@pytest.mark.gpu
def test_gpu_fit_and_cli_debug_images(tmp_path, capsys, orc):
    from sad_amd import trainer
    t = synth.make_target(96, 64, 33, n_ellipses=15)
    s0 = synth.init_sites_uniform(t, 500, 33)
    opts = config.fit_options(iterations=40, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, max_sites=500, log_freq=20)
    r = trainer.fit(t, s0, opts, want_render=True, debug_images=True)
    for img in (r["cells"], r["tau_heatmap"]):
        assert img.shape == (64, 96, 3) and img.min() >= 0.0 and img.max() <= 1.0 + 1e-6
    # the cell picture and the render share weights: where one site dominates a pixel, the
    # cell colour there is that site's hash colour
    assert not np.array_equal(r["cells"], r["render"])
    b = trainer.fit_batch([t], [s0], opts)["jobs"][0]
    assert np.array_equal(b["sites"].view(np.uint32), r["sites"].view(np.uint32))
    # the command line writes both pictures in training, and the cell picture in render mode
    img = tmp_path / "scene.png"
    sites_io.save_image(str(img), t)
    out = tmp_path / "o"
    assert cli.main([str(img), "--iters", "30", "--sites", "400", "--target-bpp", "0", "--out-dir", str(out),
                     "--include-debug-mask"]) == 0
    assert (out / "scene_cells.png").exists() and (out / "scene_tau_heatmap.png").exists()
    txt = out / "scene_sites.txt"
    assert cli.main(["--render", str(txt), "--out", str(tmp_path / "r.png"), "--render-cells"]) == 0
    cells = out / "scene_sites_cells.png"
    assert cells.exists()
    sites, w, h = sites_io.load_sites(str(txt))
    job = trainer.render_sites(sites, w, h, cand_passes=16, inject_count=config.DEFAULT_CONFIG["CAND_INJECT_COUNT"],
                               cells=True)
    assert np.array_equal(sites_io.load_image(str(cells)), sites_io.to_8bit(job["cells"]) / np.float32(255.0))
    capsys.readouterr()
