"""Sites/image files and the command line (sad_amd.sites_io, sad_amd.cli),
against the formats and flags of backends/cuda/src/sad_cuda.cu:137-420 and
:680-900.  CPU: file round trips, the reference's header lines and parsing
rules, argument parsing.  GPU: a fit and a render through the command line."""
import json
import os

import numpy as np
import pytest

from sad_amd import cli, config, sites_io, synth

HEADER = ("# SAD Sites (position_x, position_y, color_r, color_g, color_b, log_tau, radius, aniso_dir_x, "
          "aniso_dir_y, log_aniso)")


def _sites(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 10), np.float32)
    s[:, 0:2] = rng.uniform(0, 500, (n, 2))
    s[:, 2] = rng.normal(4, 2, n)
    s[:, 3] = rng.uniform(1, 30, n)
    s[:, 4:7] = rng.uniform(0, 1, (n, 3))
    a = rng.uniform(0, 2 * np.pi, n)
    s[:, 7], s[:, 8] = np.cos(a), np.sin(a)
    s[:, 9] = rng.uniform(-4, 4, n) * 10.0 ** rng.integers(-30, 3, n)
    return s


def test_sites_txt_round_trip_and_header(tmp_path):
    s = _sites(500, 1)
    p = tmp_path / "a_sites.txt"
    sites_io.write_sites_txt(s, 640, 480, str(p))
    lines = p.read_text().splitlines()
    assert lines[0] == HEADER and lines[1] == "# Image size: 640 480" and lines[2] == "# Total sites: 500"
    # column order: pos.x pos.y r g b log_tau radius dir.x dir.y log_aniso
    first = [float(v) for v in lines[3].split()]
    assert np.array_equal(np.float32(first), s[0, [0, 1, 4, 5, 6, 2, 3, 7, 8, 9]])
    back, w, h = sites_io.load_sites(str(p))
    assert (w, h) == (640, 480)
    assert np.array_equal(back.view(np.uint32), s.view(np.uint32))     # 9 significant digits round-trip fp32


def test_sites_txt_parsing_rules(tmp_path):
    p = tmp_path / "b.txt"
    p.write_text("\n".join([
        "  # a comment",
        "# IMAGE SIZE = 320x200 pixels",
        "1 2 0.1 0.2 0.3 5 4",                         # 7 columns: default direction and anisotropy
        "3 4 0.5 0.5 0.5 6 2 0 1 0.25",                 # 10 columns
        "5 6 0.5 0.5 0.5 6 2 extra words",             # reads 7 numbers, then stops
        "1 2 3",                                       # neither 7 nor 10: skipped
        "7 8 0.1 0.1 0.1 1 1 1 0",                     # 9: skipped
        "\t9 10 .5 1e-1 0.5 -2 3.5\r",
        "",
    ]))
    s, w, h = sites_io.load_sites(str(p))
    assert (w, h) == (320, 200)
    assert s.shape == (4, 10)
    assert s[0].tolist() == [1, 2, 5, 4, np.float32(0.1), np.float32(0.2), np.float32(0.3), 1, 0, 0]
    assert s[1].tolist() == [3, 4, 6, 2, 0.5, 0.5, 0.5, 0, 1, 0.25]
    assert s[2, 0:4].tolist() == [5, 6, 6, 2]
    assert s[3, 0:4].tolist() == [9, 10, -2, 3.5] and s[3, 4] == 0.5 and s[3, 5] == np.float32(0.1)
    # no size header, or a non-positive one: no size
    p.write_text("# image size: 0 100\n1 2 3 4 5 6 7\n")
    assert sites_io.load_sites(str(p))[1:] == (None, None)
    p.write_text("# nothing here\n")
    with pytest.raises(ValueError):
        sites_io.load_sites(str(p))


def test_sites_json(tmp_path):
    p = tmp_path / "c.JSON"
    p.write_text(json.dumps({"image_width": 50, "image_height": 40, "sites": [
        {"pos": [1, 2], "color": [0.1, 0.2, 0.3], "log_tau": 3.0, "radius": 5.0, "log_aniso": 0.5,
         "aniso_dir": [0.6, 0.8]},
        {"pos": [3, 4], "color": [1, 1, 1], "log_tau": 2.0, "radius_sq": 9.0},
        {"pos": [5, 6], "color": [0, 0, 0], "log_tau": 1.0, "aniso_dir": [0.0]},
    ]}))
    s, w, h = sites_io.load_sites(str(p))
    assert (w, h) == (50, 40)
    assert s[0].tolist() == [1, 2, 3, 5, np.float32(0.1), np.float32(0.2), np.float32(0.3), np.float32(0.6),
                             np.float32(0.8), 0.5]
    assert s[1].tolist() == [3, 4, 2, 9, 1, 1, 1, 1, 0, 0]
    assert s[2].tolist() == [5, 6, 1, 0, 0, 0, 0, 0, 0, 0]


def test_png_round_trip_and_rounding(tmp_path):
    rng = np.random.default_rng(2)
    k = rng.integers(0, 256, (37, 53, 3))
    img = (k.astype(np.float32) / np.float32(255.0))
    p = tmp_path / "x.png"
    sites_io.save_image(str(p), img)
    back = sites_io.load_image(str(p))
    assert back.shape == (37, 53, 3) and np.array_equal(back, img)
    # clamp, and halves away from zero on the float32 product (std::round(c * 255.0f))
    v = np.array([-0.2, 1.7, 0.5], np.float32)
    assert sites_io.to_8bit(v).tolist() == [0, 255, 128]
    ties = [np.float32(j + 0.5) / np.float32(255.0) for j in range(0, 255, 2)]
    ties = [t for t in ties if np.float32(t * np.float32(255.0)) == np.float32(np.floor(t * 255.0) + 0.5)]
    assert ties, "no float32 tie found"
    for t in ties:
        j = int(np.floor(np.float32(t * np.float32(255.0))))
        assert sites_io.to_8bit(np.array([t], np.float32))[0] == j + 1      # np.round would give the even j
    # masks: grey / 255
    from PIL import Image
    g = rng.integers(0, 256, (20, 30)).astype(np.uint8)
    Image.fromarray(g, mode="L").save(str(tmp_path / "m.png"))
    m = sites_io.load_mask(str(tmp_path / "m.png"), 30, 20)
    assert np.array_equal(m, g.astype(np.float32) / np.float32(255.0))


def test_active_sites_filter():
    s = _sites(10, 3)
    s[[1, 4], 0] = -1.0
    assert sites_io.active_sites(s).shape[0] == 8


def test_cli_parse_defaults_and_flags():
    cfg = config.DEFAULT_CONFIG
    r = cli.parse_args(["img.png"], cfg)
    o = r["opts"]
    assert r["mode"] == "train" and r["target"] == "img.png" and r["out_dir"] == "results"
    assert o["init_mode"] == 2 and o["n_sites"] == cfg["DEFAULT_SITES"] and o["target_bpp"] == cfg["DEFAULT_TARGET_BPP"]
    assert o["iterations"] == cfg["DEFAULT_ITERS"] and o["densify_enabled"] == 1
    r = cli.parse_args(["--sites", "5000", "--iters", "300", "x.png", "--target-bpp", "0.5", "--lr", "2",
                        "--cand-inject", "7", "--cand-hilbert-probes", "3", "--log-freq", "0", "--init-per-pixel",
                        "--prune-during-densify", "--densify-percentile", "0.02", "--out-dir", "o"], cfg)
    o = r["opts"]
    assert (o["n_sites"], o["iterations"], o["target_bpp"], o["lr_scale"], o["cand_inject_count"],
            o["cand_hilbert_probes"], o["log_freq"], o["init_mode"], o["prune_during_densify"]) == \
           (5000, 300, 0.5, 2.0, 7, 3, 1, 1, 1)
    assert abs(o["densify_percentile"] - 0.02) < 1e-9 and r["out_dir"] == "o"
    r = cli.parse_args(["x.png", "--init-from-sites", "s.txt"], cfg)
    assert r["opts"]["init_mode"] == 0 and r["init_sites"] == "s.txt"
    r = cli.parse_args(["--render", "s.txt", "--out", "o.png", "--render-no-jfa", "--render-cand-passes", "4",
                        "--width", "64", "--height", "32"], cfg)
    rr = r["render"]
    assert r["mode"] == "render" and (rr["sites"], rr["out"], rr["use_jfa"], rr["cand_passes"], rr["width"],
                                      rr["height"]) == ("s.txt", "o.png", False, 4, 64, 32)
    for bad in (["a.png", "b.png"], ["--bogus"], [], ["--render"], ["x.png", "--iters"], ["x.png", "--iters", "z"]):
        with pytest.raises(cli.UsageError):
            cli.parse_args(bad, cfg)
    assert cli.parse_args(["-h"], cfg)["mode"] == "help"
    o = cli.parse_args(["x.png", "--ssim", "--ssim-weight", "0.25", "--packed-psnr"], cfg)["opts"]
    assert o["ssim_metric"] == 1 and o["ssim_weight"] == 0.25 and o["packed_psnr"] == 1
    assert cli.parse_args(["--render", "s.txt", "--render-packed"], cfg)["render"]["packed"]
    with pytest.raises(cli.UsageError):                 # only after --render (cli.swift:442-450)
        cli.parse_args(["--render-packed", "--render", "s.txt"], cfg)
    assert cli.main(["--bogus"]) == 1


# ---------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_cli_fit_and_render(tmp_path, capsys):
    from sad_amd import trainer
    print("cli fit 96x64, 60 iterations", flush=True)
    t = synth.make_target(96, 64, 31, n_ellipses=15)
    img = tmp_path / "scene.png"
    sites_io.save_image(str(img), t)
    out = tmp_path / "out"
    rc = cli.main([str(img), "--iters", "60", "--sites", "600", "--target-bpp", "1.5", "--log-freq", "20",
                   "--out-dir", str(out), "--ssim"])
    assert rc == 0
    log = capsys.readouterr().out
    print(log)
    assert "Final PSNR" in log and "init=gradient" in log and "Final SSIM" in log
    png, txt = out / "scene.png", out / "scene_sites.txt"
    assert png.exists() and txt.exists()
    sites, w, h = sites_io.load_sites(str(txt))
    assert (w, h) == (96, 64) and np.all(sites[:, 0] >= 0)
    # the same fit through the library ends with the same bits (fits repeat exactly; the sites file's %.9g
    # round-trips float32), and the 8-bit PNG's PSNR is within 0.05 dB of the final PSNR
    tq = sites_io.load_image(str(img))
    final = float(log.split("Final PSNR: ")[1].split()[0])
    o = cli.parse_args([str(img), "--iters", "60", "--sites", "600", "--target-bpp", "1.5", "--log-freq", "20"],
                       config.DEFAULT_CONFIG)["opts"]
    res = trainer.fit(tq, None, o, want_render=True)
    ref = sites_io.active_sites(res["sites"][: res["actual_sites"]])
    assert sites.shape == ref.shape and np.array_equal(sites.view(np.uint32), ref.view(np.uint32))
    assert abs(trainer.psnr(sites_io.load_image(str(png)), tq) - final) <= 0.05
    assert abs(res["final_psnr"] - final) <= 5e-4
    assert abs(trainer.psnr(res["render"], tq) - res["final_psnr"]) <= 0.01
    # render mode on the written sites file, PSNR against the source image
    rpng = tmp_path / "r.png"
    rc = cli.main(["--render", str(txt), "--out", str(rpng), "--render-target", str(img)])
    assert rc == 0
    log = capsys.readouterr().out
    assert "Render PSNR" in log and rpng.exists()
    job = trainer.render_sites(sites, 96, 64, cand_passes=16, inject_count=config.DEFAULT_CONFIG["CAND_INJECT_COUNT"])
    assert np.array_equal(sites_io.load_image(str(rpng)), sites_io.to_8bit(job["image"]) / np.float32(255.0))
    psnr = float(log.split("Render PSNR: ")[1].split()[0])
    assert psnr > 20.0
    # the 16-byte quantised sites rendered by the same job
    ppng = tmp_path / "p.png"
    rc = cli.main(["--render", str(txt), "--render-packed", "--out", str(ppng), "--render-target", str(img)])
    assert rc == 0 and ppng.exists()
    plog = capsys.readouterr().out
    assert float(plog.split("Render PSNR: ")[1].split()[0]) > psnr - 1.0
    # a fit restarted from the written sites file
    rc = cli.main([str(img), "--iters", "10", "--init-from-sites", str(txt), "--out-dir", str(tmp_path / "o2"),
                   "--packed-psnr"])
    assert rc == 0 and (tmp_path / "o2" / "scene_sites.txt").exists() and (tmp_path / "o2" / "scene_packed.png").exists()
    log = capsys.readouterr().out
    assert "init=sites" in log and "Packed inference PSNR" in log
    os.remove(str(rpng))
