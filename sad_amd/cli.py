"""Command line of the fit and render jobs, with the flags of the reference's
`sad_cuda` binary (backends/cuda/src/sad_cuda.cu:680-900, main :2528-2540):

    python -m sad_amd.cli <image_path> [options]        fit an image
    python -m sad_amd.cli --render <sites.txt> [options] render a sites file

Defaults come from training_config.json in the working directory, else the
file named by $CONFIG_PATH, else the built-in copy of the reference's file.
A fit writes <out-dir>/<stem>.png (the final render) and
<out-dir>/<stem>_sites.txt (the active sites); a render writes --out, else
<sites path without extension>_render.png.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from . import config, sites_io

_INT = {
    "--sites": "n_sites", "--iters": "iterations", "--densify-start": "densify_start",
    "--densify-end": "densify_end", "--densify-freq": "densify_freq", "--prune-start": "prune_start",
    "--prune-end": "prune_end", "--prune-freq": "prune_freq", "--cand-freq": "cand_update_freq",
    "--cand-passes": "cand_update_passes", "--cand-downscale": "cand_downscale",
    "--cand-radius-probes": "cand_radius_probes", "--cand-inject-count": "cand_inject_count",
    "--cand-inject": "cand_inject_count", "--cand-hilbert-window": "cand_hilbert_window",
    "--cand-hilbert-probes": "cand_hilbert_probes", "--max-sites": "max_sites",
}
_FLOAT = {
    "--densify-percentile": "densify_percentile", "--densify-score-alpha": "densify_score_alpha",
    "--prune-percentile": "prune_percentile", "--cand-radius-scale": "cand_radius_scale", "--lr": "lr_scale",
    "--init-gradient-alpha": "init_gradient_alpha",
}
_RENDER_INT = {
    "--width": "width", "--height": "height", "--render-cand-passes": "cand_passes",
    "--render-hilbert-window": "hilbert_window", "--render-hilbert-probes": "hilbert_probes",
    "--render-cand-inject": "inject_count", "--render-jfa-rounds": "jfa_rounds",
    "--render-cand-downscale": "cand_downscale",
}

USAGE = """Usage:
  python -m sad_amd.cli <image_path> [options]
  python -m sad_amd.cli --render <sites.txt> [options]

Common options:
  --help, -h                Show this help
  --out-dir <path>          Output directory (default: results)
Training options:
  --sites N  --iters N  --target-bpp F  --mask <path>
  --init-per-pixel | --init-gradient (default) | --init-from-sites <path>  --init-gradient-alpha F
  --densify  --densify-start/-end/-freq N  --densify-percentile F  --densify-score-alpha F
  --prune-during-densify  --prune-start/-end/-freq N  --prune-percentile F
  --cand-freq N  --cand-passes N  --cand-downscale N  --cand-radius-scale F  --cand-radius-probes N
  --cand-inject N  --cand-hilbert-window N  --cand-hilbert-probes N  --max-sites N  --lr F  --log-freq N
  --ssim (log SSIM)  --ssim-weight F (loss L2 - F * SSIM)  --packed-psnr (re-render 16-byte sites, PSNR)
  --include-debug-mask (also write <image>_cells.png and <image>_tau_heatmap.png)
Render options:
  --out <path>  --width N  --height N  --render-cand-passes N  --render-no-jfa  --render-jfa-rounds N
  --render-cand-downscale N  --render-hilbert-window N  --render-hilbert-probes N  --render-cand-inject N
  --render-target <path>  --render-packed (render the 16-byte quantised sites)
  --render-cells (also write <sites>_cells.png, hashed cell colours)
"""


class UsageError(Exception):
    pass


def _load_defaults() -> dict:
    return config.load_config("training_config.json" if os.path.exists("training_config.json") else None)


def parse_args(argv: list, cfg: dict) -> dict:
    """parseArgs, sad_cuda.cu:753-900: later flags override earlier ones;
    unknown flags and a second positional argument are errors."""
    o = config.fit_options(cfg)
    o["init_mode"] = 2
    if cfg["DEFAULT_TARGET_BPP"] <= 0:
        o["target_bpp"] = 0.0
    run = dict(mode="train", opts=o, target=None, mask=None, init_sites=None, out_dir="results",
               max_dim=int(cfg["MAX_DIM"]), debug_images=False,
               render=dict(sites=None, out=None, target=None, width=0, height=0, cand_passes=16, use_jfa=True,
                           packed=False, cells=False,
                           jfa_rounds=1, cand_downscale=int(cfg["CAND_DOWNSCALE"]),
                           radius_scale=float(cfg["CAND_RADIUS_SCALE"]),
                           radius_probes=int(cfg["CAND_RADIUS_PROBES"]), inject_count=int(cfg["CAND_INJECT_COUNT"]),
                           hilbert_window=int(cfg["CAND_HILBERT_WINDOW"]),
                           hilbert_probes=int(cfg["CAND_HILBERT_PROBES"])))
    r = run["render"]
    i = 0

    def value(flag):
        nonlocal i
        if i + 1 >= len(argv):
            raise UsageError(f"Missing value for {flag}")
        i += 1
        return argv[i]

    while i < len(argv):
        a = argv[i]
        try:
            if a in ("--help", "-h"):
                run["mode"] = "help"
                return run
            if a == "--render":
                run["mode"] = "render"
                r["sites"] = value(a)
            elif a.startswith("--"):
                if a in _INT:
                    o[_INT[a]] = int(value(a))
                elif a in _FLOAT:
                    o[_FLOAT[a]] = float(value(a))
                elif a in _RENDER_INT:
                    r[_RENDER_INT[a]] = int(value(a))
                elif a == "--target-bpp":
                    o["target_bpp"] = float(value(a))
                elif a == "--mask":
                    run["mask"] = value(a)
                elif a == "--init-per-pixel":
                    o["init_mode"] = 1
                elif a == "--init-gradient":
                    o["init_mode"] = 2
                elif a == "--init-from-sites":
                    o["init_mode"] = 0
                    run["init_sites"] = value(a)
                elif a == "--densify":
                    o["densify_enabled"] = 1
                elif a == "--prune-during-densify":
                    o["prune_during_densify"] = 1
                elif a == "--ssim":
                    o["ssim_metric"] = 1
                elif a == "--ssim-weight":
                    o["ssim_weight"] = float(value(a))
                elif a == "--log-freq":
                    o["log_freq"] = max(1, int(value(a)))
                elif a == "--out-dir":
                    run["out_dir"] = value(a)
                elif a == "--out":
                    r["out"] = value(a)
                elif a == "--render-target":
                    r["target"] = value(a)
                elif a == "--render-no-jfa":
                    r["use_jfa"] = False
                elif a == "--render-packed":
                    if run["mode"] != "render":
                        raise UsageError("--render-packed is only valid in render mode")
                    r["packed"] = True
                elif a == "--packed-psnr":
                    o["packed_psnr"] = 1
                elif a in ("--render-cells", "--render-voronoi"):
                    r["cells"] = True
                elif a in ("--include-debug-mask", "--include-debug-masks"):
                    run["debug_images"] = True
                else:
                    raise UsageError(f"Unknown flag: {a}")
            elif run["mode"] == "train":
                if run["target"] is not None:
                    raise UsageError(f"Unexpected argument: {a}")
                run["target"] = a
        except ValueError:
            raise UsageError(f"Bad value for {a}") from None
        i += 1
    if run["mode"] == "render":
        if not r["sites"]:
            raise UsageError("--render requires a sites path")
    elif run["target"] is None:
        raise UsageError("Missing input image path")
    return run


def _per_pixel_sites(target: np.ndarray, log_tau: float, radius: float) -> np.ndarray:
    """sad_cuda.cu:1923-1941: one site per pixel at its integer coordinates."""
    h, w, _ = target.shape
    s = np.zeros((h * w, 10), np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    s[:, 0], s[:, 1] = xx.ravel(), yy.ravel()
    s[:, 2], s[:, 3] = log_tau, radius
    s[:, 4:7] = target.reshape(-1, 3)
    s[:, 7] = 1.0
    return s


def train(run: dict, out=None) -> dict:
    """trainVoronoi, sad_cuda.cu:1860-2526, on the native fit (sadh_fit)."""
    from . import trainer
    out = out or sys.stdout
    o = dict(run["opts"])
    t_start = time.perf_counter()
    target = sites_io.load_image(run["target"], run["max_dim"])
    h, w, _ = target.shape
    mask = sites_io.load_mask(run["mask"], w, h) if run["mask"] else None
    init = None
    if o["init_mode"] == 0:
        init, sw, sh = sites_io.load_sites(run["init_sites"])
        if sw and sh and (sw != w or sh != h):
            init[:, 0] *= np.float32(w) / np.float32(sw)
            init[:, 1] *= np.float32(h) / np.float32(sh)
        o["n_sites"] = init.shape[0]
    elif o["init_mode"] == 1:
        init = _per_pixel_sites(target, o["init_log_tau"], o["init_radius"])
    plan = trainer.fit_plan(o, w, h)
    mode = {0: "sites", 1: "per-pixel", 2: "gradient"}[o["init_mode"]]
    print(f"Training | backend=hip | image={w}x{h} | sites={o['n_sites'] if init is None else init.shape[0]} | "
          f"iters={o['iterations']} | log-freq={o['log_freq']} | mask={'yes' if mask is not None else 'no'} | "
          f"out={run['out_dir']}" + (f" | target-bpp={o['target_bpp']:.3f}" if o["target_bpp"] > 0 else ""),
          file=out)
    print(f"Schedule | init={mode} | densify={'on cap=' + str(plan['capacity']) if o['densify_enabled'] else 'off'}"
          f" | prune={'on %.3f' % plan['prune_percentile'] if plan['prune_percentile'] > 0 else 'off'}"
          f" | cand=freq {o['cand_update_freq']}, passes {o['cand_update_passes']}, "
          f"downscale {o['cand_downscale']}x", file=out)
    res = trainer.fit(target, init, o, mask=mask, want_render=True, debug_images=run["debug_images"])
    logs = list(res["log_psnr"])
    its = [k for k in range(o["iterations"]) if k % max(1, o["log_freq"]) == 0 or k == o["iterations"] - 1]
    ssims = list(res["log_ssim"]) if o["ssim_metric"] else [None] * len(logs)
    print("Logs | Iter | PSNR" + (" | SSIM" if o["ssim_metric"] else ""), file=out)
    for k, p, q in zip(its, logs, ssims):
        print(f"Logs | {k} | {p:.2f}" + (f" | {q:.4f}" if q is not None else ""), file=out)
    os.makedirs(run["out_dir"], exist_ok=True)
    stem = os.path.splitext(os.path.basename(run["target"]))[0]
    img_path = os.path.join(run["out_dir"], stem + ".png")
    sites_path = os.path.join(run["out_dir"], stem + "_sites.txt")
    sites_io.save_image(img_path, res["render"])
    print(f"Saved: {img_path}", file=out)
    final = sites_io.active_sites(res["sites"][: res["actual_sites"]])
    sites_io.write_sites_txt(final, w, h, sites_path)
    print(f"Saved: {sites_path}", file=out)
    if o["packed_psnr"]:
        packed_path = os.path.join(run["out_dir"], stem + "_packed.png")
        sites_io.save_image(packed_path, res["packed_render"])
        print(f"Saved: {packed_path}", file=out)
        print(f"Packed inference PSNR: {res['packed_psnr']:.2f} dB", file=out)
    if run["debug_images"]:
        # training.swift:665-671
        for suffix, img in (("_cells.png", res["cells"]), ("_tau_heatmap.png", res["tau_heatmap"])):
            path = os.path.join(run["out_dir"], stem + suffix)
            sites_io.save_image(path, img)
            print(f"Saved: {path}", file=out)
    print(f"Final PSNR: {res['final_psnr']} dB (best {res['best_psnr']})", file=out)
    if o["ssim_metric"]:
        print(f"Final SSIM: {res['final_ssim']} (best {res['best_ssim']})", file=out)
    print(f"Training time: {res['train_seconds']:.2f} s", file=out)
    print(f"Total time: {time.perf_counter() - t_start:.2f} s", file=out)
    return dict(result=res, image_path=img_path, sites_path=sites_path, n_active=final.shape[0])


def render(run: dict, out=None) -> dict:
    """renderVoronoiFromSites, sad_cuda.cu:1678-1858."""
    from . import trainer
    out = out or sys.stdout
    r = run["render"]
    sites, sw, sh = sites_io.load_sites(r["sites"])
    w, h = (sw, sh) if sw and sh else (r["width"], r["height"])
    if w <= 0 or h <= 0:
        raise UsageError("TXT sites require --width and --height or a header line with image size.")
    print(f"Loaded {sites.shape[0]} sites", file=out)
    print(f"Render size: {w}x{h}", file=out)
    kw = dict(cand_passes=max(0, r["cand_passes"]), use_jfa=r["use_jfa"], jfa_rounds=r["jfa_rounds"],
              cand_downscale=r["cand_downscale"], radius_scale=r["radius_scale"], radius_probes=r["radius_probes"],
              inject_count=r["inject_count"], hilbert_window=r["hilbert_window"], hilbert_probes=r["hilbert_probes"],
              cells=r["cells"])
    if r["packed"]:
        packed, quant = trainer.pack_inference(sites, w, h)
        job = trainer.render_packed_sites(packed, quant, w, h, **kw)
    else:
        job = trainer.render_sites(sites, w, h, **kw)
    psnr = None
    if r["target"]:
        t = sites_io.load_image(r["target"], max(w, h))
        if t.shape[:2] != (h, w):
            print("Render target size mismatch; skipping PSNR.", file=sys.stderr)
        else:
            d = (job["image"] - t).astype(np.float32)
            mse = float(np.sum((d * d).astype(np.float64))) / (max(1.0, float(w * h)) * 3.0)
            psnr = 100.0 if mse <= 0 else 20.0 * np.log10(1.0 / np.sqrt(mse))
            print(f"Render PSNR: {psnr:.2f} dB", file=out)
    print(f"Candidate build: {job['cand_ms']:.3f} ms | Render: {job['render_ms']:.3f} ms", file=out)
    path = r["out"] or (os.path.splitext(r["sites"])[0] + "_render.png")
    sites_io.save_image(path, job["image"])
    print(f"Saved: {path}", file=out)
    cells_path = None
    if r["cells"]:
        # cli.swift:833-836: <sites path without extension>_cells.png
        cells_path = os.path.splitext(r["sites"])[0] + "_cells.png"
        sites_io.save_image(cells_path, job["cells"])
        print(f"Saved: {cells_path}", file=out)
    return dict(image=job["image"], path=path, psnr=psnr, cells=job["cells"], cells_path=cells_path)


def main(argv: list | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    try:
        run = parse_args(argv, _load_defaults())
        if run["mode"] == "help":
            print(USAGE)
            return 0
        if run["mode"] == "render":
            render(run)
        else:
            train(run)
    except UsageError as e:
        print(e, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
