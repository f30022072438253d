"""Training configuration: the reference's training_config.json keys and the
fit options derived from them (backends/cuda/src/sad_cuda.cu:459-517,
makeTrainingOptions :654-695).

`load_config(path)` reads a training_config.json (the path in $CONFIG_PATH when
none is given, else the built-in defaults below, which equal the reference's
training_config.json:2-35).
"""
from __future__ import annotations

import json
import os

from ._lib import FitOptions

DEFAULT_CONFIG = {
    "DEFAULT_SITES": 128000, "DEFAULT_MAX_SITES": 128000, "DEFAULT_ITERS": 4000, "DEFAULT_TARGET_BPP": 1.0,
    "PRUNE_PERCENTILE": 0.033, "PRUNE_DURING_DENSIFY": True, "PRUNE_START": 100, "PRUNE_FREQ": 40,
    "PRUNE_END": 3000, "DENSIFY_START": 20, "DENSIFY_FREQ": 20, "DENSIFY_END": 3000,
    "DENSIFY_PERCENTILE": 0.01, "DENSIFY_SCORE_ALPHA": 0.7, "CAND_UPDATE_FREQ": 8, "CAND_UPDATE_PASSES": 1,
    "CAND_RADIUS_SCALE": 64.0, "CAND_RADIUS_PROBES": 0, "CAND_DOWNSCALE": 1, "CAND_INJECT_COUNT": 4,
    "CAND_HILBERT_WINDOW": 32, "CAND_HILBERT_PROBES": 1, "LR_POS_BASE": 0.025, "LR_TAU_BASE": 0.01,
    "LR_RADIUS_BASE": 0.01, "LR_COLOR_BASE": 0.01, "LR_DIR_BASE": 0.0025, "LR_ANISO_BASE": 0.025,
    "INIT_LOG_TAU": 5.0, "INIT_RADIUS": 4.0, "BETA1": 0.9, "BETA2": 0.999, "EPS": 1e-8, "MAX_DIM": 12000,
}


def load_config(path: str | None = None) -> dict:
    cfg = dict(DEFAULT_CONFIG)
    path = path or os.environ.get("CONFIG_PATH")
    if path:
        with open(path) as f:
            cfg.update(json.load(f))
    return cfg


def fit_options(cfg: dict | None = None, **over) -> dict:
    """Fit options as a dict, from a config dict plus keyword overrides.

    Keys are the SadhFitOptions field names.  Densify is on as in the reference
    CUDA host (sad_cuda.cu:609); it is effectively off when both percentiles
    are 0.
    """
    c = cfg or DEFAULT_CONFIG
    o = dict(
        iterations=c["DEFAULT_ITERS"], target_bpp=c["DEFAULT_TARGET_BPP"], init_mode=0,
        n_sites=c["DEFAULT_SITES"], max_sites=c["DEFAULT_MAX_SITES"],
        prune_percentile=c["PRUNE_PERCENTILE"], prune_start=c["PRUNE_START"], prune_freq=c["PRUNE_FREQ"],
        prune_end=c["PRUNE_END"], prune_during_densify=int(bool(c["PRUNE_DURING_DENSIFY"])), densify_enabled=1,
        densify_start=c["DENSIFY_START"], densify_freq=c["DENSIFY_FREQ"], densify_end=c["DENSIFY_END"],
        densify_percentile=c["DENSIFY_PERCENTILE"], densify_score_alpha=c["DENSIFY_SCORE_ALPHA"],
        cand_update_freq=c["CAND_UPDATE_FREQ"], cand_update_passes=c["CAND_UPDATE_PASSES"],
        cand_downscale=c["CAND_DOWNSCALE"], cand_radius_scale=c["CAND_RADIUS_SCALE"],
        cand_radius_probes=c["CAND_RADIUS_PROBES"], cand_inject_count=c["CAND_INJECT_COUNT"],
        cand_hilbert_window=c["CAND_HILBERT_WINDOW"], cand_hilbert_probes=c["CAND_HILBERT_PROBES"],
        lr_scale=1.0, lr_pos=c["LR_POS_BASE"], lr_tau=c["LR_TAU_BASE"], lr_radius=c["LR_RADIUS_BASE"],
        lr_color=c["LR_COLOR_BASE"], lr_dir=c["LR_DIR_BASE"], lr_aniso=c["LR_ANISO_BASE"],
        beta1=c["BETA1"], beta2=c["BETA2"], eps=c["EPS"], log_freq=1000, profile_gradients=0,
        init_gradient_alpha=1.0, init_log_tau=c["INIT_LOG_TAU"], init_radius=c["INIT_RADIUS"],
        ssim_weight=0.0, ssim_metric=0, packed_psnr=0, graph_ticks=0, autotune=0,
    )
    for k, v in over.items():
        if k not in o:
            raise KeyError(f"unknown fit option {k!r}")
        o[k] = v
    return o


def to_struct(opts: dict) -> FitOptions:
    s = FitOptions()
    for k, v in opts.items():
        setattr(s, k, v)
    return s
