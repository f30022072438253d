"""Seeded synthetic workloads (targets and initial sites) of the fit and render
benchmarks.  numpy only; the same arrays feed the HIP path and the CPU oracle.

Target image: mean of 3 linear colour gradients (random direction, random
endpoint colours), then `n_ellipses` filled ellipses painted in order (centre
uniform, semi-axes U[0.02, 0.15] * min(W, H), angle U[0, pi), colour U[0,1]^3),
optionally a sinusoidal texture patch (`sinus_cycles` cycles across a W/2 x H/2
square at a seeded position), then N(0, 0.02^2) noise, clamped to [0, 1].
Layout float32 [H][W][3].
"""
from __future__ import annotations

import numpy as np

SITE_FIELDS = ("pos_x", "pos_y", "log_tau", "radius", "col_r", "col_g", "col_b", "dir_x", "dir_y", "log_aniso")


def make_target(width: int, height: int, seed: int, n_ellipses: int = 40, sinus_cycles: int = 0,
                noise: float = 0.02) -> np.ndarray:
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float32)
    img = np.zeros((height, width, 3), np.float32)
    for _ in range(3):
        th = rng.uniform(0.0, 2.0 * np.pi)
        c0 = rng.uniform(0.0, 1.0, 3).astype(np.float32)
        c1 = rng.uniform(0.0, 1.0, 3).astype(np.float32)
        u = (xx - width / 2) * np.cos(th) + (yy - height / 2) * np.sin(th)
        span = 0.5 * (abs(width * np.cos(th)) + abs(height * np.sin(th))) + 1e-6
        t = np.clip(0.5 + u / (2.0 * span), 0.0, 1.0)[..., None]
        img += ((1.0 - t) * c0 + t * c1) / 3.0
    m = min(width, height)
    for _ in range(n_ellipses):
        cx, cy = rng.uniform(0, width), rng.uniform(0, height)
        a, b = rng.uniform(0.02, 0.15, 2) * m
        ang = rng.uniform(0.0, np.pi)
        col = rng.uniform(0.0, 1.0, 3).astype(np.float32)
        r = max(a, b)
        x0, x1 = max(0, int(cx - r) - 1), min(width, int(cx + r) + 2)
        y0, y1 = max(0, int(cy - r) - 1), min(height, int(cy + r) + 2)
        if x0 >= x1 or y0 >= y1:
            continue
        dx = xx[y0:y1, x0:x1] - cx
        dy = yy[y0:y1, x0:x1] - cy
        ca, sa = np.cos(ang), np.sin(ang)
        u = (dx * ca + dy * sa) / a
        v = (-dx * sa + dy * ca) / b
        inside = (u * u + v * v) <= 1.0
        img[y0:y1, x0:x1][inside] = col
    if sinus_cycles > 0:
        pw, ph = width // 2, height // 2
        px, py = int(rng.integers(0, width - pw + 1)), int(rng.integers(0, height - ph + 1))
        th = rng.uniform(0.0, np.pi)
        col = rng.uniform(0.0, 1.0, 3).astype(np.float32)
        sx = xx[py:py + ph, px:px + pw] - px
        sy = yy[py:py + ph, px:px + pw] - py
        ph_ = 2.0 * np.pi * sinus_cycles * (sx * np.cos(th) + sy * np.sin(th)) / pw
        s = (0.5 + 0.5 * np.sin(ph_))[..., None]
        img[py:py + ph, px:px + pw] = s * col + (1.0 - s) * img[py:py + ph, px:px + pw]
    img += rng.normal(0.0, noise, img.shape).astype(np.float32)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def sites_array(n: int) -> np.ndarray:
    """Zeroed float32 [n][10] site array (Site field order)."""
    return np.zeros((n, 10), np.float32)


def init_sites_uniform(target: np.ndarray, n_sites: int, seed: int, log_tau: float = 5.0,
                       radius: float | None = None) -> np.ndarray:
    """Uniform positions in [0, W-1] x [0, H-1], isotropic metric (dir (1,0),
    log_aniso 0), colour sampled at (int y, int x) of the target, radius
    max(W,H)/sqrt(N) unless given (the 'log-scale = log(S/sqrt(N))' of the
    workload definition, mapped onto the reference's radius parameter)."""
    h, w, _ = target.shape
    rng = np.random.default_rng(seed + 1_000_003)
    s = sites_array(n_sites)
    s[:, 0] = rng.uniform(0.0, w - 1, n_sites)
    s[:, 1] = rng.uniform(0.0, h - 1, n_sites)
    s[:, 2] = log_tau
    s[:, 3] = radius if radius is not None else max(w, h) / np.sqrt(n_sites)
    s[:, 4:7] = target[s[:, 1].astype(np.int64), s[:, 0].astype(np.int64)]
    s[:, 7] = 1.0
    s[:, 8] = 0.0
    s[:, 9] = 0.0
    return s


def init_sites_random(width: int, height: int, n_sites: int, seed: int, log_tau: float = 5.0) -> np.ndarray:
    """Render/bandwidth workload sites: uniform positions, two log-scales
    l1, l2 ~ N(log 4, 0.5^2) mapped to radius exp((l1+l2)/2) and
    log_aniso clamp(l2 - l1, -4, 4), direction angle U[0, pi), colour U[0,1]^3."""
    rng = np.random.default_rng(seed)
    s = sites_array(n_sites)
    s[:, 0] = rng.uniform(0.0, width - 1, n_sites)
    s[:, 1] = rng.uniform(0.0, height - 1, n_sites)
    l1 = rng.normal(np.log(4.0), 0.5, n_sites)
    l2 = rng.normal(np.log(4.0), 0.5, n_sites)
    s[:, 2] = log_tau
    s[:, 3] = np.exp(0.5 * (l1 + l2))
    s[:, 4:7] = rng.uniform(0.0, 1.0, (n_sites, 3))
    ang = rng.uniform(0.0, np.pi, n_sites)
    s[:, 7] = np.cos(ang)
    s[:, 8] = np.sin(ang)
    s[:, 9] = np.clip(l2 - l1, -4.0, 4.0)
    return s.astype(np.float32)


WORKLOADS = {
    # plain single-image fit, runnable on CPU
    "fit256": dict(width=256, height=256, seed=1, n_ellipses=40, sinus_cycles=0, n_sites=2048, iterations=100),
    # full single-image fit with densify/prune
    "fit1024": dict(width=1024, height=1024, seed=2, n_ellipses=400, sinus_cycles=64, n_sites=16384,
                    iterations=2000),
}


def fit_workload(name: str):
    """(target, initial sites, options overrides) of a named fit workload."""
    w = WORKLOADS[name]
    tgt = make_target(w["width"], w["height"], w["seed"], w["n_ellipses"], w["sinus_cycles"])
    sites = init_sites_uniform(tgt, w["n_sites"], w["seed"])
    if name == "fit256":
        over = dict(iterations=w["iterations"], n_sites=w["n_sites"], target_bpp=0.0, densify_enabled=0,
                    densify_percentile=0.0, prune_percentile=0.0, max_sites=w["n_sites"])
    else:
        over = dict(iterations=w["iterations"], n_sites=w["n_sites"], target_bpp=1.0, densify_enabled=1,
                    densify_freq=100, prune_freq=100, max_sites=262144)
    return tgt, sites, over


def batch_workload(n_images: int = 64, seed: int = 4, min_side: int = 128, max_side: int = 2048,
                   bits_per_site: float = 128.0, bpp: float = 1.0):
    """The batched-fit workload: `n_images` seeded targets whose base side s is
    log-uniform in [min_side, max_side] and aspect ratio a log-uniform in
    [0.5, 2] (W = s*sqrt(a), H = s/sqrt(a), rounded), each with
    round(bpp * W * H / bits_per_site) uniform initial sites.  Returns
    (targets, sites)."""
    rng = np.random.default_rng(seed)
    targets, sites = [], []
    for i in range(n_images):
        side = float(np.exp(rng.uniform(np.log(min_side), np.log(max_side))))
        asp = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
        w = max(16, int(round(side * np.sqrt(asp))))
        h = max(16, int(round(side / np.sqrt(asp))))
        t = make_target(w, h, seed * 1000 + i, n_ellipses=40)
        n = max(16, int(round(bpp * w * h / bits_per_site)))
        targets.append(t)
        sites.append(init_sites_uniform(t, n, seed * 1000 + i))
    return targets, sites


def serve_workload(n_requests: int = 64, rate: float = 16.0, seed: int = 5, fit_fraction: float = 0.7,
                   fit_sides=(256, 2048), render_sides=(512, 4096), render_sites=(4096, 1 << 20)):
    """The served fit/render stream: Poisson arrivals at `rate` per second; a
    request is a fit with probability `fit_fraction` (side log-uniform in
    fit_sides, aspect log-uniform in [0.5, 2], bpp-1.0 initial sites:
    W*H/128) and otherwise a render of a random site set (site count and
    output side log-uniform in render_sites / render_sides, square output).
    Returns a list of dicts {t, kind, target|sites, width, height}."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() < fit_fraction:
            side = float(np.exp(rng.uniform(np.log(fit_sides[0]), np.log(fit_sides[1]))))
            asp = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
            w = max(16, int(round(side * np.sqrt(asp))))
            h = max(16, int(round(side / np.sqrt(asp))))
            tgt = make_target(w, h, seed * 7919 + i, n_ellipses=40)
            st = init_sites_uniform(tgt, max(16, int(round(w * h / 128.0))), seed * 7919 + i)
            out.append(dict(t=t, kind="fit", target=tgt, sites=st, width=w, height=h))
        else:
            side = int(round(float(np.exp(rng.uniform(np.log(render_sides[0]), np.log(render_sides[1]))))))
            n = int(round(float(np.exp(rng.uniform(np.log(render_sites[0]), np.log(render_sites[1]))))))
            out.append(dict(t=t, kind="render", sites=init_sites_random(side, side, n, seed * 7919 + i),
                            width=side, height=side))
    return out

