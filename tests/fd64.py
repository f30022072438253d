"""FP64 closed-form softmax render and central differences over site
parameters, vectorised with numpy (render.cu:5-92 restated from its formula:
logit = -exp(log_tau) * (sqrt(max(d2 / S^2, 1e-8)) - r / S), softmax over the
valid candidates, colour blend; invalid or inactive candidates skipped).

Used to pin analytic gradients (oracle and GPU) independently of both."""
from __future__ import annotations

import numpy as np


def render_f64(c0: np.ndarray, c1: np.ndarray, sites: np.ndarray, w: int, h: int) -> np.ndarray:
    s = np.asarray(sites, np.float64)
    n = len(s)
    ids = np.concatenate([np.asarray(c0, np.uint32), np.asarray(c1, np.uint32)], 1).astype(np.int64)  # [P,8]
    valid = ids < n
    safe = np.where(valid, ids, 0)
    valid &= s[safe, 0] >= 0
    S = float(max(w, h))
    ys, xs = np.divmod(np.arange(w * h), w)
    dx = xs[:, None] - s[safe, 0]
    dy = ys[:, None] - s[safe, 1]
    ux, uy = s[safe, 7], s[safe, 8]
    proj = ux * dx + uy * dy
    perp2 = np.maximum(dx * dx + dy * dy - proj * proj, 0.0)
    la = s[safe, 9]
    d2 = np.maximum((np.exp(la) * proj * proj + np.exp(-la) * perp2) / (S * S), 1e-8)
    lg = -np.exp(s[safe, 2]) * (np.sqrt(d2) - s[safe, 3] / S)
    lg = np.where(valid, lg, -np.inf)
    mx = lg.max(1, keepdims=True)
    mx = np.where(np.isfinite(mx), mx, 0.0)
    e = np.where(valid, np.exp(lg - mx), 0.0)
    tot = e.sum(1, keepdims=True)
    wgt = np.where(tot > 0, e / np.where(tot > 0, tot, 1.0), 0.0)
    out = np.einsum("pk,pkc->pc", wgt, s[safe][:, :, 4:7])
    return out.reshape(h, w, 3)


def central_differences(loss, sites: np.ndarray, rows, cols, steps: dict) -> np.ndarray:
    """d loss / d sites[i, j] for i in rows, j in cols, by central differences in fp64."""
    base = np.asarray(sites, np.float64)
    out = np.zeros((len(rows), len(cols)))
    for a, i in enumerate(rows):
        for b, j in enumerate(cols):
            sp, sm = base.copy(), base.copy()
            sp[i, j] += steps[j]
            sm[i, j] -= steps[j]
            out[a, b] = (loss(sp) - loss(sm)) / (2 * steps[j])
    return out


FD_STEPS = {0: 1e-3, 1: 1e-3, 2: 1e-4, 3: 1e-3, 4: 1e-4, 5: 1e-4, 6: 1e-4, 7: 1e-4, 8: 1e-4, 9: 1e-4}


def fd_problem(seed: int = 11, w: int = 64, h: int = 64, n: int = 256):
    """The 64x64 / 256-site sub-problem of the reference's autograd demo
    (backends/pytorch/autograd_demo.py:16-29): random anisotropic sites with
    soft weights, |dir| = 0.999 (keeps the perp^2 >= 0 clamp inactive within
    the difference step), a random upstream gradient R, loss = sum(render * R)."""
    from tests.scenes import random_sites
    rng = np.random.default_rng(seed)
    sites = random_sites(w, h, n, seed, inactive_frac=0.0)
    sites[:, 2] = rng.uniform(1.5, 3.0, n).astype(np.float32)
    sites[:, 9] = rng.uniform(-0.8, 0.8, n).astype(np.float32)
    sites[:, 7:9] *= np.float32(0.999)
    grad_out = rng.uniform(-1, 1, (h, w, 3)).astype(np.float32)
    return sites, grad_out
