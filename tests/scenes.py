"""Seeded test scenes: sites, targets and candidate fields built with the CPU
oracle, shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

from sad_amd import synth


def same_bits(a, b) -> bool:
    """float32 arrays equal bit for bit"""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


SCHEDULE = ("n_densify", "n_prune", "actual_sites", "active_estimate", "n_cand_updates")


def assert_same_fit(g: dict, r: dict) -> None:
    """two fits (the GPU's, the oracle's or a repeat) ended with the same bits: sites, PSNR log, schedule"""
    assert tuple(g[k] for k in SCHEDULE) == tuple(r[k] for k in SCHEDULE)
    bad = int(np.count_nonzero(np.asarray(g["sites"], np.float32).view(np.uint32) !=
                               np.asarray(r["sites"], np.float32).view(np.uint32)))
    assert bad == 0, f"{bad} site values differ"
    assert same_bits(g["log_psnr"], r["log_psnr"]), (g["log_psnr"], r["log_psnr"])


def iss_of(w, h):
    m = np.float32(max(w, h))
    return float(np.float32(1.0) / (m * m))


def random_sites(w, h, n, seed, inactive_frac=0.05, aniso=True):
    s = synth.init_sites_random(w, h, n, seed)
    rng = np.random.default_rng(seed + 7)
    s[:, 2] = rng.uniform(3.0, 6.0, n).astype(np.float32)
    if not aniso:
        s[:, 7], s[:, 8], s[:, 9] = 1.0, 0.0, 0.0
    k = int(n * inactive_frac)
    if k:
        idx = rng.choice(n, k, replace=False)
        s[idx, 0] = -1.0
        s[idx, 1] = -1.0
    return s.astype(np.float32)


def converged_field(orc, sites, w, h, passes=12, inject=4, hil=True, seed=0):
    """init + seed + `passes` VPT passes (hilbert probes on when hil)."""
    n = sites.shape[0]
    iss = iss_of(w, h)
    c0, c1 = orc.init_candidates(n, seed, False, w, h)
    c0 = orc.jfa_seed(c0, sites, w, h)
    pk = orc.pack_sites(sites)
    ho = hp = None
    if hil:
        ho, hp = hilbert_tables(orc, sites, w, h)
    for i in range(passes):
        step = (orc.jump_step(i, w, h) << 16) | (i & 0xFFFF)
        c0, c1 = orc.update_candidates(c0, c1, pk, step, iss, w, h, inject_count=inject, hilbert_order=ho,
                                       hilbert_pos=hp, hilbert_probes=1 if hil else 0,
                                       hilbert_window=32 if hil else 0)
    return c0, c1


def hilbert_tables(orc, sites, w, h):
    n = sites.shape[0]
    bits = max(1, int(np.ceil(np.log2(max(w, h)))))
    padded = (n + 1023) // 1024 * 1024
    pairs = orc.hilbert_pairs(sites, padded, w, h, bits)
    mk = 0xFFFFFFFF if bits >= 16 else (1 << (2 * bits))
    srt = orc.radix_sort(pairs, mk)
    order = srt[:n, 1].copy()
    pos = np.zeros(n, np.uint32)
    pos[order] = np.arange(n, dtype=np.uint32)
    return order, pos
