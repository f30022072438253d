"""End-to-end GPU parity: the native fit loop against the oracle's restatement of
the reference schedule (final PSNR within 0.05 dB), and the render-only recipe
(JFA + VPT + render) against the oracle pipeline (image within 1e-5)."""
import numpy as np
import pytest

from sad_amd import config, synth, trainer
from tests.scenes import hilbert_tables, iss_of

pytestmark = pytest.mark.gpu


def test_fit256_psnr_parity(orc):
    tgt, sites, over = synth.fit_workload("fit256")
    opts = config.fit_options(**over)
    g = trainer.fit(tgt, sites, opts)
    r = orc.fit(tgt, sites, opts)
    print("gpu", g["log_psnr"], "oracle", r["log_psnr"])
    assert g["n_cand_updates"] == r["n_cand_updates"] == 13
    assert abs(g["final_psnr"] - r["final_psnr"]) <= 0.05
    assert g["final_psnr"] > g["log_psnr"][0] + 5.0


def test_fit_densify_prune_schedule(orc):
    """Small fit with the densify/prune machinery on (every 20 iterations)."""
    tgt = synth.make_target(192, 160, 31, n_ellipses=30)
    sites = synth.init_sites_uniform(tgt, 600, 31)
    opts = config.fit_options(iterations=200, target_bpp=1.0, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2400, log_freq=50)
    g = trainer.fit(tgt, sites, opts)
    r = orc.fit(tgt, sites, opts)
    print("gpu", g["log_psnr"], g["actual_sites"], "oracle", r["log_psnr"], r["actual_sites"])
    assert g["plan"] == r["plan"]
    assert (g["n_densify"], g["n_prune"], g["actual_sites"], g["active_estimate"]) == \
           (r["n_densify"], r["n_prune"], r["actual_sites"], r["active_estimate"])
    assert g["active_count"] == int(np.sum(g["sites"][:, 0] >= 0))
    assert abs(g["final_psnr"] - r["final_psnr"]) <= 0.05


def test_fit1024_opening_against_oracle(orc):
    """The full 1024x1024 / 16,384-site fit with its densify/prune plan, through
    the first densify and prune (iteration 100): schedule identical, PSNR within
    0.05 dB at every log.  The oracle's 120 iterations take a few seconds."""
    import os
    orc.set_threads(min(16, os.cpu_count() or 1))
    tgt, sites, over = synth.fit_workload("fit1024")
    opts = config.fit_options(**dict(over, iterations=120, log_freq=40))
    g = trainer.fit(tgt, sites, opts)
    r = orc.fit(tgt, sites, opts)
    print("gpu", g["log_psnr"], g["actual_sites"], "oracle", r["log_psnr"], r["actual_sites"])
    assert g["plan"] == r["plan"] and g["n_densify"] == r["n_densify"] == 1 and g["n_prune"] == r["n_prune"] == 1
    assert (g["actual_sites"], g["active_estimate"]) == (r["actual_sites"], r["active_estimate"])
    assert np.max(np.abs(np.asarray(g["log_psnr"]) - np.asarray(r["log_psnr"]))) <= 0.05


def test_render_sites_pipeline(orc):
    w, h, n = 200, 160, 3000
    sites = synth.init_sites_random(w, h, n, 41)
    got = trainer.render_sites(sites, w, h, cand_passes=16)
    iss = iss_of(w, h)
    # oracle: clear + seed + log2 floods (plane 1 as ping-pong), 16 VPT passes, render
    c0, c1 = orc.clear_candidates(w, h)
    c0 = orc.jfa_seed(c0, sites, w, h)
    step = 1
    while step < max(w, h):
        step <<= 1
    bufs, cur = [c0, c1], 0
    s = step // 2
    while s >= 1:
        bufs[1 - cur] = orc.jfa_flood(bufs[cur], sites, s, iss, w, h)
        cur = 1 - cur
        s //= 2
    if cur != 0:
        bufs[0] = bufs[cur].copy()
    c0, c1 = bufs
    pk = orc.pack_sites(sites)
    ho, hp = hilbert_tables(orc, sites, w, h)
    for i in range(16):
        stp = (orc.jump_step(i, w, h) << 16) | i
        c0, c1 = orc.update_candidates(c0, c1, pk, stp, iss, w, h, inject_count=4, hilbert_order=ho,
                                       hilbert_pos=hp, hilbert_probes=1, hilbert_window=32)
    ref = orc.render(c0, c1, sites, iss, w, h)
    assert np.abs(got["image"] - ref).max() <= 1e-5
