"""End-to-end GPU parity of the native fit loop against the oracle's
restatement of the reference schedule.

The library's fit is deterministic (its cross-tile sums are exact integers in
a fixed per-tile order) and the oracle forms the same sums, so the bar is
bit-identity: final sites, every logged PSNR and the densify/prune schedule.
The full 2,000-iteration 1024x1024 fit is checked against the oracle's result
stored in tests/golden/fit1024_full.npz.  The render-only recipe (JFA + VPT +
render) is held to the oracle pipeline within 1e-5."""
import os

import numpy as np
import pytest

from sad_amd import config, synth, trainer
from tests.golden import make_golden as mg
from tests.scenes import assert_same_fit, hilbert_tables, iss_of

pytestmark = pytest.mark.gpu


def test_fit256_bitexact(orc):
    tgt, sites, over = synth.fit_workload("fit256")
    opts = config.fit_options(**over)
    g = trainer.fit(tgt, sites, opts)
    r = orc.fit(tgt, sites, opts)
    print("gpu", g["log_psnr"], "oracle", r["log_psnr"])
    assert g["n_cand_updates"] == 13
    assert_same_fit(g, r)
    assert g["final_psnr"] > g["log_psnr"][0] + 5.0


@pytest.mark.parametrize("iterations", [1, 41, 200])
def test_fit_densify_prune_bitexact(orc, iterations):
    """Small fit with the densify/prune machinery on (every 20 iterations), at
    lengths that end before, between and after its densify/prune rounds."""
    tgt = synth.make_target(192, 160, 31, n_ellipses=30)
    sites = synth.init_sites_uniform(tgt, 600, 31)
    opts = config.fit_options(iterations=iterations, target_bpp=1.0, densify_freq=20, prune_freq=20,
                              densify_start=20, prune_start=40, max_sites=2400, log_freq=50)
    g = trainer.fit(tgt, sites, opts)
    r = orc.fit(tgt, sites, opts)
    print("gpu", g["log_psnr"], g["actual_sites"], "oracle", r["log_psnr"], r["actual_sites"])
    assert g["plan"] == r["plan"]
    assert g["active_count"] == int(np.sum(g["sites"][:, 0] >= 0))
    assert_same_fit(g, r)


@pytest.mark.parametrize("passes", [2, 3])
def test_multi_pass_updates_bitexact(orc, passes):
    """cand_update_passes > 1: the update runs on the second stream through a
    third field, so the concurrent gradient reads the field from before the
    update; repeats and oracle agree bit for bit."""
    tgt = synth.make_target(192, 160, 31, n_ellipses=30)
    sites = synth.init_sites_uniform(tgt, 600, 31)
    opts = config.fit_options(iterations=80, target_bpp=1.0, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2400, log_freq=40, cand_update_passes=passes)
    g = trainer.fit(tgt, sites, opts)
    assert_same_fit(g, trainer.fit(tgt, sites, opts))
    assert_same_fit(g, orc.fit(tgt, sites, opts))


def test_fit1024_300_iterations_bitexact(orc):
    """The 1024x1024 / 16,384-site fit with its densify/prune plan through 300
    iterations (densify and prune at iterations 100 and 200, the random-field opening
    whose tiles overflow the hash): two GPU runs and the oracle end with the
    same bits.  The oracle's 300 iterations take about 11 s on 16 threads."""
    orc.set_threads(min(16, os.cpu_count() or 1))
    tgt, sites, over = synth.fit_workload("fit1024")
    opts = config.fit_options(**dict(over, iterations=300, log_freq=50))
    g = trainer.fit(tgt, sites, opts)
    assert g["n_densify"] == 2 and g["n_prune"] == 2
    assert_same_fit(g, trainer.fit(tgt, sites, opts))
    assert_same_fit(g, orc.fit(tgt, sites, opts))


def test_fit1024_full_matches_oracle_golden():
    """The full 2,000-iteration 1024x1024 fit (the benchmarked workload) ends
    with the oracle's final sites, bit for bit (SHA-256 of the [262144, 10]
    buffer), every PSNR logged each 100 iterations and its schedule."""
    ref = np.load(os.path.join(os.path.dirname(mg.__file__), "fit1024_full.npz"))
    tgt, sites, over = synth.fit_workload("fit1024")
    g = trainer.fit(tgt, sites, config.fit_options(**dict(over, log_freq=mg.FIT1024_LOG_FREQ)))
    print("gpu final", g["final_psnr"], "oracle", float(ref["final_psnr"]))
    counts = [g["n_densify"], g["n_prune"], g["actual_sites"], g["active_estimate"], g["n_cand_updates"]]
    assert counts == ref["counts"].tolist()
    assert np.array_equal(np.asarray(g["log_psnr"], np.float32).view(np.uint32), ref["log_psnr"].view(np.uint32))
    assert np.array_equal(g["sites"][:256].view(np.uint32), ref["sites_head"].view(np.uint32))
    assert mg.digest(g["sites"]) == str(ref["sites_sha"])


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
