"""Batched fit (sadh_fit_batch): every job of a batch must fit as it does
alone, and as the oracle's restatement of the single-image schedule does.

Bar: bit-identity.  A job's sums do not depend on what else runs, so every job
ends with the sites, PSNR log and densify/prune schedule of the same fit run
alone, and of the oracle's fit of that image."""
import numpy as np
import pytest

from sad_amd import config, synth, trainer
from tests.scenes import assert_same_fit

pytestmark = pytest.mark.gpu


def _jobs(specs):
    tg, st = [], []
    for (w, h, seed, n) in specs:
        t = synth.make_target(w, h, seed, n_ellipses=20)
        tg.append(t)
        st.append(synth.init_sites_uniform(t, n, seed))
    return tg, st


def test_batch_of_one_matches_single_fit(orc):
    tg, st = _jobs([(96, 80, 51, 300)])
    opts = config.fit_options(iterations=80, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, n_sites=300, max_sites=300, log_freq=40)
    b = trainer.fit_batch(tg, st, opts)["jobs"][0]
    g = trainer.fit(tg[0], st[0], opts)
    r = orc.fit(tg[0], st[0], opts)
    print("batch", b["log_psnr"], "single", g["log_psnr"], "oracle", r["log_psnr"])
    assert_same_fit(b, g)
    assert_same_fit(b, r)


def test_batch_mixed_sizes_densify_prune(orc):
    """Three images of different sizes and aspect ratios with densify and
    prune every 20 iterations: per-job schedule and PSNR against the oracle."""
    specs = [(64, 48, 61, 150), (150, 100, 62, 500), (90, 170, 63, 400)]
    tg, st = _jobs(specs)
    opts = config.fit_options(iterations=120, target_bpp=1.5, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2000, log_freq=60)
    out = trainer.fit_batch(tg, st, opts)
    assert out["total_pixels"] == sum(w * h for w, h, _, _ in specs)
    for i, (t, s0) in enumerate(zip(tg, st)):
        b = out["jobs"][i]
        r = orc.fit(t, s0, dict(opts, n_sites=s0.shape[0]))
        print(i, "batch", b["log_psnr"], b["actual_sites"], "oracle", r["log_psnr"], r["actual_sites"])
        assert b["plan"] == r["plan"]
        assert b["active_count"] == int(np.sum(b["sites"][:, 0] >= 0))
        assert_same_fit(b, r)


@pytest.mark.parametrize("passes", [2, 3])
def test_batch_multi_pass_updates(orc, passes):
    """cand_update_passes > 1 in a batch: a densifying job updates before its
    gradient, the others after it, so no pass writes the field a concurrent
    gradient reads; every job equals the oracle's fit."""
    specs = [(150, 100, 62, 500), (90, 170, 63, 400)]
    tg, st = _jobs(specs)
    opts = config.fit_options(iterations=90, target_bpp=1.5, densify_freq=24, prune_freq=20, densify_start=24,
                              prune_start=40, max_sites=2000, log_freq=45, cand_update_passes=passes)
    out = trainer.fit_batch(tg, st, opts)
    for i, (t, s0) in enumerate(zip(tg, st)):
        assert_same_fit(out["jobs"][i], orc.fit(t, s0, dict(opts, n_sites=s0.shape[0])))


def test_batch_rejects_bad_jobs():
    from sad_amd._lib import SadHipError
    tg, st = _jobs([(64, 48, 71, 100)])
    with pytest.raises(ValueError):
        trainer.fit_batch([], [], config.fit_options())
    opts = config.fit_options(iterations=4, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0)
    bad = synth.sites_array(0)
    with pytest.raises(SadHipError):
        trainer.fit_batch(tg, [bad], opts)


def test_fits_repeat_exactly():
    """The fit's cross-tile sums are exact integers and every per-tile sum runs
    in a fixed order, so a fit with densify and prune repeats bit for bit,
    alone and batched."""
    specs = [(150, 100, 62, 500), (90, 170, 63, 400)]
    tg, st = _jobs(specs)
    opts = config.fit_options(iterations=120, target_bpp=1.5, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2000, log_freq=30)
    runs = [trainer.fit_batch(tg, st, opts)["jobs"] for _ in range(2)]
    for i, s0 in enumerate(st):
        single = [trainer.fit(tg[i], s0, dict(opts, n_sites=s0.shape[0])) for _ in range(2)]
        for r in (runs[1][i], single[0], single[1]):
            assert np.array_equal(r["sites"].view(np.uint32), runs[0][i]["sites"].view(np.uint32))
            assert np.array_equal(np.asarray(r["log_psnr"]), np.asarray(runs[0][i]["log_psnr"]))


@pytest.mark.gpu
@pytest.mark.parametrize("ssim_weight", [0.0, 0.5])
def test_graph_ticks_equal_eager_ticks(ssim_weight):
    """graph_ticks replays each steady iteration (no candidate update, densify,
    prune or log) as a captured hipGraph.  Those replays run between eager
    iterations, host synchronisations and recaptures (a densify changes a site
    count; with the SSIM render a candidate flip changes a launch argument), and
    every job ends bit-identical to the same batch without graphs."""
    specs = [(150, 100, 62, 500), (90, 170, 63, 400), (64, 48, 64, 120)]
    tg, st = _jobs(specs)
    opts = config.fit_options(iterations=130, target_bpp=1.5, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2000, log_freq=30, ssim_weight=ssim_weight)
    eager = trainer.fit_batch(tg, st, opts)
    graph = trainer.fit_batch(tg, st, dict(opts, graph_ticks=1))
    print("graph ticks", graph["graph_ticks"], "captures", graph["graph_captures"])
    assert eager["graph_ticks"] == 0 and graph["graph_ticks"] >= 80
    assert 4 <= graph["graph_captures"] < graph["graph_ticks"]
    for e, g in zip(eager["jobs"], graph["jobs"]):
        assert np.array_equal(g["sites"].view(np.uint32), e["sites"].view(np.uint32))
        assert np.array_equal(g["log_psnr"], e["log_psnr"])
        assert (g["n_densify"], g["n_prune"], g["actual_sites"]) == (e["n_densify"], e["n_prune"], e["actual_sites"])


def test_batch64_full_size(orc):
    """The 64-image batch at its stated size: sides log-uniform in [128, 2048],
    aspect ratios in [0.5, 2], bpp-1.0 initial sites (about 60 MP and 470k
    sites), 500 iterations with graph-captured steady ticks.  Every job ends
    bit-identical to the same image fitted alone; the three smallest images
    also to the oracle's fits."""
    targets, sites = synth.batch_workload()
    opts = config.fit_options(iterations=500, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, log_freq=250)
    out = trainer.fit_batch(targets, sites, dict(opts, graph_ticks=1))
    assert len(out["jobs"]) == 64 and out["graph_ticks"] > 400
    for i, (t, s0) in enumerate(zip(targets, sites)):
        assert_same_fit(out["jobs"][i], trainer.fit(t, s0, opts))
    small = sorted(range(64), key=lambda i: targets[i].shape[0] * targets[i].shape[1])[:3]
    for i in small:
        assert_same_fit(out["jobs"][i], orc.fit(targets[i], sites[i], dict(opts, n_sites=sites[i].shape[0])))


def test_many_small_jobs_match_single_fits():
    """40 small images in one batch: the log ticks (one render and one sum
    launch for every logging job), the Hilbert tables (one sort of all jobs'
    sites, keyed by job) and the retirement (one gather and copy) run batched.
    Every job ends with the sites, PSNR log and active count of its single fit."""
    specs = [(48 + 8 * (i % 5), 40 + 8 * (i % 3), 300 + i, 64 + 16 * (i % 4)) for i in range(40)]
    tg, st = _jobs(specs)
    opts = config.fit_options(iterations=50, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, log_freq=20)
    out = trainer.fit_batch(tg, st, opts)
    for i, (t, s0) in enumerate(zip(tg, st)):
        b = out["jobs"][i]
        assert b["n_logs"] == 4
        g = trainer.fit(t, s0, opts)
        assert_same_fit(b, g)
        assert b["active_count"] == g["active_count"] == int(np.sum(b["sites"][: s0.shape[0], 0] >= 0))


def test_masked_batch_matches_single_fits():
    """Jobs with masks (partly zero, partly fractional): the batched log sums count only pixels with mask > 0
    and divide by each job's own mask count, as the single fit does; every job ends as its single fit."""
    specs = [(80, 64, 410, 200), (72, 90, 411, 240), (96, 48, 412, 160)]
    tg, st = _jobs(specs)
    rng = np.random.default_rng(7)
    masks = []
    for (w, h, _, _) in specs:
        m = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32)
        m[: h // 3] = 0.0
        masks.append(m)
    opts = config.fit_options(iterations=40, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, log_freq=10)
    out = trainer.fit_batch(tg, st, opts, masks=masks)
    for i, (t, s0) in enumerate(zip(tg, st)):
        g = trainer.fit(t, s0, opts, mask=masks[i])
        assert_same_fit(out["jobs"][i], g)
        assert out["jobs"][i]["n_logs"] == 5


def test_autotune_keeps_every_result():
    """autotune times steady ticks (from the 41st) under each launch schedule (128- or 64-slot tile hash,
    256- or 64-thread tau blocks) and keeps the fastest; no schedule changes a result, so every job ends as in
    the batch without tuning, with and without graph replays."""
    specs = [(150, 100, 62, 500), (90, 170, 63, 400), (64, 48, 64, 120)]
    tg, st = _jobs(specs)
    opts = config.fit_options(iterations=90, target_bpp=1.5, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2000, log_freq=30)
    plain = trainer.fit_batch(tg, st, opts)
    assert plain["tuned_hash_slots"] == 0 and plain["tuned_tau_block"] == 0
    for graph in (0, 1):
        tuned = trainer.fit_batch(tg, st, dict(opts, autotune=1, graph_ticks=graph))
        print("graph", graph, "tuned schedule: hash", tuned["tuned_hash_slots"], "tau block", tuned["tuned_tau_block"])
        assert tuned["tuned_hash_slots"] in (64, 128) and tuned["tuned_tau_block"] in (64, 256)
        for a, b in zip(plain["jobs"], tuned["jobs"]):
            assert_same_fit(b, a)
