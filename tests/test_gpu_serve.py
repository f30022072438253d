"""Continuous-batching server: requests that arrive while others run are
admitted, overlap on the device, and each result equals the same request run
alone, bit for bit (fits: sites and PSNR log; renders: the image).  Fits also
equal the oracle's fits bit for bit."""
import numpy as np
import pytest

from sad_amd import config, synth, trainer
from sad_amd.serve import Server
from tests.scenes import assert_same_fit

pytestmark = pytest.mark.gpu


def _opts():
    return config.fit_options(iterations=60, target_bpp=1.5, densify_freq=20, prune_freq=20, densify_start=20,
                              prune_start=40, max_sites=2000, log_freq=30)


def _fit_case(w, h, seed, n):
    t = synth.make_target(w, h, seed, n_ellipses=20)
    return t, synth.init_sites_uniform(t, n, seed)


def test_mixed_stream_matches_alone(orc):
    opts = _opts()
    fits = [_fit_case(64, 48, 81, 150), _fit_case(120, 96, 82, 400), _fit_case(96, 160, 83, 350)]
    renders = [(synth.init_sites_random(128, 96, 500, 84), 128, 96),
               (synth.init_sites_random(200, 160, 2000, 85), 200, 160)]
    with Server(opts) as srv:
        ids = {}
        ids["f0"] = srv.submit_fit(*fits[0])
        for _ in range(5):
            srv.step()
        ids["r0"] = srv.submit_render(*renders[0])
        ids["f1"] = srv.submit_fit(*fits[1])
        for _ in range(7):
            srv.step()
        ids["f2"] = srv.submit_fit(*fits[2])
        ids["r1"] = srv.submit_render(*renders[1])
        while srv.outstanding():
            srv.step()
        got = {k: srv.result(v) for k, v in ids.items()}
    # the fits ran concurrently: f1 was admitted before f0 finished
    assert got["f1"]["admit_s"] < got["f0"]["done_s"]
    for i, (t, s0) in enumerate(fits):
        g = got[f"f{i}"]
        alone = trainer.fit(t, s0, opts)
        r = orc.fit(t, s0, dict(opts, n_sites=s0.shape[0]))
        print(i, "mix", g["log_psnr"], "alone", alone["log_psnr"], "oracle", r["log_psnr"])
        assert_same_fit(g, alone)
        assert_same_fit(g, r)
    for i, (s, w, h) in enumerate(renders):
        alone = trainer.render_sites(s, w, h, cand_passes=16)
        assert np.array_equal(got[f"r{i}"]["image"], alone["image"])


def test_memory_budget_admits_in_turn():
    """With room for one fit only, the second waits for the first; both match
    their solo runs."""
    opts = config.fit_options(iterations=40, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, log_freq=20)
    a, b = _fit_case(160, 120, 91, 600), _fit_case(150, 130, 92, 600)
    with Server(opts, memory_budget=3e6) as srv:   # one fit projects ~2 MB
        ia, ib = srv.submit_fit(*a), srv.submit_fit(*b)
        while srv.outstanding():
            srv.step()
        ra, rb = srv.result(ia), srv.result(ib)
    assert rb["admit_s"] >= ra["done_s"]
    for (t, s0), r in ((a, ra), (b, rb)):
        assert np.array_equal(r["sites"].view(np.uint32), trainer.fit(t, s0, opts)["sites"].view(np.uint32))


def test_request_larger_than_budget_is_refused():
    from sad_amd._lib import SadHipError
    opts = config.fit_options(iterations=4, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0)
    with Server(opts, memory_budget=1e5) as srv:
        with pytest.raises(SadHipError):
            srv.submit_fit(*_fit_case(128, 128, 93, 200))
        with pytest.raises(SadHipError):
            srv.submit_render(synth.init_sites_random(256, 256, 100, 94), 256, 256)


def test_served_stream_full_size():
    """A served stream at its stated sizes: fits of 256-2048 px sides (300
    iterations, bpp 1.0, densify and prune on) and renders of 4k-1M random
    sites at 512^2-4096^2, all submitted while earlier ones run.  Each result
    equals the same request run alone, bit for bit."""
    reqs = synth.serve_workload(n_requests=12, rate=1000.0, seed=5)
    kinds = [r["kind"] for r in reqs]
    assert "fit" in kinds and "render" in kinds
    opts = config.fit_options(iterations=300, target_bpp=1.0, log_freq=100)
    with Server(opts) as srv:
        ids = []
        for q in reqs:
            if q["kind"] == "fit":
                ids.append(srv.submit_fit(q["target"], q["sites"]))
            else:
                ids.append(srv.submit_render(q["sites"], q["width"], q["height"]))
            srv.step()
        while srv.outstanding():
            srv.step()
        got = [srv.result(i) for i in ids]
    for q, g in zip(reqs, got):
        print(q["kind"], q["width"], q["height"], len(q["sites"]))
        if q["kind"] == "fit":
            assert_same_fit(g, trainer.fit(q["target"], q["sites"], opts))
        else:
            alone = trainer.render_sites(q["sites"], q["width"], q["height"], cand_passes=16)
            assert np.array_equal(g["image"], alone["image"])
