"""Pins of the oracle's per-site sums and schedule solver, written independently
of sad_math.h (numpy float64 and Python's exact rationals).  CPU only.

* The fixed point of the library's cross-tile sums (orc_fix_split,
  orc_fix_value) against exact rational arithmetic, and the order
  independence that follows.
* The tile-ordered sums (gradients_fix, site_stats_fix, render_backward_fix)
  against the oracle's per-site float64 sums of the same terms.
* computeSiteStatsSimpleKernel (utils.cu:106-231): conservation laws and a
  one-site closed form.
* removalDeltaForSite (gradients.cu:82-100): brute force, by rendering the
  image again with the site removed, and the 1e6 branch.
* simulateFinalSites / solveTargetBpp (sad_cuda.cu:904-1066): hand-counted
  schedules and the solved plans.
"""
from fractions import Fraction

import numpy as np
import pytest

from tests.scenes import converged_field, iss_of, random_sites
from tests.tolerance import Rel

EMPTY = 0xFFFFFFFF


# ------------------------------------------------------------- fixed point

def _exact_split(v: float):
    """(lo, hi) of v by exact rational arithmetic: hi = trunc(v), lo = the
    fraction times 2^40 rounded half to even"""
    q = Fraction(v)
    hi = int(q)                                   # int() truncates toward zero
    f = (q - hi) * (1 << 40)
    lo = round(f)                                 # Python rounds Fractions half to even
    return lo, hi


def test_fix_split_matches_exact_rationals(orc):
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.normal(0, 1, 300), rng.normal(0, 1e6, 200), rng.normal(0, 1e-9, 200), rng.normal(0, 1e-12, 100),
        np.ldexp(rng.uniform(-1, 1, 200), rng.integers(-60, 60, 200)),
        [0.0, -0.0, 0.5, -0.5, 1.5, -1.5, 2.0 ** -41, -(2.0 ** -41), 3 * 2.0 ** -41, 2.0 ** 23 + 1, -(2.0 ** 40)],
    ]).astype(np.float32)
    for v in vals:
        assert orc.fix_split(v) == _exact_split(float(v)), float(v)
    # saturation at +-2^62 and NaN
    assert orc.fix_split(np.float32(3e19)) == (0, 1 << 62)
    assert orc.fix_split(np.float32(-np.inf)) == (0, -(1 << 62))
    assert orc.fix_split(np.float32(np.nan)) == (0, 0)


def test_fix_sums_are_exact_and_order_independent(orc):
    """Partials that are multiples of 2^-40 sum exactly (the value is the exact
    rational sum rounded once), whatever their order and magnitudes."""
    rng = np.random.default_rng(2)
    v = (np.ldexp(rng.uniform(-1, 1, 4000), rng.integers(-40, 30, 4000))).astype(np.float32)
    v = np.where(np.abs(v) < 2.0 ** -40, np.float32(0), v)
    pairs = np.array([orc.fix_split(x) for x in v], np.int64)
    total = pairs.sum(0)
    perm = pairs[rng.permutation(len(v))]
    assert np.array_equal(perm.sum(0), total)
    exact = sum((Fraction(float(x)) for x in v), Fraction(0))
    got = orc.fix_values(total[None, :])[0]
    assert got == np.float32(float(Fraction(int(total[1])) + Fraction(int(total[0]), 1 << 40)))
    assert abs(float(got) - float(exact)) <= abs(float(exact)) * 2.0 ** -23 + 2.0 ** -40


# ------------------------------------------------ tile-ordered sums vs f64

def _scene(orc, w, h, n, seed, passes):
    sites = random_sites(w, h, n, seed)
    if passes:
        c0, c1 = converged_field(orc, sites, w, h, passes=passes)
    else:
        c0, c1 = orc.init_candidates(n, seed, False, w, h)
        c0 = orc.jfa_seed(c0, sites, w, h)
    rng = np.random.default_rng(seed + 1)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    mask[rng.random((h, w)) < 0.1] = 0.0
    return sites, c0, c1, target, mask


@pytest.mark.parametrize("passes", [0, 10])
def test_tile_sums_match_per_site_f64_sums(orc, passes):
    """passes=0: a random candidate field, ~2000 distinct ids per tile (the
    library's hash rounds); 10: a converged one."""
    w, h, n = 144, 112, 1500
    sites, c0, c1, target, mask = _scene(orc, w, h, n, 41, passes)
    iss = iss_of(w, h)
    acc, ab, racc, rab = orc.gradients(c0, c1, target, mask, sites, iss, w, h, compute_removal=True)
    g, r, parts = orc.gradients_fix(c0, c1, target, mask, sites, iss, w, h, compute_removal=True)
    q = parts.astype(np.float64) * orc.FIX_HALF_STEP
    assert Rel(g, acc, ab, q[:, None]) <= 1.0
    assert Rel(r, racc, rab, q) <= 1.0
    st = orc.site_stats(c0, c1, target, mask, sites, iss, w, h)
    sf, sparts = orc.site_stats_fix(c0, c1, target, mask, sites, iss, w, h)
    assert np.array_equal(parts[sparts > 0] > 0, np.ones(int(np.sum(sparts > 0)), bool))
    bound = 1e-5 * np.abs(st).max(0) + sparts[:, None] * orc.FIX_HALF_STEP
    assert np.all(np.abs(sf - st) <= 1e-5 * np.abs(st) + bound)
    go = np.random.default_rng(43).normal(0, 1, (h, w, 3)).astype(np.float32)
    bacc, bab = orc.render_backward(c0, c1, go, sites, iss, w, h)
    bf, bparts = orc.render_backward_fix(c0, c1, go, sites, iss, w, h)
    assert Rel(bf, bacc, bab, bparts[:, None].astype(np.float64) * orc.FIX_HALF_STEP) <= 1.0


# ----------------------------------------------------------- site statistics

def _weights_f64(c0, c1, sites, w, h):
    """per pixel: candidate ids and their float64 softmax weights (render.cu:34-91 closed form; 0 for invalid)"""
    s = sites.astype(np.float64)
    S = max(w, h)
    ids = np.concatenate([c0, c1], 1).reshape(h, w, 8)
    wts = np.zeros((h, w, 8))
    for y in range(h):
        for x in range(w):
            lg = np.full(8, -np.inf)
            for k, i in enumerate(ids[y, x]):
                if i >= len(s) or s[i, 0] < 0:
                    continue
                d = np.array([x, y]) - s[i, :2]
                u = s[i, 7:9]
                proj = u @ d
                perp2 = max(d @ d - proj * proj, 0.0)
                d2 = max((np.exp(s[i, 9]) * proj * proj + np.exp(-s[i, 9]) * perp2) / S**2, 1e-8)
                lg[k] = -np.exp(s[i, 2]) * (np.sqrt(d2) - s[i, 3] / S)
            if np.isfinite(lg.max()):
                e = np.exp(lg - lg.max())
                wts[y, x] = e / e.sum()
    return ids, wts


def test_site_stats_conservation(orc):
    """Per pixel the weights sum to one, so over all sites: sum of mass = the
    number of rendered masked pixels, sum of energy = sum of the pixels' squared
    error, sum of w.err = sum of min(err, 1), and the first moments sum to the
    min(err, 1)-weighted pixel coordinates."""
    w, h, n = 40, 32, 90
    sites, c0, c1, target, mask = _scene(orc, w, h, n, 51, passes=6)
    iss = iss_of(w, h)
    st = orc.site_stats(c0, c1, target, mask, sites, iss, w, h)
    render = orc.render(c0, c1, sites, iss, w, h)
    err = np.sum((render.astype(np.float64) - target) ** 2, axis=2)
    m = mask > 0
    e1 = np.minimum(err, 1.0)
    yy, xx = np.mgrid[0:h, 0:w]
    np.testing.assert_allclose(st[:, 0].sum(), m.sum(), rtol=1e-5)
    np.testing.assert_allclose(st[:, 1].sum(), err[m].sum(), rtol=1e-4)
    np.testing.assert_allclose(st[:, 2].sum(), e1[m].sum(), rtol=1e-4)
    np.testing.assert_allclose(st[:, 3].sum(), (e1 * xx)[m].sum(), rtol=1e-4)
    np.testing.assert_allclose(st[:, 4].sum(), (e1 * yy)[m].sum(), rtol=1e-4)
    np.testing.assert_allclose(st[:, 7].sum(), (e1 * yy * yy)[m].sum(), rtol=1e-4)
    # per site against the float64 weights: mass = sum over pixels of the site's weight
    ids, wts = _weights_f64(c0, c1, sites, w, h)
    mass = np.zeros(n)
    for k in range(8):
        sel = (ids[:, :, k] < n) & m
        np.add.at(mass, ids[:, :, k][sel], wts[:, :, k][sel])
    np.testing.assert_allclose(st[:, 0], mass, rtol=1e-4, atol=1e-5)


def test_site_stats_one_site_closed_form(orc):
    """Every pixel's only candidate is site 3: its weight is 1, pred is its
    colour, and the statistics are sums over the image."""
    w, h, n = 24, 20, 6
    sites = random_sites(w, h, n, 52, inactive_frac=0.0)
    c0 = np.full((w * h, 4), EMPTY, np.uint32)
    c1 = c0.copy()
    c0[:, 0] = 3
    rng = np.random.default_rng(53)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = (rng.random((h, w)) > 0.2).astype(np.float32)
    st = orc.site_stats(c0, c1, target, mask, sites, iss_of(w, h), w, h)
    err = np.sum((sites[3, 4:7].astype(np.float64) - target) ** 2, axis=2)
    m = mask > 0
    e1 = np.minimum(err, 1.0)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    want = [m.sum(), err[m].sum(), e1[m].sum(), (e1 * xx)[m].sum(), (e1 * yy)[m].sum(), (e1 * xx * xx)[m].sum(),
            (e1 * xx * yy)[m].sum(), (e1 * yy * yy)[m].sum()]
    np.testing.assert_allclose(st[3], want, rtol=2e-6)
    assert np.all(st[np.arange(n) != 3] == 0)


# ------------------------------------------------------------- removal delta

def test_removal_delta_brute_force(orc):
    """A site's removal delta is the change of the squared error when the image
    is rendered again without it (its weight renormalised away), summed over
    the pixels it covers; a pixel it fills alone (1 - w < 1e-4) counts 1e6."""
    w, h, n = 36, 28, 70
    sites, c0, c1, target, mask = _scene(orc, w, h, n, 61, passes=6)
    sites[:, 2] = 0.5                                 # soft weights: no pixel is near w = 1
    mask[:] = 1.0
    iss = iss_of(w, h)
    _, _, rd, rab = orc.gradients(c0, c1, target, mask, sites, iss, w, h, compute_removal=True)
    ids, wts = _weights_f64(c0, c1, sites, w, h)
    base = orc.render(c0, c1, sites, iss, w, h).astype(np.float64)
    e0 = np.sum((base - target) ** 2, axis=2)
    checked = 0
    for a in range(0, n, 7):
        if sites[a, 0] < 0:
            continue
        wa = np.sum(np.where(ids == a, wts, 0.0), axis=2)
        covered = wa > 0
        if not covered.any():
            continue
        assert wa.max() < 1.0 - 1e-3
        without = sites.copy()
        without[a, 0:2] = -1.0
        e1 = np.sum((orc.render(c0, c1, without, iss, w, h).astype(np.float64) - target) ** 2, axis=2)
        want = float(np.sum((e1 - e0)[covered]))
        assert abs(rd[a] - want) <= 1e-4 * rab[a] + 1e-6, (a, rd[a], want)
        checked += 1
    assert checked >= 5
    # sites 5 and 6 share the first half of the pixels, site 5 fills the second half alone (w = 1: the 1e6
    # branch); removing one of two sites leaves exactly the other's colour
    sites[5, 2] = sites[6, 2] = 0.0
    half = w * h // 2
    c0 = np.full((w * h, 4), EMPTY, np.uint32)
    c1 = c0.copy()
    c0[:, 0] = 5
    c0[:half, 1] = 6
    _, _, rd, rab = orc.gradients(c0, c1, target, mask, sites, iss, w, h, compute_removal=True)
    ids, wts = _weights_f64(c0, c1, sites, w, h)
    w6 = np.sum(np.where(ids == 6, wts, 0.0), axis=2).reshape(-1)[:half]
    assert 1e-3 < w6.min() and w6.max() < 1 - 1e-3
    t = target.reshape(-1, 3)[:half].astype(np.float64)
    pred = orc.render(c0, c1, sites, iss, w, h).reshape(-1, 3)[:half].astype(np.float64)
    e0 = np.sum((pred - t) ** 2, 1)
    want5 = (w * h - half) * 1e6 + np.sum(np.sum((sites[6, 4:7] - t) ** 2, 1) - e0)
    want6 = np.sum(np.sum((sites[5, 4:7] - t) ** 2, 1) - e0)
    assert abs(rd[5] - want5) <= 1e-6 * want5
    assert abs(rd[6] - want6) <= 1e-4 * rab[6] + 1e-6


# ------------------------------------------------------------ bpp schedule

def test_simulate_final_sites_known_answers(orc):
    """Counted by hand.  100 sites, densify 10% at iterations 0, 10, ..., 40 and
    prune 5% at 20, 30, 40 (prune during densify):
      it 0:  +int(100*0.1)=10 -> 110        it 10: +int(11.0)=11 -> 121
      it 20: +int(12.1)=12 -> 133, -int(6.65)=6 -> 127
      it 30: +int(12.7)=12 -> 139, -int(6.95)=6 -> 133
      it 40: +int(13.3)=13 -> 146, -int(7.3)=7 -> 139
    With a 125-site capacity densify adds only up to it (prunes do not return
    capacity): 10 -> 110, 11 -> 121, min(12, 4) = 4 -> 125, then prunes only:
    125 - int(6.25)=6 = 119, 119 - int(5.95)=5 = 114, 114 - int(5.7)=5 = 109.
    Without prune during densify the prunes start at d_end = 40: densify to 133
    at 20, +13 -> 146 at 30, +14 -> 160 at 40, then -int(8.0)=8 -> 152."""
    args = dict(iters=45, densify_enabled=1, d_start=0, d_end=40, d_freq=10, d_pct=0.1, prune_during=1, p_start=20,
                p_end=45, p_freq=10, p_pct=0.05)
    assert orc.simulate_final_sites(100, 1000, **args) == 139
    assert orc.simulate_final_sites(100, 125, **args) == 109
    # prune only after densify ends (prune_during = 0 moves the start to d_end = 40): 146 - 7 = 139 at 40 only
    assert orc.simulate_final_sites(100, 1000, **dict(args, prune_during=0)) == 152
    assert orc.simulate_final_sites(100, 1000, **dict(args, d_pct=0.0, p_pct=0.0)) == 100


def test_solver_meets_the_bpp_target(orc):
    """The benchmarked 1024x1024 plan lands exactly on round(1.0 * 1024^2 / 128)
    = 8192 active sites (128 bits per site), the count the fit ends with; when
    the capacity cannot reach the target, the solver's count lies between the
    best of its own 81-point scale grid and the best of a 2001-point scan."""
    from sad_amd import config, synth
    _, _, over = synth.fit_workload("fit1024")
    o = config.fit_options(**over)
    plan = orc.fit_plan(o, 1024, 1024)
    p_end = o["prune_end"] if o["prune_end"] > 0 else o["iterations"] - 1
    sched = (o["iterations"], 1, o["densify_start"], o["densify_end"], o["densify_freq"])
    prune = (o["prune_during_densify"], o["prune_start"], p_end, o["prune_freq"])
    final = orc.simulate_final_sites(16384, plan["capacity"], *sched, plan["densify_percentile"], *prune,
                                     plan["prune_percentile"])
    assert final == round(1.0 * 1024 * 1024 / 128) == 8192
    # unreachable: 8 bpp on 256x256 wants 4096 sites; the buffer holds 1500
    base_d, base_p = o["densify_percentile"], o["prune_percentile"]
    args = (600, 1, 20, 3000, 100)
    d, p, got = orc.solve_target_bpp(8.0, 256, 256, 1000, 1500, *args, base_d, 1, 100, 599, 100, base_p)
    smax = min(0.95 / max(base_d, base_p), 50.0)

    def sim(sc):
        return orc.simulate_final_sites(1000, 1500, *args, min(0.95, base_d * sc), 1, 100, 599, 100,
                                        min(0.95, base_p * sc))
    grid = max(sim(smax * i / 80) for i in range(81))
    scan = max(sim(x) for x in np.linspace(0.0, smax, 2001))
    assert grid <= got <= scan < 4096
