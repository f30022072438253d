"""Pins the CPU oracle (oracle/, sad_amd/csrc/sad_math.h) against independent
references: libm/numpy for the arithmetic primitives, textbook algorithms and
closed forms for the kernels, brute force on tiny inputs, and central finite
differences for the analytic gradient.  CPU only."""
import numpy as np
import pytest

from tests.scenes import iss_of, random_sites


def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    return np.abs(got.astype(np.float64) - ref) / np.spacing(np.abs(ref32)).astype(np.float64)


# ---------------------------------------------------------------- primitives

def test_expf_within_one_ulp(orc):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-87, 88.7, 400000), rng.uniform(-1, 1, 100000),
                        rng.uniform(-20, 20, 100000), rng.uniform(-103.9, -87.4, 20000)]).astype(np.float32)
    y = orc.expf(x)
    ref = np.exp(x.astype(np.float64))
    ok = ref > 1.2e-38
    assert ulp_err(y[ok], ref[ok]).max() <= 1.0
    sub = (ref > 0) & (ref < 1.2e-38)
    assert np.max(np.abs(y[sub] - ref[sub]) / 2.0**-149) <= 1.0
    sp = orc.expf(np.array([0.0, -np.inf, np.inf, np.nan, 89.0, -105.0], np.float32))
    assert sp[0] == 1.0 and sp[1] == 0.0 and np.isinf(sp[2]) and np.isnan(sp[3]) and np.isinf(sp[4]) and sp[5] == 0.0


def test_logf_within_two_ulp(orc):
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-85, 85, 300000)), rng.uniform(0.5, 2.0, 100000),
                        np.float32(2.0) ** -140 * rng.uniform(1, 100, 1000)]).astype(np.float32)
    y = orc.logf(x)
    ref = np.log(x.astype(np.float64))
    assert ulp_err(y, ref).max() <= 2.0
    assert orc.logf(np.array([1.0], np.float32))[0] == 0.0


def test_fp16_round_to_nearest_even(orc):
    bits = np.arange(0, 2**32, 1013, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    nn = ~np.isnan(f)
    with np.errstate(over="ignore"):
        ref = f[nn].astype(np.float16).view(np.uint16)
    assert np.array_equal(orc.f32_to_f16_bits(f[nn]), ref)
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    back = orc.f16_bits_to_f32(h)
    refb = h.view(np.float16).astype(np.float32)
    ok = ~np.isnan(refb)
    assert np.array_equal(back[ok].view(np.uint32), refb[ok].view(np.uint32))
    # ties go to even: 2049 is exactly between 2048 and 2050 in binary16
    assert orc.f16_bits_to_f32(orc.f32_to_f16_bits(np.array([2049.0, 2051.0], np.float32))).tolist() == [2048.0, 2052.0]


def test_hilbert_curve(orc):
    # bits = 1 known answers (sad_cuda.cu:1523-1545 orientation)
    assert [orc.hilbert_index(x, y, 1) for x, y in [(0, 0), (0, 1), (1, 1), (1, 0)]] == [0, 1, 2, 3]
    for bits in (1, 2, 3, 5):
        n = 1 << bits
        d = np.array([[orc.hilbert_index(x, y, bits) for y in range(n)] for x in range(n)])
        assert sorted(d.ravel().tolist()) == list(range(n * n))          # bijection
        pos = np.zeros((n * n, 2), int)
        for x in range(n):
            for y in range(n):
                pos[d[x, y]] = (x, y)
        steps = np.abs(np.diff(pos, axis=0)).sum(1)
        assert np.all(steps == 1)                                        # a continuous curve


def test_jump_schedule(orc):
    assert [orc.jump_step(i, 1024, 1024) for i in range(12)] == [512, 256, 128, 64, 32, 16, 8, 4, 2, 1, 1, 1]
    assert [orc.jump_step(i, 300, 170) for i in range(10)] == [256, 128, 64, 32, 16, 8, 4, 2, 1, 1]
    assert orc.jump_step(0, 1, 1) == 1


# --------------------------------------------------------------------- sort

@pytest.mark.parametrize("n", [1, 17, 1024, 5000])
def test_radix_sort_is_stable_sort(orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 3] = keys[n // 2]                                   # many duplicates
    pairs = np.stack([keys, np.arange(n, dtype=np.uint32)], 1)
    assert np.array_equal(orc.radix_sort(pairs), pairs[np.argsort(keys, kind="stable")])


def test_radix_sort_two_pass_mode_sorts_low_16_bits(orc):
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 2**20, 3000).astype(np.uint32)
    pairs = np.stack([keys, np.arange(3000, dtype=np.uint32)], 1)
    out = orc.radix_sort(pairs, 1 << 16)
    assert np.array_equal(out, pairs[np.argsort(keys & 0xFFFF, kind="stable")])


# ----------------------------------------------------------- render / forward

def _render_f64(c0, c1, sites, w, h):
    """Softmax render of render.cu:5-92 in float64 from the closed form."""
    s = sites.astype(np.float64)
    out = np.zeros((h, w, 3))
    S = max(w, h)
    ids = np.concatenate([c0, c1], 1).reshape(h, w, 8)
    for y in range(h):
        for x in range(w):
            lg, cols = [], []
            for i in ids[y, x]:
                if i >= len(s) or s[i, 0] < 0:
                    continue
                d = np.array([x, y]) - s[i, :2]
                u = s[i, 7:9]
                proj = u @ d
                perp2 = max(d @ d - proj * proj, 0.0)
                d2 = max((np.exp(s[i, 9]) * proj * proj + np.exp(-s[i, 9]) * perp2) / S**2, 1e-8)
                lg.append(-np.exp(s[i, 2]) * (np.sqrt(d2) - s[i, 3] / S))
                cols.append(s[i, 4:7])
            if lg:
                lg = np.array(lg)
                e = np.exp(lg - lg.max())
                out[y, x] = (e / e.sum()) @ np.array(cols)
    return out


def test_render_matches_closed_form(orc):
    w, h, n = 24, 18, 40
    sites = random_sites(w, h, n, 2)
    c0, c1 = orc.init_candidates(n, 9, False, w, h)
    got = orc.render(c0, c1, sites, iss_of(w, h), w, h)
    np.testing.assert_allclose(got, _render_f64(c0, c1, sites, w, h), atol=2e-5)


def test_render_special_cases(orc):
    w, h, n = 8, 6, 5
    sites = random_sites(w, h, n, 3, inactive_frac=0.0)
    c0 = np.full((w * h, 4), 0xFFFFFFFF, np.uint32)
    c1 = c0.copy()
    c0[:, 0] = 2                                  # one valid candidate: exactly its colour
    out = orc.render(c0, c1, sites, iss_of(w, h), w, h)
    assert np.array_equal(out.reshape(-1, 3), np.repeat(sites[2:3, 4:7], w * h, 0))
    c0[:, 0] = 0xFFFFFFFF                         # nothing valid: black
    assert np.all(orc.render(c0, c1, sites, iss_of(w, h), w, h) == 0)
    sites[:, 4:7] = 0.375                         # constant colours: weights sum to one
    c0, c1 = orc.init_candidates(n, 1, False, w, h)
    np.testing.assert_allclose(orc.render(c0, c1, sites, iss_of(w, h), w, h), 0.375, atol=1e-7)


# ----------------------------------------------------- gradient vs finite diff

def _loss_f64(sites, c0, c1, target, w, h):
    r = _render_f64(c0, c1, sites, w, h)
    return float(np.sum((r - target) ** 2))


def test_gradient_matches_central_differences(orc):
    """The analytic per-site gradient (sad_grad_terms summed by the oracle) is
    checked against FP64 central differences of the FP64 closed-form render
    with the candidate field frozen."""
    w, h, n = 20, 16, 14
    rng = np.random.default_rng(5)
    sites = random_sites(w, h, n, 6, inactive_frac=0.0)
    sites[:, 2] = rng.uniform(1.5, 3.0, n)           # soft weights so every candidate matters
    sites[:, 9] = rng.uniform(-0.8, 0.8, n)
    # |dir| slightly below 1 keeps d^2 - (dir.d)^2 > 0, so the perp^2 >= 0 clamp is
    # inactive within the difference step and the loss is smooth in dir
    sites[:, 7:9] *= 0.999
    c0, c1 = orc.init_candidates(n, 4, False, w, h)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    acc, ab, _, _ = orc.gradients(c0, c1, target, mask, sites, iss_of(w, h), w, h)
    eps = {0: 1e-3, 1: 1e-3, 2: 1e-4, 3: 1e-3, 4: 1e-4, 5: 1e-4, 6: 1e-4, 7: 1e-4, 8: 1e-4, 9: 1e-4}
    base = sites.astype(np.float64)
    fails = []
    for i in range(n):
        for j in range(10):
            sp, sm = base.copy(), base.copy()
            sp[i, j] += eps[j]
            sm[i, j] -= eps[j]
            fd = (_loss_f64(sp, c0, c1, target, w, h) - _loss_f64(sm, c0, c1, target, w, h)) / (2 * eps[j])
            an = acc[i, j]
            if abs(an - fd) > 2e-3 * (abs(fd) + ab[i, j] * 1e-2) + 1e-6:
                fails.append((i, j, an, fd))
    assert not fails, fails[:5]


def test_gradient_invariants(orc):
    """Colour gradient = sum over pixels of w * dL/dpred (closed form): with a
    single valid candidate per pixel it is 2 (c - t) summed over the pixel set."""
    w, h, n = 12, 10, 6
    sites = random_sites(w, h, n, 7, inactive_frac=0.0)
    c0 = np.full((w * h, 4), 0xFFFFFFFF, np.uint32)
    c1 = c0.copy()
    owner = np.arange(w * h) % n
    c0[:, 0] = owner
    rng = np.random.default_rng(8)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    acc, _, _, _ = orc.gradients(c0, c1, target, np.ones((h, w), np.float32), sites, iss_of(w, h), w, h)
    for i in range(n):
        t = target.reshape(-1, 3)[owner == i]
        np.testing.assert_allclose(acc[i, 4:7], 2.0 * (sites[i, 4:7] - t).sum(0), rtol=1e-5, atol=1e-6)
    # one candidate: softmax weight is constant 1, so all geometric gradients vanish
    assert np.all(acc[:, [0, 1, 2, 3, 7, 8, 9]] == 0)


# ------------------------------------------------------ candidate search

def _keys_f64(sites, x, y, S):
    s = sites.astype(np.float64)
    d = np.stack([x - s[:, 0], y - s[:, 1]], 1)
    proj = (d * s[:, 7:9]).sum(1)
    perp2 = np.maximum((d * d).sum(1) - proj ** 2, 0)
    d2 = np.maximum((np.exp(s[:, 9]) * proj ** 2 + np.exp(-s[:, 9]) * perp2) / S**2, 1e-8)
    key = np.maximum(np.exp(s[:, 2]), 1e-4) * (np.sqrt(d2) - s[:, 3] / S)
    key[s[:, 0] < 0] = np.inf
    return key


def test_vpt_finds_exact_top8_by_brute_force(orc):
    """With injects covering every site, a few VPT passes leave every cell
    holding the 8 smallest keys among all active sites, in key order
    (the brute-force top-8 oracle of backends/metal/tests/test_jump_schedule.swift:56-104)."""
    w, h, n = 24, 20, 20
    sites = random_sites(w, h, n, 11, inactive_frac=0.1)
    # the key uses fp16-rounded values: evaluate brute force on exactly those
    sites16 = orc.f16_bits_to_f32(orc.f32_to_f16_bits(sites))
    pk = orc.pack_sites(sites)
    c0, c1 = orc.init_candidates(n, 0, False, w, h)
    for i in range(6):
        step = (orc.jump_step(i, w, h) << 16) | i
        c0, c1 = orc.update_candidates(c0, c1, pk, step, iss_of(w, h), w, h, inject_count=200)
    ids = np.concatenate([c0, c1], 1)
    for y in range(h):
        for x in range(w):
            key = _keys_f64(sites16, x, y, max(w, h))
            order = np.argsort(key, kind="stable")
            top = [int(i) for i in order[:8] if np.isfinite(key[i])]
            got = [int(i) for i in ids[y * w + x] if i != 0xFFFFFFFF]
            assert sorted(got) == sorted(top)
            gk = key[got]
            assert np.all(np.diff(gk) >= -1e-6 * np.abs(gk[1:]))


def test_jfa_flood_nearest(orc):
    """A full JFA (seed + log2 floods) puts the true nearest active site first
    in nearly every cell (JFA is approximate by design)."""
    w, h, n = 64, 48, 60
    sites = random_sites(w, h, n, 12)
    c0, c1 = orc.clear_candidates(w, h)
    c0 = orc.jfa_seed(c0, sites, w, h)
    step = 32
    while step >= 1:
        c0 = orc.jfa_flood(c0, sites, step, iss_of(w, h), w, h)
        step //= 2
    hits = 0
    for y in range(h):
        for x in range(w):
            key = _keys_f64(sites, x, y, max(w, h))
            hits += int(c0[y * w + x, 0] == int(np.argmin(key)))
    assert hits >= 0.97 * w * h


def test_jfa_seed_collision_rule(orc):
    w, h = 4, 4
    sites = np.zeros((5, 10), np.float32)
    sites[:, 0:2] = [[1.2, 1.7], [1.9, 1.1], [3.0, 3.0], [1.5, 1.5], [-1, -1]]
    c0 = orc.jfa_seed(np.full((16, 4), 7, np.uint32), sites, w, h)
    assert c0[1 * 4 + 1, 0] == 3            # three sites share cell (1,1): the largest id
    assert c0[3 * 4 + 3, 0] == 2
    assert c0[0, 0] == 7 and np.all(c0[:, 1:] == 7)


# ------------------------------------------------------- tau / adam / scores

def test_tau_diffuse_closed_form(orc):
    w, h, n = 6, 4, 5
    sites = random_sites(w, h, n, 13, inactive_frac=0.0)
    sites[4, 0:2] = -1.0
    c0 = np.tile(np.array([0, 1, 1, 2], np.uint32), (w * h, 1))
    c1 = np.tile(np.array([4, 3, 0xFFFFFFFF, 9], np.uint32), (w * h, 1))
    rng = np.random.default_rng(14)
    graw = rng.normal(0, 1, n).astype(np.float32)
    gin = rng.normal(0, 1, n).astype(np.float32)
    lam = 0.25
    out = orc.tau_diffuse(c0, c1, sites, graw, gin, lam, w, h)
    for i in range(4):
        nb = [j for j in (0, 1, 2, 3) if j != i]     # unique, active, in range, not self
        ref = (graw[i] + lam * sum(gin[j] for j in nb)) / (1 + lam * len(nb))
        assert abs(out[i] - ref) <= 1e-6 * max(1.0, abs(ref))
    assert out[4] == gin[4]                           # inactive: pass-through
    assert np.array_equal(orc.tau_diffuse(c0, c1, sites, graw, gin, 0.0, w, h), gin)


def test_adam_textbook(orc):
    """Kingma & Ba Adam in float64: scalar groups per parameter, one shared
    second moment for pos and for the tangent-projected direction."""
    n, w, h = 50, 64, 48
    rng = np.random.default_rng(15)
    sites = random_sites(w, h, n, 16, inactive_frac=0.0)
    sites[:, 4:7] = rng.uniform(0.2, 0.8, (n, 3))
    sites[:, 3] = rng.uniform(2, 10, n)
    state = np.zeros((n, 20), np.float32)
    lrs = (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025)
    b1, b2, eps = 0.9, 0.999, 1e-8
    s64 = sites.astype(np.float64)
    m = np.zeros((n, 10)); v = np.zeros((n, 10))
    for t in range(1, 4):
        grads = rng.normal(0, 30, (10, n)).astype(np.float32)
        sites, state = orc.adam(sites, state, grads, lrs, b1, b2, eps, t, w, h)
        g = grads.T.astype(np.float64) / (w * h)
        dirv = s64[:, 7:9]
        gt = g[:, 7:9] - dirv * (g[:, 7:9] * dirv).sum(1, keepdims=True)
        gg = g.copy(); gg[:, 7:9] = gt
        m = b1 * m + (1 - b1) * gg
        for a, b in ((0, 2), (7, 9)):
            v[:, a] = b2 * v[:, a] + (1 - b2) * (gg[:, a:b] ** 2).sum(1)
            v[:, a + 1] = v[:, a]
        for j in (2, 3, 4, 5, 6, 9):
            v[:, j] = b2 * v[:, j] + (1 - b2) * gg[:, j] ** 2
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        lr = np.array([lrs[0], lrs[0], lrs[1], lrs[2], lrs[3], lrs[3], lrs[3], lrs[4], lrs[4], lrs[5]])
        s64 = s64 - lr * mh / (np.sqrt(vh) + eps)
        s64[:, 0] = np.clip(s64[:, 0], 0, w - 1); s64[:, 1] = np.clip(s64[:, 1], 0, h - 1)
        s64[:, 3] = np.clip(s64[:, 3], 1, 512); s64[:, 4:7] = np.clip(s64[:, 4:7], 0, 1)
        s64[:, 7:9] /= np.linalg.norm(s64[:, 7:9], axis=1, keepdims=True)
        s64[:, 9] = np.clip(s64[:, 9], -4, 4)
        np.testing.assert_allclose(sites, s64, rtol=2e-5, atol=2e-6)


def test_adam_first_step_known_answer(orc):
    n, w, h = 4, 10, 10
    sites = random_sites(w, h, n, 17, inactive_frac=0.0)
    sites[:, 0:2] = 5.0
    sites[:, 4:7] = 0.5
    state = np.zeros((n, 20), np.float32)
    grads = np.zeros((10, n), np.float32)
    grads[2] = [100.0, -100.0, 1.0, 0.0]            # log_tau: step = lr * sign(g) at t = 1
    out, _ = orc.adam(sites, state, grads, (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025), 0.9, 0.999, 1e-8, 1, w, h)
    d = sites[:, 2] - out[:, 2]
    np.testing.assert_allclose(d[:3], [0.01, -0.01, 0.01], rtol=1e-4)   # fp32 (1 - 0.999) is 1.3e-5 off
    assert d[3] == 0.0


def test_score_keys(orc):
    n = 200
    rng = np.random.default_rng(18)
    sites = random_sites(64, 64, n, 19, inactive_frac=0.1)
    mass = rng.exponential(3.0, n).astype(np.float32)
    energy = rng.exponential(0.5, n).astype(np.float32)
    energy[::13] = np.nan
    pairs = orc.densify_pairs(sites, mass, energy, 1.0, 0.7, n + 10)
    for i in range(n + 10):
        ok = i < n and sites[i, 0] >= 0 and mass[i] > 1.0 and np.isfinite(energy[i])
        if not ok:
            assert pairs[i].tolist() == [0xFFFFFFFF, 0xFFFFFFFF]
            continue
        sc = np.float32(max(energy[i], 0.0) / (max(float(mass[i]), 1e-8) ** 0.7))
        ref = 0xFFFFFFFF - int(sc.view(np.uint32))
        assert pairs[i, 1] == i and abs(int(pairs[i, 0]) - ref) <= 8      # pow within a few ulp
    rd = rng.normal(0, 1, n).astype(np.float32)
    pp = orc.prune_pairs(sites, rd, 0.5, n)
    for i in range(n):
        if sites[i, 0] < 0:
            assert pp[i].tolist() == [0xFFFFFFFF, 0xFFFFFFFF]
        else:
            assert pp[i, 0] == np.float32(max(rd[i] * np.float32(0.5), 0.0)).view(np.uint32) and pp[i, 1] == i


def test_split_closed_form(orc):
    """Sobel fallback on a horizontal ramp: split along x by clamp(sqrt(m)/2, 1.5, 48)."""
    w, h = 32, 20
    target = np.zeros((h, w, 3), np.float32)
    target[:] = (np.arange(w, dtype=np.float32) / w)[None, :, None]
    sites = np.zeros((4, 10), np.float32)
    sites[:, 0:2] = -1.0
    sites[0] = [10.3, 7.6, 5.0, 4.0, 0.1, 0.2, 0.3, 0.0, 1.0, 0.5]
    state = np.ones((4, 20), np.float32)
    stats = np.zeros((8, 4), np.float32)
    stats[0, 0] = 36.0                                          # mass -> offset 3
    s, st = orc.split_sites(sites, state, np.array([0], np.uint32), stats, 1, target, w, h)
    np.testing.assert_allclose(s[0, 0:2], [13.3, 7.6], atol=1e-5)
    np.testing.assert_allclose(s[1, 0:2], [7.3, 7.6], atol=1e-5)
    np.testing.assert_allclose(s[:2, 2], 4.75)
    np.testing.assert_allclose(s[:2, 3], 3.4)
    np.testing.assert_allclose(s[:2, 7:9], [[1, 0], [1, 0]], atol=1e-6)
    np.testing.assert_allclose(s[:2, 9], 0.4)
    np.testing.assert_allclose(s[0, 4:7], target[7, 13])
    assert np.all(st[:2] == 0) and np.all(st[2:] == 1)
    # a sentinel parent index (>= current count) is skipped
    s2, _ = orc.split_sites(sites, state, np.array([0xFFFFFFFF], np.uint32), stats, 1, target, w, h)
    assert np.array_equal(s2, sites)


# ------------------------------------------------------- schedule / solver

def test_bpp_solver_cross_check(orc):
    """The fit plan is restated twice (C oracle, C++ library); both must agree
    (solveTargetBpp exists twice in the reference too: sad_cuda.cu:962-1066,
    training_setup.swift:382-505)."""
    from sad_amd import config, trainer
    for (w, h, n0, bpp, iters) in [(1024, 1024, 16384, 1.0, 2000), (512, 384, 4000, 0.5, 1500),
                                   (256, 256, 2048, 2.0, 600), (800, 600, 3000, 0.0, 100)]:
        o = config.fit_options(n_sites=n0, target_bpp=bpp, iterations=iters, densify_freq=100, prune_freq=100,
                               max_sites=4 * n0)
        assert trainer.fit_plan(o, w, h) == orc.fit_plan(o, w, h)
    o = config.fit_options(n_sites=100, densify_percentile=0.0, prune_percentile=0.0)
    p = orc.fit_plan(o, 64, 64)
    assert p["densify_percentile"] == 0.0 and p["prune_percentile"] == 0.0


def test_oracle_fit_improves_and_is_deterministic(orc):
    from sad_amd import config, synth
    tgt = synth.make_target(64, 48, 3, n_ellipses=8)
    sites = synth.init_sites_uniform(tgt, 120, 3)
    opts = config.fit_options(iterations=40, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, max_sites=120, log_freq=39)
    a = orc.fit(tgt, sites, opts)
    b = orc.fit(tgt, sites, opts)
    assert a["log_psnr"][-1] > a["log_psnr"][0] + 3.0
    assert np.array_equal(a["sites"], b["sites"])
