"""GPU parity of every kernel entry point of libsad_hip.so against the CPU oracle.

Bars: candidate lists, packs, sort permutations, score keys, Hilbert tables and
densify/prune decisions bit-exact; render within 1e-5 max abs; gradients,
removal deltas and site statistics within 1e-4 relative, stated for sums whose
terms may cancel as |gpu - ref| <= 1e-4 |ref| + 1e-5 sum|terms| (ref = the
oracle's double-precision sum; a plain fp32 sequential sum of the same terms
already deviates by up to ~7e-7 sum|terms|); Adam and tau diffusion bit-exact
(the same fp32 operation sequence as the oracle).
"""
import numpy as np
import pytest

from tests.scenes import converged_field, hilbert_tables, iss_of, random_sites
from oracle import orc as orc_mod
from tests.tolerance import Rel

pytestmark = pytest.mark.gpu


def _fields(gpu, c0, c1):
    return gpu.put(c0), gpu.put(c1)


# ------------------------------------------------------------ candidate field

@pytest.mark.parametrize("w,h,n,seed", [(64, 48, 300, 1), (257, 131, 5000, 2)])
def test_init_clear_seed(gpu, orc, w, h, n, seed):
    sites = random_sites(w, h, n, seed)
    # many sites per cell on the small grid: the seed collision rule is exercised
    c0, c1 = gpu.zeros((w * h, 4), np.uint32), gpu.zeros((w * h, 4), np.uint32)
    gpu.call("sadh_init_candidates", c0, c1, n, 77, 0, w, h)
    r0, r1 = orc.init_candidates(n, 77, False, w, h)
    assert np.array_equal(gpu.get(c0, np.uint32), r0) and np.array_equal(gpu.get(c1, np.uint32), r1)
    s = gpu.put(sites)
    gpu.call("sadh_jfa_seed", c0, s, n, w, h, 1)
    assert np.array_equal(gpu.get(c0, np.uint32), orc.jfa_seed(r0, sites, w, h))
    gpu.call("sadh_clear_candidates", c0, c1, w, h)
    assert np.all(gpu.get(c0, np.uint32) == 0xFFFFFFFF) and np.all(gpu.get(c1, np.uint32) == 0xFFFFFFFF)


def test_init_per_pixel(gpu, orc):
    w, h = 40, 30
    c0, c1 = gpu.zeros((w * h, 4), np.uint32), gpu.zeros((w * h, 4), np.uint32)
    gpu.call("sadh_init_candidates", c0, c1, w * h, 0, 1, w, h)
    r0, r1 = orc.init_candidates(w * h, 0, True, w, h)
    assert np.array_equal(gpu.get(c0, np.uint32), r0) and np.array_equal(gpu.get(c1, np.uint32), r1)


def test_pack_bitexact(gpu, orc):
    sites = random_sites(300, 200, 4000, 3)
    sites[::7, 8] = np.float32(3e-7)      # fp16 subnormal direction components
    sites[::11, 7] = np.float32(-2e-6)
    s = gpu.put(sites)
    out = gpu.zeros((4000, 8), np.uint16)
    gpu.call("sadh_pack_candidate_sites", s, out, 4000)
    assert np.array_equal(gpu.get(out, np.uint16), orc.pack_sites(sites))


@pytest.mark.parametrize("w,h,n", [(96, 80, 700), (256, 256, 2048)])
def test_jfa_flood_bitexact(gpu, orc, w, h, n):
    sites = random_sites(w, h, n, 4)
    iss = iss_of(w, h)
    c0, c1 = orc.clear_candidates(w, h)
    c0 = orc.jfa_seed(c0, sites, w, h)
    s = gpu.put(sites)
    cur = c0
    step = 1
    while step < max(w, h):
        step <<= 1
    step //= 2
    while step >= 1:
        gi, go = gpu.put(cur), gpu.zeros((w * h, 4), np.uint32)
        gpu.call("sadh_jfa_flood", gi, go, s, n, step, iss, w, h, 1, w, h)
        ref = orc.jfa_flood(cur, sites, step, iss, w, h)
        got = gpu.get(go, np.uint32)
        assert np.array_equal(got, ref), f"flood step {step}: {np.sum(got != ref)} ids differ"
        cur = ref
        step //= 2


@pytest.mark.parametrize("w,h,n,passes,probes", [(128, 96, 1500, 10, 0), (256, 256, 2048, 14, 3), (300, 170, 9000, 6, 0)])
def test_vpt_bitexact(gpu, orc, w, h, n, passes, probes):
    sites = random_sites(w, h, n, 5)
    sites[::5, 8] = np.float32(1e-7)
    iss = iss_of(w, h)
    c0, c1 = orc.init_candidates(n, 0, False, w, h)
    c0 = orc.jfa_seed(c0, sites, w, h)
    pk = orc.pack_sites(sites)
    ho, hp = hilbert_tables(orc, sites, w, h)
    gpk, gho, ghp = gpu.put(pk), gpu.put(ho), gpu.put(hp)
    for i in range(passes):
        step = (orc.jump_step(i, w, h) << 16) | i
        o0, o1 = gpu.zeros((w * h, 4), np.uint32), gpu.zeros((w * h, 4), np.uint32)
        i0, i1 = _fields(gpu, c0, c1)
        gpu.call("sadh_update_candidates", i0, i1, o0, o1, gpk, n, step, iss, 0, 64.0, probes, 4, gho, ghp, 1, 32,
                 1, w, h, w, h)
        r0, r1 = orc.update_candidates(c0, c1, pk, step, iss, w, h, radius_probes=probes, inject_count=4,
                                       hilbert_order=ho, hilbert_pos=hp, hilbert_probes=1, hilbert_window=32)
        g0, g1 = gpu.get(o0, np.uint32), gpu.get(o1, np.uint32)
        assert np.array_equal(g0, r0) and np.array_equal(g1, r1), \
            f"pass {i}: {np.sum(g0 != r0) + np.sum(g1 != r1)} ids differ"
        c0, c1 = r0, r1


def test_hilbert_tables(gpu, orc):
    w, h, n = 512, 300, 7000
    sites = random_sites(w, h, n, 6)
    bits = 9
    padded = (n + 1023) // 1024 * 1024
    s = gpu.put(sites)
    pairs = gpu.zeros((padded, 2), np.uint32)
    gpu.call("sadh_hilbert_pairs", s, pairs, n, padded, w, h, bits)
    ref = orc.hilbert_pairs(sites, padded, w, h, bits)
    assert np.array_equal(gpu.get(pairs, np.uint32), ref)
    scratch = gpu.zeros((padded, 2), np.uint32)
    hist = gpu.zeros((256 * (padded // 1024),), np.uint32)
    gpu.call("sadh_radix_sort_pairs", pairs, scratch, hist, hist, padded, 1 << 18)
    order, pos = gpu.zeros((n,), np.uint32), gpu.zeros((n,), np.uint32)
    gpu.call("sadh_hilbert_order", pairs, order, pos, n)
    ro, rp = hilbert_tables(orc, sites, w, h)
    assert np.array_equal(gpu.get(order, np.uint32), ro) and np.array_equal(gpu.get(pos, np.uint32), rp)


# -------------------------------------------------------------- radix sort

@pytest.mark.parametrize("n,kind", [(1, "full"), (1000, "full"), (1024, "dup"), (1025, "full"), (65536, "16bit"),
                                    (20000, "score"), (40000, "reverse"), (262144, "full"), (300001, "score"),
                                    (70000, "sorted"), (70000, "reverse")])
def test_radix_sort_bitexact(gpu, orc, n, kind):
    rng = np.random.default_rng(n)
    if kind == "full":
        keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    elif kind == "dup":
        keys = rng.integers(0, 4, n).astype(np.uint32)
    elif kind == "16bit":
        keys = rng.integers(0, 2**16, n).astype(np.uint32)
    elif kind == "score":
        sc = np.abs(rng.normal(0, 1, n)).astype(np.float32)
        keys = (0xFFFFFFFF - sc.view(np.uint32)).astype(np.uint32)
        keys[rng.random(n) < 0.2] = 0xFFFFFFFF
    elif kind == "sorted":
        keys = np.sort(rng.integers(0, 2**32, n, dtype=np.uint64)).astype(np.uint32)
    else:
        keys = np.sort(rng.integers(0, 2**32, n, dtype=np.uint64))[::-1].astype(np.uint32)
    pairs = np.stack([keys, np.arange(n, dtype=np.uint32)], 1)
    mk = (1 << 16) if kind == "16bit" else 0xFFFFFFFF
    d = gpu.put(pairs)
    scratch = gpu.zeros((n, 2), np.uint32)
    hist = gpu.zeros((256 * ((n + 1023) // 1024),), np.uint32)
    gpu.call("sadh_radix_sort_pairs", d, scratch, hist, hist, n, mk)
    got = gpu.get(d, np.uint32)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(got, pairs[order])
    assert np.array_equal(got, orc.radix_sort(pairs, mk))


# ------------------------------------------------------------ pixel kernels

def _scene(orc, w, h, n, seed, passes=10):
    sites = random_sites(w, h, n, seed)
    c0, c1 = converged_field(orc, sites, w, h, passes=passes)
    return sites, c0, c1


@pytest.mark.parametrize("w,h,n", [(128, 96, 600), (333, 257, 4000)])
def test_render_parity(gpu, orc, w, h, n):
    sites, c0, c1 = _scene(orc, w, h, n, 8)
    iss = iss_of(w, h)
    out = gpu.zeros((h, w, 3))
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_render_voronoi", g0, g1, out, gpu.put(sites), iss, n, w, h, w, h)
    ref = orc.render(c0, c1, sites, iss, w, h)
    err = np.abs(gpu.get(out) - ref).max()
    assert err <= 1e-5, err


def test_render_edge_cases(gpu, orc):
    w, h, n = 48, 40, 30
    sites = random_sites(w, h, n, 9, inactive_frac=0.0)
    sites[5:, 0:2] = -1.0            # most sites inactive
    c0, c1 = orc.init_candidates(n, 3, False, w, h)
    c1[:] = 0xFFFFFFFF               # empty second plane
    c0[::3] = 0xFFFFFFFF             # all-invalid cells -> black
    iss = iss_of(w, h)
    out = gpu.zeros((h, w, 3))
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_render_voronoi", g0, g1, out, gpu.put(sites), iss, n, w, h, w, h)
    ref = orc.render(c0, c1, sites, iss, w, h)
    assert np.abs(gpu.get(out) - ref).max() <= 1e-5


def _rel_ok(got, ref, ab, parts=None):
    """Rel with the fixed-point resolution of `parts` partials per site (rows of got)"""
    atol = 0.0
    if parts is not None:
        atol = np.asarray(parts, np.float64) * orc_mod.FIX_HALF_STEP
        if np.ndim(got) == 2:
            atol = atol[:, None]
    return Rel(got, ref, ab, atol)


def _same_bits(got, ref):
    """float32 arrays equal bit for bit (the oracle forms the library's exact per-tile sums)"""
    a = np.ascontiguousarray(got, np.float32)
    b = np.ascontiguousarray(ref, np.float32)
    bad = int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))
    assert a.shape == b.shape and bad == 0, f"{bad} of {a.size} values differ"


@pytest.mark.parametrize("removal", [False, True])
@pytest.mark.parametrize("w,h,n,passes", [(160, 128, 900, 12), (256, 256, 2048, 0)])
def test_gradients_parity(gpu, orc, w, h, n, passes, removal):
    sites, c0, c1 = _scene(orc, w, h, n, 10, passes=passes)
    rng = np.random.default_rng(11)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    mask[rng.random((h, w)) < 0.1] = 0.0
    iss = iss_of(w, h)
    grads = [gpu.zeros((n,)) for _ in range(10)]
    rd = gpu.zeros((n,))
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_compute_gradients_tiled", g0, g1, gpu.put(target), gpu.put(target), gpu.put(mask), *grads,
             gpu.put(sites), iss, n, rd, int(removal), 0.0, w, h, w, h)
    acc, ab, racc, rab = orc.gradients(c0, c1, target, mask, sites, iss, w, h, compute_removal=removal)
    got = np.stack([gpu.get(g) for g in grads], 1)
    gf, rf, parts = orc.gradients_fix(c0, c1, target, mask, sites, iss, w, h, compute_removal=removal)
    assert _rel_ok(got, acc, ab, parts) <= 1.0
    _same_bits(got, gf)
    if removal:
        assert _rel_ok(gpu.get(rd), racc, rab, parts) <= 1.0
        _same_bits(gpu.get(rd), rf)


@pytest.mark.parametrize("removal", [False, True])
@pytest.mark.parametrize("w,h,n,passes", [(160, 128, 900, 12), (256, 256, 2048, 0)])
def test_gradients_aos_parity(gpu, orc, w, h, n, passes, removal):
    """AoS [N][16] gradient lines against the oracle (passes=0: random field,
    nearly every pair overflows the tile hash)."""
    sites, c0, c1 = _scene(orc, w, h, n, 10, passes=passes)
    rng = np.random.default_rng(11)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    mask[rng.random((h, w)) < 0.1] = 0.0
    iss = iss_of(w, h)
    lines = gpu.zeros((n, 16))
    rd = gpu.zeros((n,))
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_compute_gradients_tiled_aos", g0, g1, gpu.put(target), gpu.put(mask), lines, gpu.put(sites), iss,
             n, rd, int(removal), w, h, w, h)
    acc, ab, rdr, rda = orc.gradients(c0, c1, target, mask, sites, iss, w, h, compute_removal=removal)
    got = gpu.get(lines)
    assert np.all(got[:, 10:] == 0)
    gf, rf, parts = orc.gradients_fix(c0, c1, target, mask, sites, iss, w, h, compute_removal=removal)
    assert _rel_ok(got[:, :10], acc, ab, parts) <= 1.0
    _same_bits(got[:, :10], gf)
    if removal:
        assert _rel_ok(gpu.get(rd), rdr, rda, parts) <= 1.0
        _same_bits(gpu.get(rd), rf)


def test_render_backward_parity(gpu, orc):
    w, h, n = 144, 112, 800
    sites, c0, c1 = _scene(orc, w, h, n, 12)
    rng = np.random.default_rng(13)
    go = rng.normal(0, 1, (h, w, 3)).astype(np.float32)
    iss = iss_of(w, h)
    out = gpu.zeros((n, 10))
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_render_backward", g0, g1, gpu.put(go), out, gpu.put(sites), iss, n, w, h, w, h)
    acc, ab = orc.render_backward(c0, c1, go, sites, iss, w, h)
    gf, parts = orc.render_backward_fix(c0, c1, go, sites, iss, w, h)
    assert _rel_ok(gpu.get(out), acc, ab, parts) <= 1.0
    _same_bits(gpu.get(out), gf)


def test_site_stats_parity(gpu, orc):
    w, h, n = 200, 150, 1200
    sites, c0, c1 = _scene(orc, w, h, n, 14)
    rng = np.random.default_rng(15)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    iss = iss_of(w, h)
    st = [gpu.zeros((n,)) for _ in range(8)]
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_compute_site_stats_simple", g0, g1, gpu.put(target), gpu.put(mask), gpu.put(sites), iss, n,
             *st, w, h, w, h)
    ref = orc.site_stats(c0, c1, target, mask, sites, iss, w, h)
    got = np.stack([gpu.get(x) for x in st], 1)
    err = np.abs(got - ref) / (np.abs(ref) + 1e-3 * np.abs(ref).max(0) + 1e-20)
    assert err.max() <= 1e-4
    _same_bits(got, orc.site_stats_fix(c0, c1, target, mask, sites, iss, w, h)[0])


# ------------------------------------------------------------- site kernels

def test_tau_diffuse_parity(gpu, orc):
    w, h, n = 160, 120, 1500
    sites, c0, c1 = _scene(orc, w, h, n, 16)
    rng = np.random.default_rng(17)
    graw = rng.normal(0, 1, n).astype(np.float32)
    gin = rng.normal(0, 1, n).astype(np.float32)
    out = gpu.zeros((n,))
    g0, g1 = _fields(gpu, c0, c1)
    gpu.call("sadh_tau_diffuse", g0, g1, gpu.put(sites), gpu.put(graw), gpu.put(gin), out, n, 0.037, w, h, 1)
    ref = orc.tau_diffuse(c0, c1, sites, graw, gin, 0.037, w, h)
    assert np.array_equal(gpu.get(out), ref)


@pytest.mark.parametrize("clear", [False, True])
def test_adam_parity(gpu, orc, clear):
    w, h, n = 300, 200, 5000
    rng = np.random.default_rng(18)
    sites = random_sites(w, h, n, 18)
    state = np.abs(rng.normal(0, 1e-3, (n, 20))).astype(np.float32)
    state[:, 18:] = 0.0
    grads = rng.normal(0, 50, (10, n)).astype(np.float32)
    lrs = (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025)
    gs, gst = gpu.put(sites), gpu.put(state)
    gg = [gpu.put(grads[j]) for j in range(10)]
    name = "sadh_adam_update_clear" if clear else "sadh_adam_update"
    gpu.call(name, gs, gst, *gg, *lrs, 0.9, 0.999, 1e-8, 7, n, w, h)
    rs, rst = orc.adam(sites, state, grads, lrs, 0.9, 0.999, 1e-8, 7, w, h)
    got_s, got_st = gpu.get(gs), gpu.get(gst)
    assert np.array_equal(got_s, rs)
    assert np.array_equal(got_st, rst)
    gz = np.stack([gpu.get(g) for g in gg])
    if clear:
        assert np.all(gz == 0)
    else:
        assert np.array_equal(gz, grads)


@pytest.mark.parametrize("w,h,n,lam", [(160, 120, 1500, 0.037), (300, 200, 70000, 0.05), (64, 64, 300, 0.0)])
def test_tau_diffuse_adam_fused(gpu, orc, w, h, n, lam):
    """The fused tail (tau passes, the last one fused with Adam + clear) against the five
    separate launches: bit-identical sites, state and cleared gradients; the
    separate chain is itself pinned to the oracle by the tests above."""
    sites, c0, c1 = _scene(orc, w, h, n, 16, passes=6)
    rng = np.random.default_rng(31)
    state = np.abs(rng.normal(0, 1e-3, (n, 20))).astype(np.float32)
    state[:, 18:] = 0.0
    grads = rng.normal(0, 50, (10, n)).astype(np.float32)
    lrs = (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025)
    g0, g1 = _fields(gpu, c0, c1)
    # separate: g2 -> tmp -> g2 -> tmp -> g2 with raw = the first input, then Adam with clear
    s_a, st_a = gpu.put(sites), gpu.put(state)
    ga = [gpu.put(grads[j]) for j in range(10)]
    raw, tmp = gpu.put(grads[2]), gpu.zeros((n,))
    gpu.call("sadh_tau_diffuse", g0, g1, s_a, raw, ga[2], tmp, n, lam, w, h, 1)
    gpu.call("sadh_tau_diffuse", g0, g1, s_a, raw, tmp, ga[2], n, lam, w, h, 1)
    gpu.call("sadh_tau_diffuse", g0, g1, s_a, raw, ga[2], tmp, n, lam, w, h, 1)
    gpu.call("sadh_tau_diffuse", g0, g1, s_a, raw, tmp, ga[2], n, lam, w, h, 1)
    gpu.call("sadh_adam_update_clear", s_a, st_a, *ga, *lrs, 0.9, 0.999, 1e-8, 7, n, w, h)
    # fused
    s_b, st_b = gpu.put(sites), gpu.put(state)
    gb = [gpu.put(grads[j]) for j in range(10)]
    gpu.call("sadh_tau_diffuse_adam", g0, g1, s_b, st_b, *gb, gpu.zeros((10 * n,)), lam, w, h, 1, *lrs, 0.9, 0.999, 1e-8,
             7, n, w, h)
    assert np.array_equal(gpu.get(s_b), gpu.get(s_a))
    assert np.array_equal(gpu.get(st_b), gpu.get(st_a))
    assert all(np.all(gpu.get(g) == 0) for g in gb)


def test_score_pairs_and_edits(gpu, orc):
    w, h, n, cap = 256, 256, 3000, 4096
    sites = np.zeros((cap, 10), np.float32)
    sites[:, 0:2] = -1.0
    sites[:n] = random_sites(w, h, n, 19)
    rng = np.random.default_rng(20)
    mass = rng.exponential(3.0, cap).astype(np.float32)
    energy = rng.exponential(0.5, cap).astype(np.float32)
    energy[::97] = np.inf
    rdelta = rng.normal(0, 2, cap).astype(np.float32)
    gs = gpu.put(sites)
    pairs = gpu.zeros((cap, 2), np.uint32)
    gpu.call("sadh_compute_densify_score_pairs", gs, gpu.put(mass), gpu.put(energy), pairs, n, 1.0, 0.7, cap)
    ref_d = orc.densify_pairs(sites[:n], mass, energy, 1.0, 0.7, cap)
    assert np.array_equal(gpu.get(pairs, np.uint32), ref_d)
    gpu.call("sadh_compute_prune_score_pairs", gs, gpu.put(rdelta), pairs, n, 1.0 / 65536, cap)
    ref_p = orc.prune_pairs(sites[:n], rdelta, 1.0 / 65536, cap)
    assert np.array_equal(gpu.get(pairs, np.uint32), ref_p)
    # sort the densify pairs, take the first k (the split decision) and split
    dp = gpu.put(ref_d)
    scratch = gpu.zeros((cap, 2), np.uint32)
    hist = gpu.zeros((256 * 4,), np.uint32)
    gpu.call("sadh_radix_sort_pairs", dp, scratch, hist, hist, cap, 0xFFFFFFFF)
    srt = orc.radix_sort(ref_d)
    assert np.array_equal(gpu.get(dp, np.uint32), srt)
    k = 500
    idx = gpu.zeros((k,), np.uint32)
    gpu.call("sadh_write_split_indices", dp, idx, k)
    ref_idx = srt[:k, 1]
    assert np.array_equal(gpu.get(idx, np.uint32), ref_idx)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    stats = np.abs(rng.normal(0, 1, (8, cap))).astype(np.float32)
    stats[3] = stats[2] * rng.uniform(0, w - 1, cap)
    stats[4] = stats[2] * rng.uniform(0, h - 1, cap)
    stats[5] = stats[2] * rng.uniform(0, w - 1, cap) ** 2
    stats[6] = stats[3] * stats[4] / np.maximum(stats[2], 1e-3)
    stats[7] = stats[2] * rng.uniform(0, h - 1, cap) ** 2
    stats[2, ::3] = 0.0                       # Sobel fallback branch
    state = rng.normal(0, 1, (cap, 20)).astype(np.float32)
    gst = gpu.put(state)
    gstats = [gpu.put(stats[j]) for j in range(8)]
    gpu.call("sadh_split_sites", gs, gst, idx, k, gstats[0], gstats[2], gstats[3], gstats[4], gstats[5],
             gstats[6], gstats[7], n, gpu.put(target), w, h)
    rs, rst = orc.split_sites(sites, state, ref_idx, stats, n, target, w, h)
    got = gpu.get(gs)
    np.testing.assert_allclose(got, rs, rtol=1e-5, atol=1e-5)
    assert np.array_equal(got[:, 0] >= 0, rs[:, 0] >= 0)
    assert np.array_equal(gpu.get(gst), rst)
    # prune the 100 lowest removal deltas
    pp = gpu.put(ref_p)
    gpu.call("sadh_radix_sort_pairs", pp, scratch, hist, hist, cap, 0xFFFFFFFF)
    pidx = gpu.zeros((100,), np.uint32)
    gpu.call("sadh_write_split_indices", pp, pidx, 100)
    gpu.call("sadh_prune_sites_by_index", gs, pidx, 100)
    rp = orc.radix_sort(ref_p)[:100, 1]
    assert np.array_equal(gpu.get(pidx, np.uint32), rp)
    after = gpu.get(gs)
    assert np.all(after[rp[rp != 0xFFFFFFFF], 0] == -1.0)


def test_metrics(gpu, orc):
    w, h = 123, 77
    rng = np.random.default_rng(21)
    a = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    b = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    m = (rng.random((h, w)) > 0.3).astype(np.float32)
    acc = gpu.zeros((1,))
    gpu.call("sadh_compute_psnr", gpu.put(a), gpu.put(b), gpu.put(m), acc, w, h)
    ref = orc.sse(a, b, m)
    assert abs(float(gpu.get(acc)[0]) - ref) <= 1e-5 * ref
    _same_bits(gpu.get(acc)[:1], np.array([orc.sse_fix(a, b, m)], np.float32))
    sites = random_sites(w, h, 999, 22, inactive_frac=0.3)
    cnt = gpu.zeros((1,), np.uint32)
    gpu.call("sadh_count_active_sites", gpu.put(sites), cnt, 999)
    assert int(gpu.get(cnt, np.uint32)[0]) == int(np.sum(sites[:, 0] >= 0))
