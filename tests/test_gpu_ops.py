"""GPU parity of `torch.ops.sad` and of the autograd render against the
oracle and against FP64 central differences (the differentiable-render
pattern: 64x64 / 256 sites, as backends/pytorch/autograd_demo.py:16-29)."""
import numpy as np
import pytest

from tests.fd64 import FD_STEPS, central_differences, fd_problem, render_f64
from tests.scenes import converged_field, iss_of, random_sites
from tests.tolerance import Rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sad(gpu):
    from sad_amd import ops  # noqa: F401
    return gpu.torch.ops.sad


def _u32(gpu, t):
    return gpu.get(t).view(np.uint32)


def test_init_and_render(gpu, orc, sad):
    w, h, n = 96, 72, 400
    sites = random_sites(w, h, n, 21)
    c0, c1 = sad.init_candidates(w, h, n, 5)
    r0, r1 = orc.init_candidates(n, 5, False, w, h)
    assert np.array_equal(_u32(gpu, c0), r0) and np.array_equal(_u32(gpu, c1), r1)
    out = sad.render_sad(c0, c1, gpu.put(sites), iss_of(w, h), w, h)
    assert tuple(out.shape) == (h, w, 3)
    assert np.abs(gpu.get(out) - orc.render(r0, r1, sites, iss_of(w, h), w, h)).max() <= 1e-5
    # per-pixel mode (site_count == width * height)
    p0, p1 = sad.init_candidates(16, 8, 128, 3)
    q0, q1 = orc.init_candidates(128, 3, True, 16, 8)
    assert np.array_equal(_u32(gpu, p0), q0) and np.array_equal(_u32(gpu, p1), q1)


def test_vpt_pass_bit_exact(gpu, orc, sad):
    w, h, n = 80, 64, 500
    sites = random_sites(w, h, n, 22)
    c0, c1 = sad.init_candidates(w, h, n, 9)
    r0, r1 = orc.init_candidates(n, 9, False, w, h)
    gs = gpu.put(sites)
    pk = orc.pack_sites(sites)
    for jump in (32, 16, 8, 4, 2, 1, 1):
        c0, c1 = sad.vpt_pass(c0, c1, gs, iss_of(w, h), w, h, jump, 0, 64.0, 4, 123)
        r0, r1 = orc.update_candidates(r0, r1, pk, jump << 16, iss_of(w, h), w, h, inject_count=4)
    assert np.array_equal(_u32(gpu, c0), r0) and np.array_equal(_u32(gpu, c1), r1)


def test_autograd_render_matches_oracle_and_fd64(gpu, orc):
    """The differentiable render's backward vs the oracle backward (bit for bit
    against its tile-ordered sums; 1e-4 relative, with the cancellation
    allowance, against its float64 sums) and vs FP64 central differences."""
    from sad_amd.differentiable import CandidateField, inv_scale_sq, render
    torch = gpu.torch
    w, h = 64, 64
    sites, go = fd_problem(w=w, h=h)
    c0, c1 = converged_field(orc, sites, w, h, passes=10)
    iss = iss_of(w, h)
    assert inv_scale_sq(w, h) == iss
    ts = gpu.put(sites).requires_grad_(True)
    field = CandidateField(w, h, gpu.put(c0), gpu.put(c1))
    out = render(ts, field)
    assert np.abs(gpu.get(out) - orc.render(c0, c1, sites, iss, w, h)).max() <= 1e-5
    (out * gpu.put(go)).sum().backward()
    got = gpu.get(ts.grad)
    acc, ab = orc.render_backward(c0, c1, go, sites, iss, w, h)
    exact, parts = orc.render_backward_fix(c0, c1, go, sites, iss, w, h)
    assert np.array_equal(got.view(np.uint32), exact.view(np.uint32))
    assert Rel(got, acc, ab, parts[:, None] * orc.FIX_HALF_STEP) <= 1.0
    R = go.astype(np.float64)

    def loss(s):
        return float(np.sum(render_f64(c0, c1, s, w, h) * R))

    rows = list(np.random.default_rng(4).choice(len(sites), 16, replace=False))
    fd = central_differences(loss, sites, rows, list(range(10)), FD_STEPS)
    bound = 2e-3 * (np.abs(fd) + 1e-2 * ab[rows]) + 1e-6
    bad = np.argwhere(np.abs(got[rows] - fd) > bound)
    assert len(bad) == 0, [(rows[i], j, got[rows][i, j], fd[i, j]) for i, j in bad[:5]]
    assert torch.isfinite(ts.grad).all()


def test_training_ops_vs_oracle(gpu, orc, sad):
    """compute_gradients_tiled, tau_diffuse and adam_update through the torch
    schema, one iteration, against the oracle."""
    torch = gpu.torch
    w, h, n = 128, 96, 700
    sites = random_sites(w, h, n, 23)
    c0, c1 = converged_field(orc, sites, w, h, passes=10)
    iss = iss_of(w, h)
    rng = np.random.default_rng(24)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    mask = np.ones((h, w), np.float32)
    g0, g1, gt, gm, gs = gpu.put(c0), gpu.put(c1), gpu.put(target), gpu.put(mask), gpu.put(sites)
    grads = [torch.zeros(n, device=gpu.dev) for _ in range(10)]
    removal = torch.zeros(n, device=gpu.dev)
    sad.compute_gradients_tiled(g0, g1, gt, gt, gm, gs, *grads, removal, iss, n, 1, 0.0, w, h)
    acc, ab, rd, ra = orc.gradients(c0, c1, target, mask, sites, iss, w, h, compute_removal=True)
    got = np.stack([gpu.get(g) for g in grads], 1)
    assert Rel(got, acc, ab) <= 1.0
    assert Rel(gpu.get(removal), rd, ra) <= 1.0
    # tau diffusion and Adam consume the GPU gradients, so both sides start from the same inputs
    graw = got[:, 2].copy()
    out = torch.zeros(n, device=gpu.dev)
    sad.tau_diffuse(g0, g1, gs, grads[2].clone(), grads[2], out, n, 0.05, 1, w, h)
    ref_tau = orc.tau_diffuse(c0, c1, sites, graw, graw, 0.05, w, h)
    assert np.array_equal(gpu.get(out), ref_tau)
    adam = torch.zeros((n, 20), device=gpu.dev)
    lrs = (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025)
    sad.adam_update(gs, adam, *grads, *lrs, 0.9, 0.999, 1e-8, 1, w, h)
    ref_sites, ref_state = orc.adam(sites, np.zeros((n, 20), np.float32), got.T.astype(np.float32), lrs, 0.9, 0.999,
                                    1e-8, 1, w, h)
    assert np.array_equal(gpu.get(gs), ref_sites)
    assert np.array_equal(gpu.get(adam), ref_state)


def test_radix_sort_and_hilbert_ops(gpu, orc, sad):
    torch = gpu.torch
    rng = np.random.default_rng(25)
    n = 5000
    pairs = np.stack([rng.integers(0, 1 << 20, n), np.arange(n)], 1).astype(np.uint32)
    t = gpu.put(pairs)
    sad.radix_sort_pairs(t, 0xFFFFFFFF)
    assert np.array_equal(_u32(gpu, t), pairs[np.argsort(pairs[:, 0], kind="stable")])
    w, h, ns = 200, 150, 900
    sites = random_sites(w, h, ns, 26)
    padded = 1024
    hp = torch.zeros((padded, 2), dtype=torch.int32, device=gpu.dev)
    sad.build_hilbert_pairs(gpu.put(sites), hp, ns, padded, w, h, 8)
    assert np.array_equal(_u32(gpu, hp), orc.hilbert_pairs(sites, padded, w, h, 8))


def test_ops_reject_bad_shapes(gpu, sad):
    torch = gpu.torch
    c = torch.zeros((16, 4), dtype=torch.int32, device=gpu.dev)
    with pytest.raises(RuntimeError, match="sites must be"):
        sad.render_sad(c, c, torch.zeros((2, 9), device=gpu.dev), 0.1, 4, 4)
    with pytest.raises(RuntimeError, match="cand0 size must match"):
        sad.render_sad(c, c, torch.zeros((2, 10), device=gpu.dev), 0.1, 5, 4)


def test_candidate_field_seeded_and_refined(gpu, orc):
    """CandidateField.seeded is init + JFA seed, refine() runs VPT passes on the
    jump schedule; both equal the oracle's fields bit for bit."""
    from sad_amd.differentiable import CandidateField, jump_schedule
    w, h, n = 80, 64, 500
    sites = random_sites(w, h, n, 23)
    ts = gpu.put(sites)
    f = CandidateField.seeded(ts, w, h, seed=3)
    r0, r1 = orc.init_candidates(n, 3, False, w, h)
    r0 = orc.jfa_seed(r0, sites, w, h)
    assert np.array_equal(_u32(gpu, f.cand0), r0) and np.array_equal(_u32(gpu, f.cand1), r1)
    assert jump_schedule(w, h, 8) == [orc.jump_step(i, w, h) for i in range(8)]
    f.refine(ts, passes=5)
    pk = orc.pack_sites(sites)
    for jump in jump_schedule(w, h, 5):
        r0, r1 = orc.update_candidates(r0, r1, pk, jump << 16, iss_of(w, h), w, h, inject_count=4)
    assert f.passes_done == 5
    assert np.array_equal(_u32(gpu, f.cand0), r0) and np.array_equal(_u32(gpu, f.cand1), r1)
