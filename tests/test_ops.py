"""CPU checks of the torch operator layer and of the FP64 finite-difference pin
of the render backward at the 64x64 / 256-site autograd shape."""
import numpy as np
import pytest
import torch

from tests.fd64 import FD_STEPS, central_differences, fd_problem, render_f64
from tests.scenes import converged_field, iss_of, random_sites

# The operator names of backends/pytorch/torch_ext/torch_binding.cpp:7-30, in order.
REFERENCE_OPS = [
    "render_sad", "render_sad_padded", "init_candidates", "vpt_pass", "render_sad_backward", "clear_buffer",
    "pack_candidate_sites", "update_candidates_compact", "jfa_clear", "jfa_seed", "jfa_flood",
    "compute_gradients_tiled", "tau_diffuse", "adam_update", "compute_site_stats_tiled",
    "compute_densify_score_pairs", "compute_prune_score_pairs", "radix_sort_pairs", "write_split_indices",
    "split_sites", "prune_sites", "build_hilbert_pairs", "write_hilbert_order", "init_gradient_weighted",
]


def test_all_reference_ops_registered():
    from sad_amd import ops
    assert ops.OP_NAMES == REFERENCE_OPS
    for name in REFERENCE_OPS:
        schema = getattr(torch.ops.sad, name).default._schema
        assert schema.name == f"sad::{name}"


def test_op_argument_lists():
    from sad_amd import ops  # noqa: F401
    s = torch.ops.sad.compute_gradients_tiled.default._schema
    assert [a.name for a in s.arguments][:6] == ["cand0", "cand1", "target", "rendered", "mask", "sites"]
    assert [a.name for a in s.arguments][-6:] == ["inv_scale_sq", "site_count", "compute_removal", "ssim_weight",
                                                  "cand_width", "cand_height"]
    s = torch.ops.sad.vpt_pass.default._schema
    assert [a.name for a in s.arguments] == ["cand0", "cand1", "sites", "inv_scale_sq", "width", "height", "jump",
                                             "radius_probes", "radius_scale", "inject_count", "seed"]
    assert len(s.returns) == 2


def test_cpu_tensors_have_no_kernel():
    """No CPU fallback: the dispatcher refuses CPU tensors."""
    from sad_amd import ops  # noqa: F401
    c = torch.zeros((16, 4), dtype=torch.int32)
    with pytest.raises(NotImplementedError):
        torch.ops.sad.render_sad(c, c, torch.zeros((2, 10)), 0.1, 4, 4)


def test_render_f64_matches_oracle(orc):
    w, h, n = 40, 28, 90
    sites = random_sites(w, h, n, 4)
    c0, c1 = orc.init_candidates(n, 3, False, w, h)
    np.testing.assert_allclose(orc.render(c0, c1, sites, iss_of(w, h), w, h), render_f64(c0, c1, sites, w, h),
                               atol=2e-5)


def test_render_backward_matches_fd64(orc):
    """The oracle's autograd backward (sad_bindings.cu:41-207 restated) against
    FP64 central differences of loss = sum(render * R) at 64x64 / 256 sites,
    candidates frozen."""
    w, h = 64, 64
    sites, go = fd_problem(w=w, h=h)
    c0, c1 = converged_field(orc, sites, w, h, passes=10)
    acc, ab = orc.render_backward(c0, c1, go, sites, iss_of(w, h), w, h)
    R = go.astype(np.float64)

    def loss(s):
        return float(np.sum(render_f64(c0, c1, s, w, h) * R))

    rows = list(np.random.default_rng(2).choice(len(sites), 20, replace=False))
    fd = central_differences(loss, sites, rows, list(range(10)), FD_STEPS)
    an = acc[rows]
    bound = 2e-3 * (np.abs(fd) + 1e-2 * ab[rows]) + 1e-6
    bad = np.argwhere(np.abs(an - fd) > bound)
    assert len(bad) == 0, [(rows[i], j, an[i, j], fd[i, j]) for i, j in bad[:5]]


def test_tiled_gradient_equals_autograd_backward(orc):
    """With grad_output = 2 (pred - target) and a full mask, the tiled training
    gradient and the autograd backward are the same derivative."""
    w, h, n = 48, 40, 150
    sites = random_sites(w, h, n, 8, inactive_frac=0.0)
    c0, c1 = converged_field(orc, sites, w, h, passes=8)
    iss = iss_of(w, h)
    rng = np.random.default_rng(3)
    target = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    pred = orc.render(c0, c1, sites, iss, w, h)
    go = (2.0 * (pred - target)).astype(np.float32)
    g_tiled, ab_t, _, _ = orc.gradients(c0, c1, target, np.ones((h, w), np.float32), sites, iss, w, h)
    g_bwd, ab_b = orc.render_backward(c0, c1, go, sites, iss, w, h)
    err = np.abs(g_tiled - g_bwd) / (1e-4 * np.abs(g_bwd) + 1e-5 * ab_b + 1e-30)
    assert err.max() <= 1.0
