"""`torch.ops.sad`: the reference's PyTorch operator library on the HIP C ABI.

Registers the 24 schemas of backends/pytorch/torch_ext/torch_binding.cpp:6-30
under the same names and argument lists, with implementations on the CUDA
dispatch key (HIP tensors use it on ROCm builds of PyTorch).  Each
implementation checks its arguments as the reference wrapper does
(backends/pytorch/torch_ext/cuda/sad_bindings.cu:496-1400, same messages,
raised as RuntimeError), makes inputs contiguous, allocates what the reference
allocates, and calls one `sadh_*` entry point of libsad_hip.so on the current
stream of the tensor's device.  There is no CPU implementation: a CPU tensor
fails in the dispatcher.

Differences from the reference wrappers:
  * work is issued on the current stream rather than the device's default
    stream, so the ops can be captured in HIP graphs and run on side streams;
  * `init_gradient_weighted` hands out the R2 seeds in site order, where the
    reference's threads race for them (same sites, deterministic order);
  * `compute_gradients_tiled` with ssim_weight > 0 checks that `rendered` has
    the target's size (the SSIM term reads it at every pixel).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

_SCHEMAS = [
    "render_sad(Tensor cand0, Tensor cand1, Tensor sites, float inv_scale_sq, int width, int height) -> Tensor",
    "render_sad_padded(Tensor cand0, Tensor cand1, Tensor sites, float inv_scale_sq, int width, int height, "
    "int cand_width, int cand_height) -> Tensor",
    "init_candidates(int width, int height, int site_count, int seed) -> (Tensor, Tensor)",
    "vpt_pass(Tensor cand0, Tensor cand1, Tensor sites, float inv_scale_sq, int width, int height, int jump, "
    "int radius_probes, float radius_scale, int inject_count, int seed) -> (Tensor, Tensor)",
    "render_sad_backward(Tensor cand0, Tensor cand1, Tensor sites, Tensor grad_output, float inv_scale_sq, "
    "int width, int height) -> Tensor",
    "clear_buffer(Tensor buffer, int count) -> Tensor",
    "pack_candidate_sites(Tensor sites, Tensor packed, int site_count) -> Tensor",
    "update_candidates_compact(Tensor cand0_in, Tensor cand1_in, Tensor cand0_out, Tensor cand1_out, "
    "Tensor packed_sites, Tensor hilbert_order, Tensor hilbert_pos, float inv_scale_sq, int site_count, int step, "
    "int step_high, float radius_scale, int radius_probes, int inject_count, int hilbert_probes, "
    "int hilbert_window, int cand_downscale, int target_width, int target_height, int cand_width, "
    "int cand_height) -> Tensor",
    "jfa_clear(Tensor cand0, Tensor cand1, int cand_width, int cand_height) -> Tensor",
    "jfa_seed(Tensor cand0, Tensor sites, int site_count, int cand_downscale, int cand_width, "
    "int cand_height) -> Tensor",
    "jfa_flood(Tensor cand0_in, Tensor cand0_out, Tensor sites, float inv_scale_sq, int site_count, "
    "int step_size, int cand_downscale, int cand_width, int cand_height) -> Tensor",
    "compute_gradients_tiled(Tensor cand0, Tensor cand1, Tensor target, Tensor rendered, Tensor mask, "
    "Tensor sites, Tensor grad_pos_x, Tensor grad_pos_y, Tensor grad_log_tau, Tensor grad_radius, "
    "Tensor grad_color_r, Tensor grad_color_g, Tensor grad_color_b, Tensor grad_dir_x, Tensor grad_dir_y, "
    "Tensor grad_log_aniso, Tensor removal_delta, float inv_scale_sq, int site_count, int compute_removal, "
    "float ssim_weight, int cand_width, int cand_height) -> Tensor",
    "tau_diffuse(Tensor cand0, Tensor cand1, Tensor sites, Tensor grad_raw, Tensor grad_in, Tensor grad_out, "
    "int site_count, float lambda, int cand_downscale, int cand_width, int cand_height) -> Tensor",
    "adam_update(Tensor sites, Tensor adam, Tensor grad_pos_x, Tensor grad_pos_y, Tensor grad_log_tau, "
    "Tensor grad_radius, Tensor grad_color_r, Tensor grad_color_g, Tensor grad_color_b, Tensor grad_dir_x, "
    "Tensor grad_dir_y, Tensor grad_log_aniso, float lr_pos, float lr_tau, float lr_radius, float lr_color, "
    "float lr_dir, float lr_aniso, float beta1, float beta2, float eps, int t, int width, int height) -> Tensor",
    "compute_site_stats_tiled(Tensor cand0, Tensor cand1, Tensor target, Tensor mask, Tensor sites, "
    "Tensor mass, Tensor energy, Tensor err_w, Tensor err_wx, Tensor err_wy, Tensor err_wxx, Tensor err_wxy, "
    "Tensor err_wyy, float inv_scale_sq, int site_count, int cand_width, int cand_height) -> Tensor",
    "compute_densify_score_pairs(Tensor sites, Tensor mass, Tensor energy, Tensor pairs, int site_count, "
    "float min_mass, float score_alpha, int pair_count) -> Tensor",
    "compute_prune_score_pairs(Tensor sites, Tensor removal_delta, Tensor pairs, int site_count, "
    "float delta_norm, int pair_count) -> Tensor",
    "radix_sort_pairs(Tensor pairs, int max_key_exclusive) -> Tensor",
    "write_split_indices(Tensor pairs, Tensor indices, int num_to_split) -> Tensor",
    "split_sites(Tensor sites, Tensor adam, Tensor split_indices, Tensor mass, Tensor err_w, Tensor err_wx, "
    "Tensor err_wy, Tensor err_wxx, Tensor err_wxy, Tensor err_wyy, int current_site_count, int num_to_split, "
    "Tensor target) -> Tensor",
    "prune_sites(Tensor sites, Tensor indices, int count) -> Tensor",
    "build_hilbert_pairs(Tensor sites, Tensor pairs, int site_count, int padded_count, int width, int height, "
    "int bits) -> Tensor",
    "write_hilbert_order(Tensor pairs, Tensor order, Tensor pos, int site_count) -> Tensor",
    "init_gradient_weighted(Tensor sites, Tensor target, Tensor mask, int site_count, float init_log_tau, "
    "float init_radius, float gradient_alpha) -> Tensor",
]

OP_NAMES = [s.split("(")[0] for s in _SCHEMAS]

# --- argument checks (sad_bindings.cu:24-40 and the per-op TORCH_CHECKs) ---------------------------------


def _check(cond: bool, msg: str) -> None:
    if not cond:
        raise RuntimeError(msg)


def _cuda(t: torch.Tensor, name: str) -> None:
    _check(t.is_cuda, f"{name} must be a CUDA tensor")


def _f32(t: torch.Tensor, name: str) -> None:
    _check(t.dtype == torch.float32, f"{name} must be float32")


def _i32(t: torch.Tensor, name: str) -> None:
    _check(t.dtype == torch.int32, f"{name} must be int32")


def _f16(t: torch.Tensor, name: str) -> None:
    _check(t.dtype == torch.float16, f"{name} must be float16")


def _cand(t: torch.Tensor, name: str) -> None:
    _cuda(t, name)
    _i32(t, name)
    _check(t.dim() == 2 and t.size(1) == 4, f"{name} must be [M,4]")


def _sites(t: torch.Tensor, name: str = "sites") -> None:
    _cuda(t, name)
    _f32(t, name)
    _check(t.dim() == 2 and t.size(1) == 10, f"{name} must be [N,10]")


def _image(t: torch.Tensor, name: str) -> None:
    _cuda(t, name)
    _f32(t, name)
    _check(t.dim() == 3 and t.size(2) == 3, f"{name} must be [H,W,3]")


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _p(t):
    return None if t is None else t.data_ptr()


def _run(name: str, dev_tensor: torch.Tensor, *args) -> None:
    with torch.cuda.device(dev_tensor.device):
        _lib.call(name, *args, _stream(dev_tensor))


# --- implementations --------------------------------------------------------------------------------------


def render_sad(cand0, cand1, sites, inv_scale_sq, width, height):
    """renderVoronoi, sad_bindings.cu:496-539."""
    return render_sad_padded(cand0, cand1, sites, inv_scale_sq, width, height, width, height, _padded=False)


def render_sad_padded(cand0, cand1, sites, inv_scale_sq, width, height, cand_width, cand_height, _padded=True):
    """renderVoronoiPadded, sad_bindings.cu:541-589."""
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _sites(sites)
    _check(width > 0 and height > 0, "width/height must be positive")
    if _padded:
        _check(cand_width > 0 and cand_height > 0, "invalid cand size")
        _check(cand0.size(0) == cand_width * cand_height, "cand0 size must match cand_width*cand_height")
        _check(cand1.size(0) == cand_width * cand_height, "cand1 size must match cand_width*cand_height")
    else:
        _check(cand0.size(0) == width * height, "cand0 size must match width*height")
        _check(cand1.size(0) == width * height, "cand1 size must match width*height")
    cand0, cand1, sites = cand0.contiguous(), cand1.contiguous(), sites.contiguous()
    out = torch.empty((height, width, 3), dtype=torch.float32, device=cand0.device)
    _run("sadh_render_voronoi", cand0, _p(cand0), _p(cand1), _p(out), _p(sites), float(inv_scale_sq),
         sites.size(0), width, height, cand_width, cand_height)
    return out


def init_candidates(width, height, site_count, seed):
    """initCandidates, sad_bindings.cu:591-621: per-pixel mode when site_count == width*height."""
    _check(width > 0 and height > 0, "invalid cand size")
    _check(site_count >= 0, "site_count must be non-negative")
    dev = torch.device("cuda", torch.cuda.current_device())
    count = width * height
    cand0 = torch.empty((count, 4), dtype=torch.int32, device=dev)
    cand1 = torch.empty((count, 4), dtype=torch.int32, device=dev)
    _run("sadh_init_candidates", cand0, _p(cand0), _p(cand1), site_count, seed & 0xFFFFFFFF,
         int(site_count == count), width, height)
    return cand0, cand1


def vpt_pass(cand0, cand1, sites, inv_scale_sq, width, height, jump, radius_probes, radius_scale, inject_count,
             seed):
    """vptPass, sad_bindings.cu:623-696: fp16 pack + one VPT pass with step = jump << 16.

    `seed` is accepted and ignored, as in the reference."""
    del seed
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _sites(sites)
    _check(cand0.size(0) == width * height, "cand0 size must match width*height")
    _check(cand1.size(0) == width * height, "cand1 size must match width*height")
    cand0, cand1, sites = cand0.contiguous(), cand1.contiguous(), sites.contiguous()
    out0 = torch.empty_like(cand0)
    out1 = torch.empty_like(cand1)
    n = sites.size(0)
    packed = torch.empty((n, 8), dtype=torch.float16, device=sites.device)
    _run("sadh_pack_candidate_sites", sites, _p(sites), _p(packed), n)
    _run("sadh_update_candidates", cand0, _p(cand0), _p(cand1), _p(out0), _p(out1), _p(packed), n,
         (jump << 16) & 0xFFFFFFFF, float(inv_scale_sq), 0, float(radius_scale), radius_probes, inject_count,
         None, None, 0, 0, 1, width, height, width, height)
    return out0, out1


def render_sad_backward(cand0, cand1, sites, grad_output, inv_scale_sq, width, height):
    """renderVoronoiBackward, sad_bindings.cu:698-749: grad_sites [N,10] (Site field order)."""
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _sites(sites)
    _image(grad_output, "grad_output")
    _check(grad_output.size(0) == height and grad_output.size(1) == width,
           "grad_output size must match height/width")
    _check(cand0.size(0) == width * height, "cand0 size must match width*height")
    _check(cand1.size(0) == width * height, "cand1 size must match width*height")
    cand0, cand1, sites, grad_output = cand0.contiguous(), cand1.contiguous(), sites.contiguous(), \
        grad_output.contiguous()
    grad_sites = torch.zeros((sites.size(0), 10), dtype=torch.float32, device=sites.device)
    _run("sadh_render_backward", cand0, _p(cand0), _p(cand1), _p(grad_output), _p(grad_sites), _p(sites),
         float(inv_scale_sq), sites.size(0), width, height, width, height)
    return grad_sites


def clear_buffer(buffer, count):
    """clearBuffer, sad_bindings.cu:751-764: count <= 0 or > numel clears everything."""
    _cuda(buffer, "buffer")
    _f32(buffer, "buffer")
    _check(buffer.is_contiguous(), "buffer must be contiguous")
    numel = buffer.numel()
    if count <= 0 or count > numel:
        count = numel
    _run("sadh_clear_buffer", buffer, _p(buffer), count)
    return buffer


def pack_candidate_sites(sites, packed, site_count):
    """packCandidateSites, sad_bindings.cu:766-784."""
    _sites(sites)
    _cuda(packed, "packed")
    _f16(packed, "packed")
    _check(packed.dim() == 2 and packed.size(1) == 8, "packed must be [N,8]")
    _check(packed.is_contiguous(), "packed must be contiguous")
    sites = sites.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    _check(site_count <= min(sites.size(0), packed.size(0)), "site_count exceeds sites/packed rows")
    _run("sadh_pack_candidate_sites", sites, _p(sites), _p(packed), site_count)
    return packed


def update_candidates_compact(cand0_in, cand1_in, cand0_out, cand1_out, packed_sites, hilbert_order, hilbert_pos,
                              inv_scale_sq, site_count, step, step_high, radius_scale, radius_probes, inject_count,
                              hilbert_probes, hilbert_window, cand_downscale, target_width, target_height,
                              cand_width, cand_height):
    """updateCandidatesCompact, sad_bindings.cu:786-863 (one VPT pass on prepacked sites)."""
    for t, n in ((cand0_in, "cand0_in"), (cand1_in, "cand1_in"), (cand0_out, "cand0_out"), (cand1_out, "cand1_out")):
        _cand(t, n)
        _check(t.size(0) == cand_width * cand_height, f"{n} size mismatch")
    _check(cand0_out.is_contiguous() and cand1_out.is_contiguous(), "cand outputs must be contiguous")
    _cuda(packed_sites, "packed_sites")
    _f16(packed_sites, "packed_sites")
    _check(packed_sites.dim() == 2 and packed_sites.size(1) == 8, "packed_sites must be [N,8]")
    cand0_in, cand1_in, packed_sites = cand0_in.contiguous(), cand1_in.contiguous(), packed_sites.contiguous()
    if site_count <= 0:
        site_count = packed_sites.size(0)
    ho = hp = None
    if hilbert_probes > 0 and hilbert_window > 0:
        _i32(hilbert_order, "hilbert_order")
        _i32(hilbert_pos, "hilbert_pos")
        hilbert_order, hilbert_pos = hilbert_order.contiguous(), hilbert_pos.contiguous()
        ho, hp = hilbert_order, hilbert_pos
    _run("sadh_update_candidates", cand0_in, _p(cand0_in), _p(cand1_in), _p(cand0_out), _p(cand1_out),
         _p(packed_sites), site_count, step & 0xFFFFFFFF, float(inv_scale_sq), step_high & 0xFFFFFFFF,
         float(radius_scale), radius_probes, inject_count, _p(ho), _p(hp), hilbert_probes, hilbert_window,
         cand_downscale, target_width, target_height, cand_width, cand_height)
    return cand0_out


def jfa_clear(cand0, cand1, cand_width, cand_height):
    """jfaClear, sad_bindings.cu:865-881."""
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _check(cand0.is_contiguous() and cand1.is_contiguous(), "cand0/cand1 must be contiguous")
    _check(cand0.size(0) >= cand_width * cand_height and cand1.size(0) >= cand_width * cand_height,
           "cand0/cand1 smaller than cand_width*cand_height")
    _run("sadh_clear_candidates", cand0, _p(cand0), _p(cand1), cand_width, cand_height)
    return cand0


def jfa_seed(cand0, sites, site_count, cand_downscale, cand_width, cand_height):
    """jfaSeed, sad_bindings.cu:883-906."""
    _sites(sites)
    _cand(cand0, "cand0")
    _check(cand0.size(0) == cand_width * cand_height, "cand0 size must match cand_width*cand_height")
    _check(cand0.is_contiguous(), "cand0 must be contiguous")
    sites = sites.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    _run("sadh_jfa_seed", cand0, _p(cand0), _p(sites), site_count, cand_width, cand_height, cand_downscale)
    return cand0


def jfa_flood(cand0_in, cand0_out, sites, inv_scale_sq, site_count, step_size, cand_downscale, cand_width,
              cand_height):
    """jfaFlood, sad_bindings.cu:908-942."""
    _sites(sites)
    _cand(cand0_in, "cand0_in")
    _cand(cand0_out, "cand0_out")
    _check(cand0_in.size(0) == cand_width * cand_height, "cand0_in size must match cand_width*cand_height")
    _check(cand0_out.size(0) == cand_width * cand_height, "cand0_out size must match cand_width*cand_height")
    _check(cand0_out.is_contiguous(), "cand0_out must be contiguous")
    cand0_in, sites = cand0_in.contiguous(), sites.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    _run("sadh_jfa_flood", cand0_in, _p(cand0_in), _p(cand0_out), _p(sites), site_count, step_size,
         float(inv_scale_sq), cand_width, cand_height, cand_downscale, cand_width * cand_downscale,
         cand_height * cand_downscale)
    return cand0_out


_GRAD_NAMES = ["grad_pos_x", "grad_pos_y", "grad_log_tau", "grad_radius", "grad_color_r", "grad_color_g",
               "grad_color_b", "grad_dir_x", "grad_dir_y", "grad_log_aniso"]


def _grad_arrays(grads, count):
    for g, n in zip(grads, _GRAD_NAMES):
        _cuda(g, n)
        _f32(g, n)
        _check(g.is_contiguous() and g.numel() >= count, f"{n} must be contiguous with >= site_count elements")


def compute_gradients_tiled(cand0, cand1, target, rendered, mask, sites, grad_pos_x, grad_pos_y, grad_log_tau,
                            grad_radius, grad_color_r, grad_color_g, grad_color_b, grad_dir_x, grad_dir_y,
                            grad_log_aniso, removal_delta, inv_scale_sq, site_count, compute_removal, ssim_weight,
                            cand_width, cand_height):
    """computeGradientsTiled, sad_bindings.cu:944-1013 (adds into the 10 SoA gradient arrays)."""
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _image(target, "target")
    _image(rendered, "rendered")
    _cuda(mask, "mask")
    _f32(mask, "mask")
    _check(mask.dim() == 2, "mask must be [H,W]")
    _check(mask.size(0) == target.size(0) and mask.size(1) == target.size(1), "mask size must match target")
    _sites(sites)
    cand0, cand1, target, rendered, mask, sites = (t.contiguous() for t in (cand0, cand1, target, rendered, mask,
                                                                            sites))
    if site_count <= 0:
        site_count = sites.size(0)
    grads = [grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r, grad_color_g, grad_color_b,
             grad_dir_x, grad_dir_y, grad_log_aniso]
    _grad_arrays(grads, site_count)
    _cuda(removal_delta, "removal_delta")
    _check(removal_delta.is_contiguous(), "removal_delta must be contiguous")
    _check(cand0.size(0) == cand_width * cand_height, "cand0 size must match cand_width*cand_height")
    _check(not ssim_weight > 0.0 or rendered.shape == target.shape, "rendered size must match target")
    _run("sadh_compute_gradients_tiled", cand0, _p(cand0), _p(cand1), _p(target), _p(rendered), _p(mask),
         *[_p(g) for g in grads], _p(sites), float(inv_scale_sq), site_count, _p(removal_delta),
         int(compute_removal), float(ssim_weight), target.size(1), target.size(0), cand_width, cand_height)
    return grad_pos_x


def tau_diffuse(cand0, cand1, sites, grad_raw, grad_in, grad_out, site_count, lambda_, cand_downscale, cand_width,
                cand_height):
    """tauDiffuse, sad_bindings.cu:1015-1057."""
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _sites(sites)
    for t, n in ((grad_raw, "grad_raw"), (grad_in, "grad_in"), (grad_out, "grad_out")):
        _cuda(t, n)
        _f32(t, n)
    _check(grad_out.is_contiguous(), "grad_out must be contiguous")
    cand0, cand1, sites, grad_raw, grad_in = (t.contiguous() for t in (cand0, cand1, sites, grad_raw, grad_in))
    if site_count <= 0:
        site_count = sites.size(0)
    _run("sadh_tau_diffuse", cand0, _p(cand0), _p(cand1), _p(sites), _p(grad_raw), _p(grad_in), _p(grad_out),
         site_count, float(lambda_), cand_width, cand_height, cand_downscale)
    return grad_out


def adam_update(sites, adam, grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r, grad_color_g,
                grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso, lr_pos, lr_tau, lr_radius, lr_color, lr_dir,
                lr_aniso, beta1, beta2, eps, t, width, height):
    """adamUpdate, sad_bindings.cu:1059-1104 (all sites rows, t = iteration + 1)."""
    _sites(sites)
    _cuda(adam, "adam")
    _f32(adam, "adam")
    _check(adam.dim() == 2 and adam.size(1) == 20, "adam must be [N,20]")
    _check(sites.is_contiguous() and adam.is_contiguous(), "sites/adam must be contiguous")
    n = sites.size(0)
    _check(adam.size(0) >= n, "adam must have a row per site")
    grads = [grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r, grad_color_g, grad_color_b,
             grad_dir_x, grad_dir_y, grad_log_aniso]
    _grad_arrays(grads, n)
    _run("sadh_adam_update", sites, _p(sites), _p(adam), *[_p(g) for g in grads], float(lr_pos), float(lr_tau),
         float(lr_radius), float(lr_color), float(lr_dir), float(lr_aniso), float(beta1), float(beta2), float(eps),
         t, n, width, height)
    return sites


def compute_site_stats_tiled(cand0, cand1, target, mask, sites, mass, energy, err_w, err_wx, err_wy, err_wxx,
                             err_wxy, err_wyy, inv_scale_sq, site_count, cand_width, cand_height):
    """computeSiteStatsTiled, sad_bindings.cu:1106-1161 (clears and fills the 8 stat arrays)."""
    _cand(cand0, "cand0")
    _cand(cand1, "cand1")
    _image(target, "target")
    _cuda(mask, "mask")
    _check(mask.dim() == 2, "mask must be [H,W]")
    _check(mask.size(0) == target.size(0) and mask.size(1) == target.size(1), "mask size must match target")
    _sites(sites)
    cand0, cand1, target, mask, sites = (t.contiguous() for t in (cand0, cand1, target, mask, sites))
    if site_count <= 0:
        site_count = sites.size(0)
    stats = [mass, energy, err_w, err_wx, err_wy, err_wxx, err_wxy, err_wyy]
    for s in stats:
        _cuda(s, "stats")
        _f32(s, "stats")
        _check(s.is_contiguous() and s.numel() >= site_count, "stat arrays must hold site_count floats")
    _run("sadh_compute_site_stats_simple", cand0, _p(cand0), _p(cand1), _p(target), _p(mask), _p(sites),
         float(inv_scale_sq), site_count, *[_p(s) for s in stats], target.size(1), target.size(0), cand_width,
         cand_height)
    return mass


def _pairs(t, name="pairs"):
    _cuda(t, name)
    _i32(t, name)
    _check(t.dim() == 2 and t.size(1) == 2, f"{name} must be [M,2]")
    _check(t.is_contiguous(), f"{name} must be contiguous")


def compute_densify_score_pairs(sites, mass, energy, pairs, site_count, min_mass, score_alpha, pair_count):
    """computeDensifyScorePairs, sad_bindings.cu:1163-1188."""
    _sites(sites)
    _pairs(pairs)
    sites, mass, energy = sites.contiguous(), mass.contiguous(), energy.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    if pair_count <= 0:
        pair_count = pairs.size(0)
    _check(pair_count <= pairs.size(0), "pair_count exceeds pairs rows")
    _run("sadh_compute_densify_score_pairs", sites, _p(sites), _p(mass), _p(energy), _p(pairs), site_count,
         float(min_mass), float(score_alpha), pair_count)
    return pairs


def compute_prune_score_pairs(sites, removal_delta, pairs, site_count, delta_norm, pair_count):
    """computePruneScorePairs, sad_bindings.cu:1190-1212."""
    _sites(sites)
    _pairs(pairs)
    sites, removal_delta = sites.contiguous(), removal_delta.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    if pair_count <= 0:
        pair_count = pairs.size(0)
    _check(pair_count <= pairs.size(0), "pair_count exceeds pairs rows")
    _run("sadh_compute_prune_score_pairs", sites, _p(sites), _p(removal_delta), _p(pairs), site_count,
         float(delta_norm), pair_count)
    return pairs


def radix_sort_pairs(pairs, max_key_exclusive):
    """radixSortPairs, sad_bindings.cu:1214-1245: stable ascending sort by column 0, in place."""
    _pairs(pairs)
    count = pairs.size(0)
    scratch = torch.empty_like(pairs)
    tiles = (count + 1023) // 1024
    hist = torch.empty((max(1, tiles) * 256,), dtype=torch.int32, device=pairs.device)
    block_sums = torch.empty((max(1, (tiles * 256 + 255) // 256),), dtype=torch.int32, device=pairs.device)
    _run("sadh_radix_sort_pairs", pairs, _p(pairs), _p(scratch), _p(hist), _p(block_sums), count,
         max_key_exclusive & 0xFFFFFFFF)
    return pairs


def write_split_indices(pairs, indices, num_to_split):
    """writeSplitIndices, sad_bindings.cu:1247-1257."""
    _pairs(pairs)
    _check(indices.is_contiguous() and indices.numel() >= num_to_split, "indices must hold num_to_split entries")
    _check(num_to_split <= pairs.size(0), "num_to_split exceeds pairs rows")
    _run("sadh_write_split_indices", pairs, _p(pairs), _p(indices), num_to_split)
    return indices


def split_sites(sites, adam, split_indices, mass, err_w, err_wx, err_wy, err_wxx, err_wxy, err_wyy,
                current_site_count, num_to_split, target):
    """splitSites, sad_bindings.cu:1259-1286: children at current_site_count + i."""
    _sites(sites)
    _image(target, "target")
    _check(sites.is_contiguous() and adam.is_contiguous(), "sites/adam must be contiguous")
    _check(current_site_count + num_to_split <= sites.size(0), "split children exceed the sites buffer")
    target = target.contiguous()
    _run("sadh_split_sites", sites, _p(sites), _p(adam), _p(split_indices.contiguous()), num_to_split, _p(mass),
         _p(err_w), _p(err_wx), _p(err_wy), _p(err_wxx), _p(err_wxy), _p(err_wyy), current_site_count, _p(target),
         target.size(1), target.size(0))
    return sites


def prune_sites(sites, indices, count):
    """pruneSites, sad_bindings.cu:1288-1298."""
    _sites(sites)
    _check(sites.is_contiguous(), "sites must be contiguous")
    _run("sadh_prune_sites_by_index", sites, _p(sites), _p(indices.contiguous()), count)
    return sites


def build_hilbert_pairs(sites, pairs, site_count, padded_count, width, height, bits):
    """buildHilbertPairs, sad_bindings.cu:1300-1330."""
    _sites(sites)
    _pairs(pairs)
    sites = sites.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    if padded_count <= 0:
        padded_count = pairs.size(0)
    _check(padded_count <= pairs.size(0), "padded_count exceeds pairs rows")
    _run("sadh_hilbert_pairs", sites, _p(sites), _p(pairs), site_count, padded_count, width, height, bits)
    return pairs


def write_hilbert_order(pairs, order, pos, site_count):
    """writeHilbertOrder, sad_bindings.cu:1332-1350."""
    _pairs(pairs)
    _check(order.is_contiguous() and pos.is_contiguous(), "order/pos must be contiguous")
    if site_count <= 0:
        site_count = order.size(0)
    _run("sadh_hilbert_order", pairs, _p(pairs), _p(order), _p(pos), site_count)
    return order


def init_gradient_weighted(sites, target, mask, site_count, init_log_tau, init_radius, gradient_alpha):
    """initGradientWeighted, sad_bindings.cu:1399-1437: threshold 0.01 * alpha,
    256 attempts, seeds from 0.  Sites receive their seeds in site order (the
    reference's threads race for them), so the result is deterministic."""
    _sites(sites)
    _image(target, "target")
    _cuda(mask, "mask")
    _f32(mask, "mask")
    _check(mask.dim() == 2, "mask must be [H,W]")
    _check(mask.size(0) == target.size(0) and mask.size(1) == target.size(1), "mask size must match target")
    sites, target, mask = sites.contiguous(), target.contiguous(), mask.contiguous()
    if site_count <= 0:
        site_count = sites.size(0)
    _check(site_count <= sites.size(0), "site_count exceeds sites rows")
    counter = torch.zeros(1, dtype=torch.int32, device=sites.device)
    _run("sadh_init_gradient_weighted", sites, _p(sites), site_count, _p(counter), _p(target), _p(mask),
         float(np.float32(0.01) * np.float32(gradient_alpha)), 256, float(init_log_tau), float(init_radius),
         target.size(1), target.size(0))
    return sites


_IMPLS = {
    "render_sad": render_sad, "render_sad_padded": render_sad_padded, "init_candidates": init_candidates,
    "vpt_pass": vpt_pass, "render_sad_backward": render_sad_backward, "clear_buffer": clear_buffer,
    "pack_candidate_sites": pack_candidate_sites, "update_candidates_compact": update_candidates_compact,
    "jfa_clear": jfa_clear, "jfa_seed": jfa_seed, "jfa_flood": jfa_flood,
    "compute_gradients_tiled": compute_gradients_tiled, "tau_diffuse": tau_diffuse, "adam_update": adam_update,
    "compute_site_stats_tiled": compute_site_stats_tiled,
    "compute_densify_score_pairs": compute_densify_score_pairs,
    "compute_prune_score_pairs": compute_prune_score_pairs, "radix_sort_pairs": radix_sort_pairs,
    "write_split_indices": write_split_indices, "split_sites": split_sites, "prune_sites": prune_sites,
    "build_hilbert_pairs": build_hilbert_pairs, "write_hilbert_order": write_hilbert_order,
    "init_gradient_weighted": init_gradient_weighted,
}

_LIB = None


def register() -> None:
    """Define the `sad` namespace and register the CUDA-key implementations (idempotent)."""
    global _LIB
    if _LIB is not None:
        return
    lib = torch.library.Library("sad", "DEF")
    for schema in _SCHEMAS:
        lib.define(schema)
    for name in OP_NAMES:
        # init_candidates has no tensor argument: like the reference, register it for every backend
        key = "CompositeExplicitAutograd" if name == "init_candidates" else "CUDA"
        lib.impl(name, _IMPLS[name], key)
    _LIB = lib


register()
