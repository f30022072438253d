"""numpy front end of the CPU oracle (liboracle.so) - TEST INFRASTRUCTURE.

Only the tests, the smoke check and the CPU-baseline leg of bench.py use this
module; it is the checker, never the measured or shipped path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

P = C.c_void_p
U32, I32, F32, I64 = C.c_uint32, C.c_int, C.c_float, C.c_int64


class OrcFitConfig(C.Structure):
    _fields_ = [
        ("iterations", I32), ("target_bpp", F32), ("init_mode", I32), ("n_sites", I32), ("max_sites", I32),
        ("prune_percentile", F32), ("prune_start", I32), ("prune_freq", I32), ("prune_end", I32),
        ("prune_during_densify", I32), ("densify_enabled", I32), ("densify_start", I32), ("densify_freq", I32),
        ("densify_end", I32), ("densify_percentile", F32), ("densify_score_alpha", F32),
        ("cand_update_freq", I32), ("cand_update_passes", I32), ("cand_downscale", I32),
        ("cand_radius_scale", F32), ("cand_radius_probes", U32), ("cand_inject_count", U32),
        ("cand_hilbert_window", U32), ("cand_hilbert_probes", U32),
        ("lr_scale", F32), ("lr_pos", F32), ("lr_tau", F32), ("lr_radius", F32), ("lr_color", F32),
        ("lr_dir", F32), ("lr_aniso", F32), ("beta1", F32), ("beta2", F32), ("eps", F32), ("log_freq", I32),
        ("init_gradient_alpha", F32), ("init_log_tau", F32), ("init_radius", F32), ("ssim_weight", F32),
        ("ssim_metric", I32),
    ]


class OrcFitResult(C.Structure):
    _fields_ = [
        ("final_psnr", F32), ("best_psnr", F32), ("actual_sites", I32), ("active_estimate", I32),
        ("densify_percentile", F32), ("prune_percentile", F32), ("capacity", I32), ("n_logs", I32),
        ("n_densify", I32), ("n_prune", I32), ("n_cand_updates", I32), ("final_ssim", F32), ("best_ssim", F32),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_sse.restype = C.c_double
        _lib.orc_ssim_sum.restype = C.c_double
        _lib.orc_hilbert_index.restype = U32
        _lib.orc_jump_step.restype = U32
        _lib.orc_get_threads.restype = I32
        _lib.orc_fit.restype = I32
        _lib.orc_init_gradient_weighted.restype = U32
        _lib.orc_init_grad_eval.restype = I32
        _lib.orc_fix_value.restype = F32
        _lib.orc_sse_fix.restype = F32
        _lib.orc_simulate_final.restype = I32
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def set_threads(n: int) -> None:
    lib().orc_set_threads(I32(n))


def get_threads() -> int:
    return lib().orc_get_threads()


# ---------------------------------------------------------------- primitives
def expf(x):
    x = _f32(x); y = np.empty_like(x); lib().orc_expf_array(_p(x), _p(y), I64(x.size)); return y


def logf(x):
    x = _f32(x); y = np.empty_like(x); lib().orc_logf_array(_p(x), _p(y), I64(x.size)); return y


def f32_to_f16_bits(x):
    x = _f32(x); y = np.empty(x.shape, np.uint16); lib().orc_f32_to_f16_array(_p(x), _p(y), I64(x.size)); return y


def f16_bits_to_f32(h):
    h = np.ascontiguousarray(h, np.uint16); y = np.empty(h.shape, np.float32)
    lib().orc_f16_to_f32_array(_p(h), _p(y), I64(h.size)); return y


def hilbert_index(x, y, bits):
    return lib().orc_hilbert_index(U32(x), U32(y), I32(bits))


def jump_step(i, w, h):
    return lib().orc_jump_step(U32(i), I32(w), I32(h))


# ------------------------------------------------------------------- kernels
def render(cand0, cand1, sites, iss, width, height, cand_w=None, cand_h=None):
    cand_w = cand_w or width; cand_h = cand_h or height
    c0, c1, s = _u32(cand0), _u32(cand1), _f32(sites)
    out = np.zeros((height, width, 3), np.float32)
    lib().orc_render(_p(c0), _p(c1), _p(out), _p(s), F32(iss), U32(s.shape[0]), I32(width), I32(height),
                     I32(cand_w), I32(cand_h))
    return out


def render_cells(cand0, cand1, sites, iss, width, height, cand_w=None, cand_h=None):
    """render with each site's colour replaced by sad_hash_color(id)"""
    cand_w = cand_w or width; cand_h = cand_h or height
    c0, c1, s = _u32(cand0), _u32(cand1), _f32(sites)
    out = np.zeros((height, width, 3), np.float32)
    lib().orc_render_cells(_p(c0), _p(c1), _p(out), _p(s), F32(iss), U32(s.shape[0]), I32(width), I32(height),
                           I32(cand_w), I32(cand_h))
    return out


def tau_heatmap(cand0, cand1, sites, tmin, tmean, tmax, dot, width, height, cand_w=None, cand_h=None):
    c0, c1, s = _u32(cand0), _u32(cand1), _f32(sites)
    out = np.zeros((height, width, 3), np.float32)
    lib().orc_tau_heatmap(_p(c0), _p(c1), _p(out), _p(s), U32(s.shape[0]), F32(tmin), F32(tmean), F32(tmax),
                          F32(dot), I32(width), I32(height), I32(cand_w or width), I32(cand_h or height))
    return out


def hash_colors(ids):
    ids = np.ascontiguousarray(ids, np.uint32)
    out = np.zeros((ids.size, 3), np.float32)
    lib().orc_hash_colors(_p(ids), U32(ids.size), _p(out))
    return out


def init_candidates(n, seed, per_pixel, width, height):
    c0 = np.zeros((width * height, 4), np.uint32); c1 = np.zeros_like(c0)
    lib().orc_init_candidates(_p(c0), _p(c1), U32(n), U32(seed), I32(int(per_pixel)), I32(width), I32(height))
    return c0, c1


def clear_candidates(width, height):
    c0 = np.zeros((width * height, 4), np.uint32); c1 = np.zeros_like(c0)
    lib().orc_clear_candidates(_p(c0), _p(c1), I32(width), I32(height))
    return c0, c1


def jfa_seed(cand0, sites, width, height, downscale=1):
    c0 = _u32(cand0).copy(); s = _f32(sites)
    lib().orc_jfa_seed(_p(c0), _p(s), U32(s.shape[0]), I32(width), I32(height), I32(downscale))
    return c0


def jfa_flood(cand0, sites, step, iss, width, height, downscale=1, target_w=None, target_h=None):
    c0 = _u32(cand0); s = _f32(sites); out = np.zeros_like(c0)
    lib().orc_jfa_flood(_p(c0), _p(out), _p(s), U32(s.shape[0]), U32(step), F32(iss), I32(width), I32(height),
                        I32(downscale), I32(target_w or width), I32(target_h or height))
    return out


def pack_sites(sites):
    s = _f32(sites); out = np.zeros((s.shape[0], 8), np.uint16)
    lib().orc_pack_sites(_p(s), _p(out), U32(s.shape[0]))
    return out


def update_candidates(cand0, cand1, packed, step, iss, width, height, step_high=0, radius_scale=64.0,
                      radius_probes=0, inject_count=4, hilbert_order=None, hilbert_pos=None, hilbert_probes=0,
                      hilbert_window=0, downscale=1, target_w=None, target_h=None, n=None):
    c0, c1 = _u32(cand0), _u32(cand1)
    pk = np.ascontiguousarray(packed, np.uint16)
    o0, o1 = np.zeros_like(c0), np.zeros_like(c1)
    ho = None if hilbert_order is None else _u32(hilbert_order)
    hp = None if hilbert_pos is None else _u32(hilbert_pos)
    lib().orc_update_candidates(_p(c0), _p(c1), _p(o0), _p(o1), _p(pk), U32(n if n is not None else pk.shape[0]),
                                U32(step), F32(iss), U32(step_high), F32(radius_scale), U32(radius_probes),
                                U32(inject_count), _p(ho), _p(hp), U32(hilbert_probes), U32(hilbert_window),
                                I32(downscale), I32(target_w or width), I32(target_h or height), I32(width),
                                I32(height))
    return o0, o1


def gradients(cand0, cand1, target, mask, sites, iss, width, height, compute_removal=False, cand_w=None,
              cand_h=None, rendered=None, ssim_weight=0.0):
    """Returns (grad [N,10] float64, abs-sum [N,10], removal [N], removal abs-sum [N]).
    With ssim_weight > 0 the loss gains -ssim_weight * SSIM, read from `rendered`."""
    c0, c1, t, m, s = _u32(cand0), _u32(cand1), _f32(target), _f32(mask), _f32(sites)
    n = s.shape[0]
    acc = np.zeros((n, 10)); ab = np.zeros((n, 10)); rd = np.zeros(n); ra = np.zeros(n)
    if ssim_weight > 0.0:
        r = _f32(rendered)
        lib().orc_gradients_ssim(_p(c0), _p(c1), _p(t), _p(r), _p(m), _p(s), F32(iss), U32(n),
                                 I32(int(compute_removal)), F32(ssim_weight), I32(width), I32(height),
                                 I32(cand_w or width), I32(cand_h or height), _p(acc), _p(ab), _p(rd), _p(ra))
    else:
        lib().orc_gradients(_p(c0), _p(c1), _p(t), _p(m), _p(s), F32(iss), U32(n), I32(int(compute_removal)),
                            I32(width), I32(height), I32(cand_w or width), I32(cand_h or height), _p(acc), _p(ab),
                            _p(rd), _p(ra))
    return acc, ab, rd, ra


def render_backward(cand0, cand1, grad_out, sites, iss, width, height):
    c0, c1, go, s = _u32(cand0), _u32(cand1), _f32(grad_out), _f32(sites)
    n = s.shape[0]
    acc = np.zeros((n, 10)); ab = np.zeros((n, 10))
    lib().orc_render_backward(_p(c0), _p(c1), _p(go), _p(s), F32(iss), U32(n), I32(width), I32(height),
                              I32(width), I32(height), _p(acc), _p(ab))
    return acc, ab


# ------------------------------------------------- the library's exact sums
def fix_split(v):
    """float32 -> (lo, hi) int64 words of the library's fixed point"""
    lo, hi = C.c_int64(), C.c_int64()
    lib().orc_fix_split(F32(float(v)), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def fix_values(pairs):
    """int64 [..., 2] (lo, hi) pairs -> float32 [...]"""
    p = np.ascontiguousarray(pairs, np.int64)
    out = np.zeros(p.shape[:-1], np.float32)
    lib().orc_fix_values(_p(p), _p(out), I64(out.size))
    return out


# resolution of the library's fixed point: a sum of P float partials is within P * FIX_HALF_STEP of exact
FIX_HALF_STEP = 2.0 ** -41


def gradients_fix(cand0, cand1, target, mask, sites, iss, width, height, compute_removal=False, cand_w=None,
                  cand_h=None, rendered=None, ssim_weight=0.0):
    """The tiled gradient summed as the HIP library sums it: (grad [N,10] float32, removal [N] float32,
    partials per site [N] int32)."""
    c0, c1, t, m, s = _u32(cand0), _u32(cand1), _f32(target), _f32(mask), _f32(sites)
    n = s.shape[0]
    g = np.zeros((n, 10, 2), np.int64)
    r = np.zeros((n, 2), np.int64)
    parts = np.zeros(n, np.int32)
    ren = _f32(rendered) if ssim_weight > 0.0 else None
    lib().orc_gradients_fix(_p(c0), _p(c1), _p(t), _p(m), _p(s), F32(iss), U32(n), I32(int(compute_removal)),
                            I32(width), I32(height), I32(cand_w or width), I32(cand_h or height), _p(ren),
                            F32(ssim_weight), _p(g), _p(r), _p(parts))
    return fix_values(g), fix_values(r), parts


def render_backward_fix(cand0, cand1, grad_out, sites, iss, width, height):
    """-> (grad [N,10] float32, partials per site [N] int32)"""
    c0, c1, go, s = _u32(cand0), _u32(cand1), _f32(grad_out), _f32(sites)
    n = s.shape[0]
    g = np.zeros((n, 10, 2), np.int64)
    parts = np.zeros(n, np.int32)
    lib().orc_render_backward_fix(_p(c0), _p(c1), _p(go), _p(s), F32(iss), U32(n), I32(width), I32(height),
                                  I32(width), I32(height), _p(g), _p(parts))
    return fix_values(g), parts


def site_stats_fix(cand0, cand1, target, mask, sites, iss, width, height):
    """-> (stats [N,8] float32, partials per site [N] int32)"""
    c0, c1, t, m, s = _u32(cand0), _u32(cand1), _f32(target), _f32(mask), _f32(sites)
    n = s.shape[0]
    g = np.zeros((n, 8, 2), np.int64)
    parts = np.zeros(n, np.int32)
    lib().orc_site_stats_fix(_p(c0), _p(c1), _p(t), _p(m), _p(s), F32(iss), U32(n), I32(width), I32(height),
                             I32(width), I32(height), _p(g), _p(parts))
    return fix_values(g), parts


def sse_fix(rendered, target, mask):
    r, t, m = _f32(rendered), _f32(target), _f32(mask)
    h, w = m.shape
    return lib().orc_sse_fix(_p(r), _p(t), _p(m), I32(w), I32(h))


def tau_diffuse(cand0, cand1, sites, graw, gin, lam, width, height, downscale=1):
    c0, c1, s, gr, gi = _u32(cand0), _u32(cand1), _f32(sites), _f32(graw), _f32(gin)
    out = np.zeros_like(gi)
    lib().orc_tau_diffuse(_p(c0), _p(c1), _p(s), _p(gr), _p(gi), _p(out), U32(s.shape[0]), F32(lam), I32(width),
                          I32(height), I32(downscale))
    return out


def adam(sites, state, grads, lrs, beta1, beta2, eps, t, width, height):
    """grads: float32 [10][N]; lrs: (pos, tau, radius, color, dir, aniso). Returns new (sites, state)."""
    s = _f32(sites).copy(); st = _f32(state).copy(); g = [_f32(x).copy() for x in grads]
    arr = (C.c_void_p * 10)(*[x.ctypes.data for x in g])
    lib().orc_adam(_p(s), _p(st), arr, *[F32(v) for v in lrs], F32(beta1), F32(beta2), F32(eps), U32(t),
                   U32(s.shape[0]), I32(width), I32(height))
    return s, st


def densify_pairs(sites, mass, energy, min_mass, alpha, count):
    s, m, e = _f32(sites), _f32(mass), _f32(energy)
    out = np.zeros((count, 2), np.uint32)
    lib().orc_densify_pairs(_p(s), _p(m), _p(e), _p(out), U32(s.shape[0]), F32(min_mass), F32(alpha), U32(count))
    return out


def prune_pairs(sites, rdelta, delta_norm, count):
    s, r = _f32(sites), _f32(rdelta)
    out = np.zeros((count, 2), np.uint32)
    lib().orc_prune_pairs(_p(s), _p(r), _p(out), U32(s.shape[0]), F32(delta_norm), U32(count))
    return out


def radix_sort(pairs, max_key_exclusive=0xFFFFFFFF):
    p = _u32(pairs).copy()
    lib().orc_radix_sort(_p(p), U32(p.shape[0]), U32(max_key_exclusive))
    return p


def split_sites(sites, state, idx, stats, cur, target, width, height):
    """stats: float32 [8][N] (mass, energy, w, wx, wy, wxx, wxy, wyy)."""
    s = _f32(sites).copy(); st = _f32(state).copy(); ix = _u32(idx); t = _f32(target)
    sts = [_f32(x) for x in stats]
    lib().orc_split_sites(_p(s), _p(st), _p(ix), U32(ix.shape[0]), _p(sts[0]), _p(sts[2]), _p(sts[3]), _p(sts[4]),
                          _p(sts[5]), _p(sts[6]), _p(sts[7]), U32(cur), _p(t), I32(width), I32(height))
    return s, st


def hilbert_pairs(sites, padded, width, height, bits):
    s = _f32(sites); out = np.zeros((padded, 2), np.uint32)
    lib().orc_hilbert_pairs(_p(s), _p(out), U32(s.shape[0]), U32(padded), I32(width), I32(height), I32(bits))
    return out


def site_stats(cand0, cand1, target, mask, sites, iss, width, height):
    c0, c1, t, m, s = _u32(cand0), _u32(cand1), _f32(target), _f32(mask), _f32(sites)
    out = np.zeros((s.shape[0], 8))
    lib().orc_site_stats(_p(c0), _p(c1), _p(t), _p(m), _p(s), F32(iss), U32(s.shape[0]), I32(width), I32(height),
                         I32(width), I32(height), _p(out))
    return out


def sse(rendered, target, mask):
    r, t, m = _f32(rendered), _f32(target), _f32(mask)
    h, w = m.shape
    return lib().orc_sse(_p(r), _p(t), _p(m), I32(w), I32(h))


# ----------------------------------------------------------------------- fit
def _cfg(opts: dict) -> OrcFitConfig:
    c = OrcFitConfig()
    for k, _ in OrcFitConfig._fields_:
        setattr(c, k, opts[k])
    return c


def simulate_final_sites(init_sites, max_sites, iters, densify_enabled, d_start, d_end, d_freq, d_pct,
                         prune_during, p_start, p_end, p_freq, p_pct, max_split=65536):
    """simulateFinalSites (sad_cuda.cu:904-960): the active-site count a schedule ends with"""
    return int(lib().orc_simulate_final(I32(init_sites), I32(max_sites), I32(iters), I32(int(densify_enabled)),
                                        I32(d_start), I32(d_end), I32(d_freq), F32(d_pct), I32(int(prune_during)),
                                        I32(p_start), I32(p_end), I32(p_freq), F32(p_pct), I32(max_split)))


def solve_target_bpp(target_bpp, width, height, init_sites, max_sites, iters, densify_enabled, d_start, d_end,
                     d_freq, base_d, prune_during, p_start, p_end, p_freq, base_p, max_split=65536):
    """solveTargetBpp (sad_cuda.cu:962-1066) -> (densify pct, prune pct, simulated final sites)"""
    d, p, n = C.c_float(), C.c_float(), C.c_int()
    lib().orc_solve_target_bpp(F32(target_bpp), I32(width), I32(height), I32(init_sites), I32(max_sites), I32(iters),
                               I32(int(densify_enabled)), I32(d_start), I32(d_end), I32(d_freq), F32(base_d),
                               I32(int(prune_during)), I32(p_start), I32(p_end), I32(p_freq), F32(base_p),
                               I32(max_split), C.byref(d), C.byref(p), C.byref(n))
    return d.value, p.value, n.value


def fit_plan(opts: dict, width: int, height: int) -> dict:
    c = _cfg(opts)
    d, p, cap, pc = C.c_float(), C.c_float(), C.c_int(), C.c_int()
    lib().orc_fit_plan(C.byref(c), I32(width), I32(height), C.byref(d), C.byref(p), C.byref(cap), C.byref(pc))
    return dict(densify_percentile=d.value, prune_percentile=p.value, capacity=cap.value, pair_count=pc.value)


def fit(target, sites, opts: dict, mask=None) -> dict:
    """sites: initial sites [N][10], or None with init_mode 2 (gradient-weighted)."""
    h, w, _ = target.shape
    o = dict(opts)
    if o["init_mode"] == 0:
        o["n_sites"] = int(sites.shape[0])
    plan = fit_plan(o, w, h)
    n_in = 0 if sites is None else sites.shape[0]
    cap = max(plan["capacity"], n_in)
    buf = np.zeros((cap, 10), np.float32)
    buf[:, 0:2] = -1.0
    if n_in:
        buf[:n_in] = sites
    t = _f32(target)
    m = None if mask is None else _f32(mask)
    logs = np.zeros(64, np.float32)
    slogs = np.zeros(64, np.float32)
    res = OrcFitResult()
    rc = lib().orc_fit(C.byref(_cfg(o)), _p(t), _p(m), I32(w), I32(h), _p(buf), I32(cap), _p(logs), I32(64),
                       C.byref(res), _p(slogs))
    if rc != 0:
        raise RuntimeError(f"orc_fit failed: {rc}")
    out = {k: getattr(res, k) for k, _ in OrcFitResult._fields_}
    out["sites"] = buf
    out["log_psnr"] = logs[: min(res.n_logs, 64)].copy()
    out["log_ssim"] = slogs[: min(res.n_logs, 64)].copy()
    out["plan"] = plan
    return out


def init_gradient_weighted(target, n, mask=None, threshold=0.01, attempts=256, log_tau0=5.0, radius0=4.0, seed0=0):
    """-> (sites [n][10], one past the last seed drawn)"""
    t = _f32(target)
    h, w, _ = t.shape
    m = None if mask is None else _f32(mask)
    out = np.zeros((n, 10), np.float32)
    end = lib().orc_init_gradient_weighted(_p(out), U32(n), U32(seed0), _p(t), _p(m), F32(threshold), U32(attempts),
                                           F32(log_tau0), F32(radius0), I32(w), I32(h))
    return out, int(end)


def init_grad_eval(target, seed, mask=None, threshold=0.01):
    """-> (accepted, x, y, pixel index, gradient magnitude) of one R2 seed"""
    t = _f32(target)
    h, w, _ = t.shape
    m = None if mask is None else _f32(mask)
    o = np.zeros(4, np.float32)
    ok = lib().orc_init_grad_eval(_p(t), _p(m), I32(w), I32(h), U32(seed), F32(threshold), _p(o))
    return bool(ok), float(o[0]), float(o[1]), int(o[2]), float(o[3])


def ssim_sum(rendered, target, mask=None):
    """-> (sum over masked pixels of the window SSIM, per-pixel SSIM [H][W])"""
    r, t = _f32(rendered), _f32(target)
    h, w, _ = r.shape
    m = np.ones((h, w), np.float32) if mask is None else _f32(mask)
    per = np.zeros((h, w), np.float32)
    s = lib().orc_ssim_sum(_p(r), _p(t), _p(m), I32(w), I32(h), _p(per))
    return float(s), per


def ssim_dcentre(rendered, target, mask=None):
    """dSSIM(window of k)/dx_k per pixel [H][W]"""
    r, t = _f32(rendered), _f32(target)
    h, w, _ = r.shape
    m = np.ones((h, w), np.float32) if mask is None else _f32(mask)
    out = np.zeros((h, w), np.float32)
    lib().orc_ssim_dcentre(_p(r), _p(t), _p(m), I32(w), I32(h), _p(out))
    return out


QUANT_FIELDS = ("log_tau_min", "log_tau_scale", "radius_min", "radius_scale", "r_min", "r_scale", "g_min", "g_scale",
                "b_min", "b_scale")


def pack_inference(sites, width, height):
    """-> (packed uint32 [N][4], quant float32 [10] in QUANT_FIELDS order)"""
    s = _f32(sites)
    n = s.shape[0]
    out = np.zeros((n, 4), np.uint32)
    q = np.zeros(10, np.float32)
    lib().orc_pack_inference(_p(s), U32(n), I32(width), I32(height), _p(out), _p(q))
    return out, q


def unpack_inference(packed, quant, width, height):
    p, q = _u32(packed), _f32(quant)
    n = p.shape[0]
    out = np.zeros((n, 10), np.float32)
    lib().orc_unpack_inference(_p(p), _p(q), U32(n), I32(width), I32(height), _p(out))
    return out


def atan2(y, x):
    y, x = _f32(y), _f32(x)
    out = np.zeros_like(y)
    lib().orc_atan2_array(_p(y), _p(x), _p(out), I64(y.size))
    return out
