"""Host-side driver of the native fit and render loops (sadh_fit,
sadh_render_sites): numpy in, numpy out."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import fit_options, to_struct


def fit_plan(opts: dict, width: int, height: int) -> dict:
    s = to_struct(opts)
    d, p, cap, pairs = C.c_float(), C.c_float(), C.c_int(), C.c_int()
    _lib.call("sadh_fit_plan", C.byref(s), width, height, C.byref(d), C.byref(p), C.byref(cap), C.byref(pairs))
    return dict(densify_percentile=d.value, prune_percentile=p.value, capacity=cap.value, pair_count=pairs.value)


def fit(target: np.ndarray, sites: np.ndarray | None, opts: dict | None = None, mask: np.ndarray | None = None,
        want_render: bool = False, debug_images: bool = False, **over) -> dict:
    """Run the fit on the current HIP device.

    target float32 [H][W][3]; sites float32 [N][10] initial sites (None with
    init_mode 2, where the fit makes n_sites gradient-weighted sites).  Returns
    a dict with the final sites [capacity][10], logged PSNRs and the
    SadhFitResult fields."""
    o = dict(opts or fit_options())
    o.update(over)
    h, w, _ = target.shape
    if o["init_mode"] == 0:
        o["n_sites"] = int(sites.shape[0])
    plan = fit_plan(o, w, h)
    n_in = 0 if sites is None else sites.shape[0]
    cap = max(plan["capacity"], n_in)
    buf = np.zeros((cap, 10), np.float32)
    buf[:, 0:2] = -1.0
    if n_in:
        buf[:n_in] = sites
    tgt = np.ascontiguousarray(target, np.float32)
    msk = None if mask is None else np.ascontiguousarray(mask, np.float32)
    logs = np.zeros(64, np.float32)
    res = _lib.FitResult()
    trace = np.zeros(max(1, o["iterations"]), np.float32) if o.get("profile_gradients") else None
    if trace is not None:
        res.grad_ms_trace = trace.ctypes.data_as(C.POINTER(C.c_float))
        res.grad_ms_trace_cap = trace.size
    render = np.zeros((h, w, 3), np.float32) if want_render else None
    if render is not None:
        res.final_render = render.ctypes.data_as(C.POINTER(C.c_float))
    slogs = np.zeros(len(logs), np.float32)
    res.log_ssim = slogs.ctypes.data_as(C.POINTER(C.c_float))
    prender = np.zeros((h, w, 3), np.float32) if (want_render and o.get("packed_psnr")) else None
    if prender is not None:
        res.packed_render = prender.ctypes.data_as(C.POINTER(C.c_float))
    cells = np.zeros((h, w, 3), np.float32) if debug_images else None
    heat = np.zeros((h, w, 3), np.float32) if debug_images else None
    if debug_images:
        res.cells_render = cells.ctypes.data_as(C.POINTER(C.c_float))
        res.tau_heatmap = heat.ctypes.data_as(C.POINTER(C.c_float))
    s = to_struct(o)
    _lib.call("sadh_fit", C.byref(s), _lib.ptr(tgt), _lib.ptr(msk), w, h, _lib.ptr(buf), cap, _lib.ptr(logs),
              len(logs), C.byref(res))
    out = _lib.fit_result_dict(res)
    out["grad_ms_trace"] = trace
    out["render"] = render
    out["packed_render"] = prender
    out["cells"] = cells
    out["tau_heatmap"] = heat
    out["sites"] = buf
    out["log_psnr"] = logs[: min(res.n_logs, len(logs))].copy()
    out["log_ssim"] = slogs[: min(res.n_logs, len(logs))].copy()
    out["plan"] = plan
    return out


def fit_batch(targets: list, sites: list, opts: dict | None = None, masks: list | None = None,
              want_render: bool = False, **over) -> dict:
    """Fit many images in lock step with one launch per kernel per iteration
    (sadh_fit_batch).  targets[i] float32 [H][W][3], sites[i] float32 [N_i][10];
    all jobs share the options.  Returns {"jobs": [per-job dicts as fit()],
    "train_seconds", "grad_kernel_ms", "grad_launches", "grad_bytes", "total_pixels", "graph_ticks",
    "graph_captures"}."""
    o = dict(opts or fit_options())
    o.update(over)
    if o["init_mode"] != 0:
        raise ValueError("fit_batch takes caller-provided initial sites (init_mode 0)")
    n = len(targets)
    if n == 0 or len(sites) != n or (masks is not None and len(masks) != n):
        raise ValueError("targets, sites (and masks) must be non-empty lists of equal length")
    keep = []
    jobs = (_lib.BatchJob * n)()
    results = [_lib.FitResult() for _ in range(n)]
    outs = []
    for i, (tgt, st) in enumerate(zip(targets, sites)):
        h, w, _ = tgt.shape
        oi = dict(o, n_sites=int(st.shape[0]))
        plan = fit_plan(oi, w, h)
        cap = max(plan["capacity"], st.shape[0])
        buf = np.zeros((cap, 10), np.float32)
        buf[:, 0:2] = -1.0
        buf[: st.shape[0]] = st
        t = np.ascontiguousarray(tgt, np.float32)
        m = None if masks is None or masks[i] is None else np.ascontiguousarray(masks[i], np.float32)
        logs = np.zeros(64, np.float32)
        keep += [t, m, buf, logs]
        jobs[i].target = _lib.ptr(t)
        jobs[i].mask = _lib.ptr(m)
        jobs[i].width, jobs[i].height = w, h
        jobs[i].sites_io = _lib.ptr(buf)
        jobs[i].n_sites = int(st.shape[0])
        jobs[i].sites_cap = cap
        jobs[i].log_psnr = _lib.ptr(logs)
        jobs[i].log_cap = len(logs)
        jobs[i].result = C.pointer(results[i])
        slogs = np.zeros(len(logs), np.float32)
        results[i].log_ssim = slogs.ctypes.data_as(C.POINTER(C.c_float))
        render = np.zeros((h, w, 3), np.float32) if want_render else None
        if render is not None:
            results[i].final_render = render.ctypes.data_as(C.POINTER(C.c_float))
        prender = np.zeros((h, w, 3), np.float32) if (want_render and o.get("packed_psnr")) else None
        if prender is not None:
            results[i].packed_render = prender.ctypes.data_as(C.POINTER(C.c_float))
        keep += [slogs, render, prender]
        outs.append(dict(sites=buf, plan=plan, _logs=logs, _slogs=slogs, render=render, packed_render=prender))
    br = _lib.BatchResult()
    s = to_struct(o)
    _lib.call("sadh_fit_batch", C.byref(s), jobs, n, C.byref(br))
    for out, res in zip(outs, results):
        out.update(_lib.fit_result_dict(res))
        logs = out.pop("_logs")
        slogs = out.pop("_slogs")
        out["log_psnr"] = logs[: min(res.n_logs, len(logs))].copy()
        out["log_ssim"] = slogs[: min(res.n_logs, len(logs))].copy()
    ret = {k: getattr(br, k) for k, _ in _lib.BatchResult._fields_}
    ret["jobs"] = outs
    return ret


def render_sites(sites: np.ndarray, width: int, height: int, cand_passes: int = 16, use_jfa: bool = True,
                 jfa_rounds: int = 1, cand_downscale: int = 1, radius_scale: float = 64.0, radius_probes: int = 0,
                 inject_count: int = 4, hilbert_window: int = 32, hilbert_probes: int = 1, cells: bool = False) -> dict:
    """Render-only job: JFA + VPT candidate build, then the softmax render
    (renderVoronoiFromSites, backends/cuda/src/sad_cuda.cu:1678-1858)."""
    s = np.ascontiguousarray(sites, np.float32)
    out = np.zeros((height, width, 3), np.float32)
    cell = np.zeros((height, width, 3), np.float32) if cells else None
    cms, rms = C.c_float(), C.c_float()
    _lib.call("sadh_render_sites", _lib.ptr(s), s.shape[0], width, height, cand_passes, int(use_jfa), jfa_rounds,
              cand_downscale, radius_scale, radius_probes, inject_count, hilbert_window, hilbert_probes,
              _lib.ptr(out), C.byref(cms), C.byref(rms), _lib.ptr(cell))
    return dict(image=out, cand_ms=cms.value, render_ms=rms.value, cells=cell)


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 100.0 if mse <= 0 else 20.0 * np.log10(1.0 / np.sqrt(mse))


def pack_inference(sites: np.ndarray, width: int, height: int):
    """Quantised 16-byte inference sites of `sites` (sadh_pack_inference_sites, on
    the current HIP device).  -> (packed uint32 [N][4], quant float32 [10]:
    log_tau, radius, r, g, b as (min, scale) pairs)."""
    import torch
    s = torch.from_numpy(np.ascontiguousarray(sites, np.float32)).cuda()
    n = s.shape[0]
    packed = torch.zeros((max(n, 1), 4), dtype=torch.int32, device=s.device)
    quant = torch.zeros(10, dtype=torch.float32, device=s.device)
    _lib.call("sadh_pack_inference_sites", s.data_ptr(), n, width, height, packed.data_ptr(), quant.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return packed[:n].cpu().numpy().view(np.uint32), quant.cpu().numpy()


def render_packed_sites(packed: np.ndarray, quant: np.ndarray, width: int, height: int, cand_passes: int = 16,
                        use_jfa: bool = True, jfa_rounds: int = 1, cand_downscale: int = 1, radius_scale: float = 64.0,
                        radius_probes: int = 0, inject_count: int = 4, hilbert_window: int = 32,
                        hilbert_probes: int = 1, cells: bool = False) -> dict:
    """render_sites on quantised inference sites (sadh_render_packed_sites;
    renderPackedInference, backends/metal/sources/rendering/packed_renderer.swift:15-137)."""
    p = np.ascontiguousarray(packed, np.uint32)
    q = np.ascontiguousarray(quant, np.float32)
    out = np.zeros((height, width, 3), np.float32)
    cell = np.zeros((height, width, 3), np.float32) if cells else None
    cms, rms = C.c_float(), C.c_float()
    _lib.call("sadh_render_packed_sites", _lib.ptr(p), _lib.ptr(q), p.shape[0], width, height, cand_passes,
              int(use_jfa), jfa_rounds, cand_downscale, radius_scale, radius_probes, inject_count, hilbert_window,
              hilbert_probes, _lib.ptr(out), C.byref(cms), C.byref(rms), _lib.ptr(cell))
    return dict(image=out, cand_ms=cms.value, render_ms=rms.value, cells=cell)
