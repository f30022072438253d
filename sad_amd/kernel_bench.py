"""Per-kernel timings of the bandwidth-bound render + single-backward workload.

Workload "render4096": a 4096x4096 seeded synthetic target (the fit
generator, 40 ellipses) and 1,048,576 sites with uniform positions,
log-scales l1, l2 ~ N(log 4, 0.5^2) (radius exp((l1+l2)/2), log_aniso
clamp(l2-l1, -4, 4)), direction angle U[0, pi), colour U[0,1]^3, log_tau 5.
The candidate field is built the render-job way (clear, JFA seed, 12 flood
passes, VPT passes with Hilbert probes), then each kernel is launched `reps`
times on its own and timed with HIP events on the stream it runs on; the
median is reported with its algorithmic bytes, GB/s and fraction of the
8 TB/s HBM peak.

    python -m sad_amd.kernel_bench [--size 4096] [--sites 1048576] [--reps 20]

Algorithmic bytes (P pixels = candidate cells, N sites, n sort pairs):
  JFA flood pass        32 P + 40 N        VPT pass           64 P + 24 N
  render forward        44 P + 40 N        gradient backward  48 P + 120 N
  tau diffusion pass    52 N               Adam + clear       320 N
  radix sort (4 passes) 96 n
`python bench.py --workload render4096` prints these rows and times the CPU
oracle on the same inputs (keep_inputs) as its baseline.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

HBM_PEAK_GBS = 8000.0


def _args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--sites", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--json", default=None, help="also write the rows to this file")
    return ap.parse_args(argv)


def run(size: int = 4096, n_sites: int = 1 << 20, reps: int = 20, seed: int = 3,
        keep_inputs: dict | None = None) -> list[dict]:
    import torch

    from . import _lib, synth

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    lib = _lib.load()
    stream = torch.cuda.current_stream()
    st = stream.cuda_stream

    def call(name, *a):
        rc = getattr(lib, name)(*[x.data_ptr() if hasattr(x, "data_ptr") else x for x in a], st)
        if rc != 0:
            raise _lib.SadHipError(f"{name}: {lib.sadh_last_error().decode()}")

    w = h = size
    n = n_sites
    P = w * h
    m = np.float32(max(w, h))
    iss = float(np.float32(1.0) / (m * m))
    target = synth.make_target(w, h, seed, 40)
    sites = synth.init_sites_random(w, h, n, seed)
    gt = torch.from_numpy(target).to(dev)
    gs = torch.from_numpy(sites).to(dev)
    gm = torch.ones((h, w), device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    a0, a1 = torch.empty((P, 4), **i32), torch.empty((P, 4), **i32)
    b0, b1 = torch.empty((P, 4), **i32), torch.empty((P, 4), **i32)

    def timed(fn, label, nbytes, units):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e3)
        us = float(np.median(ts))
        gbs = nbytes / (us * 1e-6) / 1e9
        return {"kernel": label, "median_us": round(us, 2), "min_us": round(float(np.min(ts)), 2),
                "bytes": int(nbytes), "units": units, "gb_s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}

    rows = []
    # ---- candidate field: init + seed + flood (render-job recipe)
    call("sadh_clear_candidates", a0, a1, w, h)
    call("sadh_jfa_seed", a0, gs, n, w, h, 1)
    steps = []
    s = 1 << int(np.ceil(np.log2(max(w, h))))
    s >>= 1
    while s >= 1:
        steps.append(s)
        s >>= 1
    seeded = a0.clone()
    rows.append(timed(lambda: call("sadh_jfa_flood", seeded, b0, gs, n, steps[0], iss, w, h, 1, w, h),
                      f"k_jfa_flood (step {steps[0]})", 32 * P + 40 * n, f"P={P}, N={n}"))
    src, dst = a0, b0
    for stp in steps:
        call("sadh_jfa_flood", src, dst, gs, n, stp, iss, w, h, 1, w, h)
        src, dst = dst, src
    a0 = src
    # ---- Hilbert tables, packed sites, VPT passes
    bits = max(1, int(np.ceil(np.log2(max(w, h)))))
    padded = (n + 1023) // 1024 * 1024
    hp = torch.empty((padded, 2), **i32)
    hs = torch.empty((padded, 2), **i32)
    hh = torch.empty((256 * (padded // 1024),), **i32)
    call("sadh_hilbert_pairs", gs, hp, n, padded, w, h, bits)
    call("sadh_radix_sort_pairs", hp, hs, hh, hh, padded, 0xFFFFFFFF if bits >= 16 else (1 << (2 * bits)))
    order, pos = torch.empty(n, **i32), torch.empty(n, **i32)
    call("sadh_hilbert_order", hp, order, pos, n)
    pk = torch.empty((n, 8), dtype=torch.float16, device=dev)
    rows.append(timed(lambda: call("sadh_pack_candidate_sites", gs, pk, n), "k_pack_sites", 56 * n, f"N={n}"))

    def vpt(c0, c1, o0, o1, idx, inject=4, hil_probes=1):
        j = max(1, (1 << (bits - 1)) >> min(idx, bits - 1))
        call("sadh_update_candidates", c0, c1, o0, o1, pk, n, (j << 16) | (idx & 0xFFFF), iss, 0, 64.0, 0, inject,
             order, pos, hil_probes, 32, 1, w, h, w, h)

    for idx in range(bits + 4):
        vpt(a0, a1, b0, b1, idx)
        a0, b0, a1, b1 = b0, a0, b1, a1
    rows.append(timed(lambda: vpt(a0, a1, b0, b1, 100), "k_vpt (jump 1, Hilbert 1/32, inject 4)",
                      64 * P + 24 * n, f"P={P}, N={n}"))
    # the same pass without its random-id sources: Hilbert neighbours, then also the injected ids
    rows.append(timed(lambda: vpt(a0, a1, b0, b1, 100, 4, 0), "k_vpt (jump 1, no Hilbert, inject 4)",
                      64 * P + 24 * n, f"P={P}, N={n}"))
    rows.append(timed(lambda: vpt(a0, a1, b0, b1, 100, 0, 0), "k_vpt (jump 1, no Hilbert, no injection)",
                      64 * P + 24 * n, f"P={P}, N={n}"))
    # ---- render, gradient, tau, Adam
    out = torch.empty((h, w, 3), device=dev)
    rows.append(timed(lambda: call("sadh_render_voronoi", a0, a1, out, gs, iss, n, w, h, w, h),
                      "k_pixel<RENDER>", 44 * P + 40 * n, f"P={P}, N={n}"))
    grads = [torch.zeros(n, device=dev) for _ in range(10)]
    rd = torch.zeros(n, device=dev)
    rows.append(timed(lambda: call("sadh_compute_gradients_tiled", a0, a1, gt, gt, gm, *grads, gs, iss, n, rd, 0, 0.0,
                                   w, h, w, h), "k_pixel<GRAD>", 48 * P + 120 * n, f"P={P}, N={n}"))
    go = torch.randn((h, w, 3), device=dev)
    gaos = torch.zeros((n, 10), device=dev)
    rows.append(timed(lambda: call("sadh_render_backward", a0, a1, go, gaos, gs, iss, n, w, h, w, h),
                      "k_pixel<BWD> (autograd)", 44 * P + 120 * n, f"P={P}, N={n}"))
    raw, tmp = grads[2].clone(), torch.zeros(n, device=dev)
    rows.append(timed(lambda: call("sadh_tau_diffuse", a0, a1, gs, raw, grads[2], tmp, n, 0.05, w, h, 1),
                      "k_tau_diffuse (one pass)", 52 * n, f"N={n}"))
    adam = torch.zeros((n, 20), device=dev)
    lrs = (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025)
    sites_copy = gs.clone()
    rows.append(timed(lambda: call("sadh_adam_update_clear", sites_copy, adam, *grads, *lrs, 0.9, 0.999, 1e-8, 1, n,
                                   w, h), "k_adam<clear>", 320 * n, f"N={n}"))
    pairs = torch.from_numpy(np.stack([np.random.default_rng(seed).integers(0, 2**32, n, dtype=np.uint64)
                                       .astype(np.uint32), np.arange(n, dtype=np.uint32)], 1).view(np.int32)).to(dev)
    work = pairs.clone()
    scratch = torch.empty_like(pairs)
    hist = torch.empty((256 * ((n + 1023) // 1024),), **i32)

    def sort():
        work.copy_(pairs)
        call("sadh_radix_sort_pairs", work, scratch, hist, hist, n, 0xFFFFFFFF)
    rows.append(timed(sort, "radix sort (4 passes, incl. 8 B/pair reset copy)", 96 * n + 16 * n, f"n={n}"))

    if keep_inputs is not None:
        keep_inputs.update(target=target, sites=sites, iss=iss, width=w, height=h,
                           cand0=a0.cpu().numpy().view(np.uint32), cand1=a1.cpu().numpy().view(np.uint32),
                           hilbert_order=order.cpu().numpy().view(np.uint32),
                           hilbert_pos=pos.cpu().numpy().view(np.uint32), lrs=lrs)
    return rows


def main(argv=None) -> int:
    a = _args(argv)
    rows = run(a.size, a.sites, a.reps, a.seed)
    for r in rows:
        print(json.dumps(r), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
