"""Benchmark: fit megapixel-iterations/s of the SAD fit loop on one MI355X.

Workload "fit1024": 1024x1024 seeded synthetic target (400 ellipses + 64-cycle
sinusoid patch), 16,384 initial sites, densify/prune every 100 iterations up
to a 262,144-site buffer, target 1.0 bpp, 2,000 iterations, fp32.  One step =
one complete fit (all 2,000 iterations: candidate updates, gradients, tau
diffusion, Adam, densify/prune with radix sorts, log renders).

value = sum(W*H/1e6 * iterations) / sum(fit-loop seconds) over the timed steps;
the loop time starts after the target and initial sites are resident in HBM.
roofline: the tiled gradient kernel (the dominant kernel), algorithmic bytes
48*P + 120*N per launch (+8*N on prune iterations) over its average duration,
measured with HIP events on its stream inside one more fit after the timed
steps (the timed fits record no per-launch events); traffic: HBM bytes per
launch of the same kernel from rocprofv3 FETCH_SIZE/WRITE_SIZE passes, scaled
by a copy kernel of known size (profiles/k_pixel_grad_<workload>_traffic.md),
used only while that summary's kernel-source hash matches the current source.
cpu_baseline: the CPU oracle (C, OpenMP, every thread the host grants) on the
first 120 iterations of the same fit (the random-field opening and the first
densify and prune), as a bounded sample.  The fits are deterministic; the line
reports whether the timed fits ended with the oracle's own result for the
workload (tests/golden/fit1024_full.npz).

Other workloads (not the headline line):
  --workload fit256      the 256x256 / 2,048-site plain fit (100 iterations);
                         with --cpu-iters 100 the CPU leg is the whole fit.
  --workload render4096  per-kernel timings at 4096x4096 / 1,048,576 sites
                         (sad_amd.kernel_bench) with the CPU oracle on the same
                         inputs; the line carries every kernel's row.
  --workload batch64     64 images (about 60 MP) fitted as one batch for 500
                         iterations, ticks launched on the stream
                         (--graph-ticks 1: steady ticks replayed as captured
                         hipGraphs), and one at a time on the same inputs.
  --workload smallbatch  --images N (256) images of --side S (128) px square,
                         S*S/32 initial sites each, 100 iterations without
                         densify: the batched path against the same fits one
                         at a time, where single fits cannot fill the device.
  --workload serve       a Poisson stream of fit and render requests replayed
                         against the continuous-batching server: throughput
                         and p50/p99 latency per request kind.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the gradient kernel's source and every header it includes
KERNEL_SOURCES = [os.path.join(ROOT, *p.split("/")) for p in (
    "sad_amd/csrc/k_pixel.hip", "sad_amd/csrc/fit_common.h", "sad_amd/csrc/batch.h", "sad_amd/csrc/sad_device.h",
    "sad_amd/csrc/sad_math.h", "include/sad_hip.h", "include/sad_types.h")]


def kernel_source_sha() -> str:
    """first 16 hex digits of the SHA-256 of the gradient kernel's source files, concatenated in a fixed order"""
    import hashlib
    h = hashlib.sha256()
    for path in KERNEL_SOURCES:
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profile_value(name: str, key: str):
    """`key: value` from a committed rocprofv3 summary under profiles/, or None when the summary is
    missing or was measured on another version of the kernel's source (its kernel_source_sha256 line)"""
    path = os.path.join(ROOT, "profiles", name)
    vals = {}
    try:
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith(("|", "#", " ")):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
    except OSError:
        return None
    if vals.get("kernel_source_sha256") != kernel_source_sha() or key not in vals:
        return None
    return vals[key]


def pmc_traffic(workload: str):
    """HBM bytes per gradient launch (FETCH_SIZE + WRITE_SIZE of the kernel over the same fit, each
    scaled by what the counter reads for a 256 MiB copy), or None"""
    v = profile_value(f"k_pixel_grad_{workload}_traffic.md", "traffic_bytes_per_launch")
    return None if v is None else int(v)


def pmc_valu_busy(workload: str):
    """VALU busy fraction of the gradient kernel (SQ_ACTIVE_INST_VALU x 4 over SIMDs x busy cycles), or None"""
    v = profile_value(f"k_pixel_grad_{workload}_valu.md", "valu_busy")
    return None if v is None else float(v)


def host_threads() -> int:
    """threads for the CPU oracle: the CPUs this process may run on, capped by OMP_NUM_THREADS when the
    host sets it (a shared GPU host grants each GPU a share of its cores that way)"""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def golden_match(workload: str, sites) -> bool | None:
    """whether `sites` equal the oracle's final sites of the workload's full fit (SHA-256 of the buffer);
    None when no such fixture exists"""
    import hashlib

    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", f"{workload}_full.npz")
    if not os.path.exists(path):
        return None
    ref = np.load(path)
    return hashlib.sha256(np.ascontiguousarray(sites).tobytes()).hexdigest() == str(ref["sites_sha"])


def kernel_workload(args) -> int:
    """--workload render4096: per-kernel timings at 4096x4096 / 1,048,576 sites
    (sad_amd.kernel_bench), and the CPU oracle on the same inputs."""
    from sad_amd import kernel_bench
    inputs: dict = {}
    rows = kernel_bench.run(keep_inputs=inputs)
    cpu = None
    if not args.no_cpu:
        import numpy as np
        from oracle import orc
        threads = host_threads()
        orc.set_threads(threads)
        i = inputs
        w, h, n = i["width"], i["height"], len(i["sites"])
        timings = {}
        t0 = time.perf_counter()
        orc.render(i["cand0"], i["cand1"], i["sites"], i["iss"], w, h)
        timings["render"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.gradients(i["cand0"], i["cand1"], i["target"], np.ones((h, w), np.float32), i["sites"], i["iss"], w, h)
        timings["gradient"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.update_candidates(i["cand0"], i["cand1"], orc.pack_sites(i["sites"]), (1 << 16) | 100, i["iss"], w, h,
                              inject_count=4, hilbert_order=i["hilbert_order"], hilbert_pos=i["hilbert_pos"],
                              hilbert_probes=1, hilbert_window=32)
        timings["vpt"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.adam(i["sites"], np.zeros((n, 20), np.float32), np.zeros((10, n), np.float32), i["lrs"], 0.9, 0.999,
                 1e-8, 1, w, h)
        timings["adam"] = time.perf_counter() - t0
        cpu = {"value": round(w * h / 1e6 / timings["gradient"], 3), "unit": "MP/s (one gradient backward)",
               "cores": threads, "kind": "port",
               "sample": "one launch each of render, gradient, VPT pass, Adam on the same 4096x4096 inputs",
               "seconds": {k: round(v, 4) for k, v in timings.items()}}
    grad = next(r for r in rows if r["kernel"] == "k_pixel<GRAD>")
    line = {
        "metric": "gradient backward GB/s (algorithmic) at 4096x4096 / 1,048,576 sites",
        "value": grad["gb_s"], "unit": "GB/s", "n_gpus": 1, "steps": 20, "warmup": 1,
        "ms_per_step": round(grad["median_us"] / 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "render4096", "width": 4096, "height": 4096, "sites": 1 << 20},
        "roofline": {"bound": "hbm", "achieved": grad["gb_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": grad["frac_hbm"], "traffic": None, "kernel": "k_pixel<MODE_GRAD>"},
        "kernels": rows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def batch_workload(args) -> int:
    """--workload batch64: 64 images (sides log-uniform in [128, 2048], aspect
    log-uniform in [0.5, 2], bpp 1.0 initial sites), 500 iterations, fitted as
    one batch (sadh_fit_batch) and, for comparison, one image at a time
    (sadh_fit) on the same inputs.  --workload smallbatch: --images square
    images of --side px with side*side/32 sites each, 100 iterations."""
    import numpy as np
    import torch

    from sad_amd import config, synth, trainer
    torch.cuda.set_device(0)
    if args.workload == "batch64":
        targets, sites = synth.batch_workload()
        iters = 500
    else:
        side = args.side
        targets = [synth.make_target(side, side, 100 + i, 40) for i in range(args.images)]
        sites = [synth.init_sites_uniform(t, side * side // 32, 100 + i) for i, t in enumerate(targets)]
        iters = 100
    opts = config.fit_options(iterations=iters, target_bpp=0.0, densify_enabled=0, densify_percentile=0.0,
                              prune_percentile=0.0, log_freq=10**9)
    # the timed batches record no per-launch events (a profiled tick is never graph-replayed); one more
    # batch with profile_gradients times the batched gradient launches for the roofline
    bopts = dict(opts, graph_ticks=args.graph_ticks, autotune=args.autotune)
    mp = sum(t.shape[0] * t.shape[1] for t in targets) / 1e6
    for _ in range(args.warmup):
        trainer.fit_batch(targets[:4], sites[:4], bopts)
    torch.cuda.synchronize()
    runs = [trainer.fit_batch(targets, sites, bopts) for _ in range(args.steps)]
    prof = trainer.fit_batch(targets, sites, dict(opts, profile_gradients=1))
    loop_s = sum(r["train_seconds"] for r in runs)
    value = mp * iters * args.steps / loop_s
    last = runs[-1]
    single_s, dpsnr, all_same = 0.0, 0.0, True
    for i, (t, s0) in enumerate(zip(targets, sites)):
        r = trainer.fit(t, s0, opts)
        single_s += r["train_seconds"]
        dpsnr = max(dpsnr, abs(r["final_psnr"] - last["jobs"][i]["final_psnr"]))
        all_same = all_same and np.array_equal(r["sites"].view(np.uint32), last["jobs"][i]["sites"].view(np.uint32))
    single = mp * iters / single_s
    gms, gb = prof["grad_kernel_ms"], prof["grad_bytes"]
    achieved = gb / (gms * 1e-3) / 1e9 if gms > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        from oracle import orc
        threads = host_threads()
        orc.set_threads(threads)
        k, ci = 4, max(1, args.cpu_iters // 6)
        c0 = time.perf_counter()
        for t, s0 in zip(targets[:k], sites[:k]):
            orc.fit(t, s0, dict(opts, iterations=ci, n_sites=s0.shape[0], profile_gradients=0))
        cs = time.perf_counter() - c0
        cmp = sum(t.shape[0] * t.shape[1] for t in targets[:k]) / 1e6
        cpu = {"value": cmp * ci / cs, "unit": "MP-iters/s", "cores": threads, "kind": "port",
               "sample": f"first {ci} iterations of the first {k} images ({cmp:.2f} MP, {cs:.1f} s, OpenMP oracle)"}
    workload = {"workload": args.workload}
    if args.workload == "smallbatch":
        workload.update(width=args.side, height=args.side)
    line = {
        "metric": "fit megapixel-iters/sec (fwd+bwd+Adam)", "value": round(value, 3), "unit": "MP-iters/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * loop_s / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {**workload, "images": len(targets), "megapixels": round(mp, 3), "iterations": iters,
                   "sites": int(sum(len(s0) for s0 in sites)), "graph_ticks": last["graph_ticks"],
                   "graph_captures": last["graph_captures"],
                   "autotune": {"hash_slots": last["tuned_hash_slots"], "tau_block": last["tuned_tau_block"]}
                   if args.autotune else None,
                   "jobs_bit_identical_to_single": bool(all_same),
                   "mean_final_psnr_db": round(float(np.mean([j["final_psnr"] for j in last["jobs"]])), 3)},
        "single_image_path": {"value": round(single, 3), "unit": "MP-iters/s", "seconds": round(single_s, 3),
                              "max_final_psnr_diff_db": round(dpsnr, 4)},
        "speedup_vs_single_image_path": round(value / single, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_pixel_grad_batch", "avg_launch_us": round(1e3 * gms / max(1, prof["grad_launches"]), 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def serve_workload(args) -> int:
    """--workload serve: a Poisson stream (16 requests/s, 64 requests) of 70%
    fits (sides 256-2048, bpp 1.0, 300 iterations) and 30% renders (4k-1M
    random sites, 512^2-4096^2), replayed in wall-clock time against the
    continuous-batching server; throughput and p50/p99 latency per kind."""
    import torch

    from sad_amd import config, serve, synth
    torch.cuda.set_device(0)
    reqs = synth.serve_workload(n_requests=args.requests, rate=args.rate)
    opts = config.fit_options(iterations=300, target_bpp=1.0, log_freq=10**9)
    warm = synth.serve_workload(n_requests=4, rate=1000.0, seed=99, fit_sides=(128, 256), render_sides=(256, 512),
                                render_sites=(1024, 4096))
    for _ in range(args.warmup):
        serve.run_stream(warm, opts)
    r = serve.run_stream(reqs, opts)
    line = {
        "metric": "fit megapixel-iters/sec (fwd+bwd+Adam)", "value": round(r["fit_mp_iters_per_s"], 3),
        "unit": "MP-iters/s", "n_gpus": 1, "steps": 1, "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["wall_s"], 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serve", "requests": len(reqs), "rate_per_s": args.rate, "fits": r["n_fit"],
                   "renders": r["n_render"], "fit_iterations": 300},
        "latency_s": {"fit_p50": r["fit_latency_p50_s"], "fit_p99": r["fit_latency_p99_s"],
                      "render_p50": r["render_latency_p50_s"], "render_p99": r["render_latency_p99_s"]},
        "render_mp_per_s": round(r["render_mp_per_s"], 3),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fit1024", choices=["fit1024", "fit256", "render4096", "batch64", "smallbatch", "serve"])
    ap.add_argument("--autotune", type=int, default=0, choices=[0, 1],
                    help="batch64, smallbatch: time steady ticks under each result-neutral launch schedule "
                         "(tile hash size, tau block size) and keep the fastest")
    ap.add_argument("--images", type=int, default=256, help="smallbatch: images in the batch")
    ap.add_argument("--side", type=int, default=128, help="smallbatch: image side in pixels")
    ap.add_argument("--requests", type=int, default=64, help="serve: requests in the stream")
    ap.add_argument("--rate", type=float, default=16.0, help="serve: Poisson arrival rate per second")
    ap.add_argument("--cpu-iters", type=int, default=120)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-ticks", type=int, default=0, choices=[0, 1],
                    help="batch64, smallbatch: replay steady ticks as captured hipGraphs (1) or launch them on the stream (0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the fit path runs on one device)",
              file=sys.stderr)
        return 2

    if args.workload == "render4096":
        return kernel_workload(args)
    if args.workload in ("batch64", "smallbatch"):
        return batch_workload(args)
    if args.workload == "serve":
        return serve_workload(args)

    import numpy as np
    import torch

    from sad_amd import config, synth, trainer

    torch.cuda.set_device(0)
    tgt, sites, over = synth.fit_workload(args.workload)
    h, w, _ = tgt.shape
    opts = config.fit_options(**over)
    iters = opts["iterations"]
    mpix = w * h / 1e6

    for _ in range(args.warmup):
        trainer.fit(tgt, sites, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runs = [trainer.fit(tgt, sites, opts) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    loop_s = sum(r["train_seconds"] for r in runs)
    value = mpix * iters * args.steps / loop_s
    last = runs[-1]
    # The gradient kernel's launch time: one more fit, outside the timed steps, with HIP events recorded on the
    # kernel's stream around every gradient launch (each record costs the queue a few microseconds, so the
    # timed steps run without them).
    prof = trainer.fit(tgt, sites, dict(opts, profile_gradients=1))
    gms, glaunch, gbytes = prof["grad_kernel_ms"], prof["grad_launches"], prof["grad_bytes"]
    achieved = gbytes / (gms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle import orc
        threads = host_threads()
        orc.set_threads(threads)
        copts = dict(opts)
        cpu_iters = min(args.cpu_iters, iters)
        copts["iterations"] = cpu_iters
        copts["log_freq"] = 10**9
        # keep the densify/prune plan of the full fit: same percentiles, same buffers
        plan = trainer.fit_plan(opts, w, h)
        copts.update(target_bpp=0.0, densify_percentile=plan["densify_percentile"],
                     prune_percentile=plan["prune_percentile"])
        c0 = time.perf_counter()
        orc.fit(tgt, sites, copts)
        cs = time.perf_counter() - c0
        edits = ", through the first densify and prune" if opts["densify_enabled"] and cpu_iters > 100 else ""
        cpu = {"value": mpix * cpu_iters / cs, "unit": "MP-iters/s", "cores": threads, "kind": "port",
               "sample": f"first {cpu_iters} of the {iters} iterations of the {args.workload} fit{edits} "
                         f"({cs:.1f} s, OpenMP oracle)"}

    line = {
        "metric": "fit megapixel-iters/sec (fwd+bwd+Adam)",
        "value": round(value, 3),
        "unit": "MP-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * loop_s / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "width": w, "height": h, "init_sites": int(sites.shape[0]),
                   "iterations": iters, "final_sites": last["actual_sites"], "active_sites": last["active_count"],
                   "final_psnr_db": round(last["final_psnr"], 4),
                   "final_psnr_db_steps": [round(r["final_psnr"], 4) for r in runs], "wall_s": round(wall, 3),
                   "steps_bit_identical": all(np.array_equal(r["sites"].view(np.uint32), last["sites"].view(np.uint32))
                                              for r in runs),
                   "oracle_golden_match": golden_match(args.workload, last["sites"])},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.workload),
                     "algorithmic_bytes_per_launch": int(gbytes / max(1, glaunch)),
                     "kernel": "k_pixel<MODE_GRAD> (compute_gradients_tiled)",
                     "avg_launch_us": round(1e3 * gms / max(1, glaunch), 3),
                     "valu_busy": pmc_valu_busy(args.workload)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
