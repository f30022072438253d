"""Continuous-batching fit/render server on the native engine (sadh_server_*).

    srv = Server(fit_options(iterations=300, ...))
    fid = srv.submit_fit(target, sites)        # float32 [H][W][3], [N][10]
    rid = srv.submit_render(sites, 1024, 1024)
    while srv.outstanding():
        srv.step()                             # admit, one batched iteration of every running fit, retire
    srv.result(fid)["final_psnr"], srv.result(rid)["image"]

`run_stream` replays a timed request list (synth.serve_workload) against a
server in wall-clock time and reports throughput and p50/p99 latency per
request kind.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .config import fit_options, to_struct
from .trainer import fit_plan


class Server:
    def __init__(self, opts: dict | None = None, memory_budget: float = 0.0, max_fit_jobs: int = 64,
                 render_passes: int = 16):
        o = dict(opts or fit_options())
        o["init_mode"] = 0
        self.opts = o
        so = _lib.ServerOptions()
        so.fit = to_struct(o)
        so.memory_budget = float(memory_budget)
        so.max_fit_jobs = int(max_fit_jobs)
        so.render_passes = int(render_passes)
        h = C.c_void_p()
        _lib.call("sadh_server_create", C.byref(so), C.byref(h))
        self._h = h
        self._reqs: dict[int, dict] = {}
        self._pending = 0

    def submit_fit(self, target: np.ndarray, sites: np.ndarray, mask: np.ndarray | None = None) -> int:
        h, w, _ = target.shape
        plan = fit_plan(dict(self.opts, n_sites=int(sites.shape[0])), w, h)
        cap = max(plan["capacity"], sites.shape[0])
        buf = np.zeros((cap, 10), np.float32)
        buf[:, 0:2] = -1.0
        buf[: sites.shape[0]] = sites
        t = np.ascontiguousarray(target, np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, np.float32)
        logs = np.zeros(64, np.float32)
        res = _lib.FitResult()
        job = _lib.BatchJob()
        job.target, job.mask = _lib.ptr(t), _lib.ptr(m)
        job.width, job.height = w, h
        job.sites_io, job.n_sites, job.sites_cap = _lib.ptr(buf), int(sites.shape[0]), cap
        job.log_psnr, job.log_cap = _lib.ptr(logs), len(logs)
        job.result = C.pointer(res)
        rid = C.c_longlong()
        _lib.call("sadh_server_submit_fit", self._h, C.byref(job), C.byref(rid))
        self._reqs[rid.value] = dict(kind="fit", keep=(t, m, buf, logs, res, job), sites=buf, logs=logs, res=res,
                                     pixels=w * h, done=False)
        self._pending += 1
        return rid.value

    def submit_render(self, sites: np.ndarray, width: int, height: int) -> int:
        s = np.ascontiguousarray(sites, np.float32)
        out = np.zeros((height, width, 3), np.float32)
        rid = C.c_longlong()
        _lib.call("sadh_server_submit_render", self._h, _lib.ptr(s), s.shape[0], width, height, _lib.ptr(out),
                  C.byref(rid))
        self._reqs[rid.value] = dict(kind="render", keep=(s, out), image=out, pixels=width * height, done=False)
        self._pending += 1
        return rid.value

    def step(self) -> list[dict]:
        """Admit, advance, retire; returns the events of requests finished now."""
        n = C.c_int()
        _lib.call("sadh_server_step", self._h, C.byref(n))
        evs = (_lib.ServerEvent * 64)()
        out = []
        while True:
            k = C.c_int()
            _lib.call("sadh_server_poll", self._h, evs, 64, C.byref(k))
            for i in range(k.value):
                e = evs[i]
                r = self._reqs[e.id]
                r.update(done=True, submit_s=e.submit_s, admit_s=e.admit_s, done_s=e.done_s)
                self._pending -= 1
                out.append(dict(id=e.id, kind=r["kind"], submit_s=e.submit_s, admit_s=e.admit_s, done_s=e.done_s))
            if k.value < 64:
                break
        return out

    def outstanding(self) -> int:
        return self._pending

    def now(self) -> float:
        return _lib.load().sadh_server_time(self._h)

    def result(self, rid: int) -> dict:
        r = self._reqs[rid]
        if not r["done"]:
            raise RuntimeError(f"request {rid} has not finished")
        times = dict(submit_s=r["submit_s"], admit_s=r["admit_s"], done_s=r["done_s"],
                     latency_s=r["done_s"] - r["submit_s"])
        if r["kind"] == "render":
            return dict(image=r["image"], **times)
        res = r["res"]
        out = _lib.fit_result_dict(res)
        out["sites"] = r["sites"]
        out["log_psnr"] = r["logs"][: min(res.n_logs, len(r["logs"]))].copy()
        out.update(times)
        return out

    def close(self) -> None:
        if self._h:
            _lib.call("sadh_server_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else None


def run_stream(requests: list[dict], opts: dict | None = None, memory_budget: float = 0.0, max_fit_jobs: int = 64,
               render_passes: int = 16) -> dict:
    """Submit each request at its arrival time `t` (seconds from the start)
    and step the server until all finish.  Returns throughput and latency
    percentiles per kind, and the per-request results."""
    o = dict(opts or fit_options())
    with Server(o, memory_budget, max_fit_jobs, render_passes) as srv:
        ids = []
        t0 = time.perf_counter()
        nxt = 0
        while nxt < len(requests) or srv.outstanding():
            now = time.perf_counter() - t0
            while nxt < len(requests) and requests[nxt]["t"] <= now:
                r = requests[nxt]
                if r["kind"] == "fit":
                    ids.append(srv.submit_fit(r["target"], r["sites"]))
                else:
                    ids.append(srv.submit_render(r["sites"], r["width"], r["height"]))
                nxt += 1
            if srv.outstanding():
                srv.step()
            elif nxt < len(requests):
                time.sleep(min(0.002, max(0.0, requests[nxt]["t"] - now)))
        wall = time.perf_counter() - t0
        results = [srv.result(i) for i in ids]
    fit_lat = [r["latency_s"] for r, q in zip(results, requests) if q["kind"] == "fit"]
    ren_lat = [r["latency_s"] for r, q in zip(results, requests) if q["kind"] == "render"]
    iters = o["iterations"]
    fit_mpi = sum(q["width"] * q["height"] for q in requests if q["kind"] == "fit") * iters / 1e6
    ren_mp = sum(q["width"] * q["height"] for q in requests if q["kind"] == "render") / 1e6
    return dict(wall_s=wall, fit_mp_iters_per_s=fit_mpi / wall, render_mp_per_s=ren_mp / wall,
                fit_latency_p50_s=_pct(fit_lat, 50), fit_latency_p99_s=_pct(fit_lat, 99),
                render_latency_p50_s=_pct(ren_lat, 50), render_latency_p99_s=_pct(ren_lat, 99),
                n_fit=len(fit_lat), n_render=len(ren_lat), results=results)
