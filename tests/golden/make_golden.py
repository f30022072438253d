"""Write the golden fixtures of tests/golden/ with the CPU oracle.

    python tests/golden/make_golden.py

The reference commits no golden vectors (its tests draw seeded inputs and
compare with CPU oracles at run time), so these fixtures are written by this
repository's oracle, after the oracle has been pinned by tests/test_oracle.py.
They freeze its outputs against later refactors and give the GPU tests fixed
expected outputs.  Inputs are regenerated from the seeds stored with them.

Fixtures (one .npz each, compressed):
  cand_fit256.npz   256x256 / 2048-site plain-fit scene: the initial field
                    (init seed 0 + jfa seed) and the field after 16 VPT passes
                    (inject 4, Hilbert window 32 / probes 1), as SHA-256 digests
                    of the full planes plus the first 32 rows verbatim;
  cand_small.npz    64x48 / 300 sites: both planes after 10 VPT passes, whole;
  sort_pairs.npz    4096 densify score pairs, sorted (4 passes), and a 16-bit
                    key set sorted with the 2-pass rule;
  adam_step.npz     one Adam step (t = 7) on 256 sites with seeded state and
                    gradients: new sites and state;
  fit1024_full.npz  the full 1024x1024 fit (16,384 initial sites, densify and
                    prune every 100 iterations, 262,144-site buffer, 1.0 bpp,
                    2,000 iterations): SHA-256 of the final [262144, 10] sites,
                    their first 256 rows, the PSNR logged every 100 iterations
                    and the schedule counts.  Written only on request
                    (`make_golden.py fit1024`): the oracle needs a few minutes.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fit256_scene():
    from sad_amd import synth
    tgt = synth.make_target(256, 256, 1, 40)
    sites = synth.init_sites_uniform(tgt, 2048, 1)
    return tgt, sites


def vpt_passes(orc, c0, c1, sites, w, h, passes, hil=True):
    from tests.scenes import hilbert_tables, iss_of
    iss = iss_of(w, h)
    pk = orc.pack_sites(sites)
    ho, hp = hilbert_tables(orc, sites, w, h) if hil else (None, None)
    for i in range(passes):
        step = (orc.jump_step(i, w, h) << 16) | (i & 0xFFFF)
        c0, c1 = orc.update_candidates(c0, c1, pk, step, iss, w, h, inject_count=4, hilbert_order=ho,
                                       hilbert_pos=hp, hilbert_probes=1 if hil else 0,
                                       hilbert_window=32 if hil else 0)
    return c0, c1


def small_scene():
    from tests.scenes import random_sites
    return random_sites(64, 48, 300, 31)


def score_pairs_inputs():
    from tests.scenes import random_sites
    cap, n = 4096, 3500
    sites = np.zeros((cap, 10), np.float32)
    sites[:, 0:2] = -1.0
    sites[:n] = random_sites(256, 256, n, 32)
    rng = np.random.default_rng(33)
    mass = rng.exponential(3.0, cap).astype(np.float32)
    energy = rng.exponential(0.5, cap).astype(np.float32)
    keys16 = np.stack([rng.integers(0, 1 << 16, 3000), np.arange(3000)], 1).astype(np.uint32)
    return sites, mass, energy, keys16


def adam_inputs():
    from tests.scenes import random_sites
    rng = np.random.default_rng(34)
    n = 256
    sites = random_sites(200, 160, n, 35)
    state = np.abs(rng.normal(0, 1e-3, (n, 20))).astype(np.float32)
    state[:, 18:] = 0.0
    grads = rng.normal(0, 50, (10, n)).astype(np.float32)
    return sites, state, grads


ADAM_LRS = (0.025, 0.01, 0.01, 0.01, 0.0025, 0.025)


FIT1024_LOG_FREQ = 100


def fit1024_full(orc) -> dict:
    """The oracle's full 1024x1024 fit, as stored in fit1024_full.npz."""
    from sad_amd import config, synth
    tgt, sites, over = synth.fit_workload("fit1024")
    opts = config.fit_options(**dict(over, log_freq=FIT1024_LOG_FREQ))
    r = orc.fit(tgt, sites, opts)
    return dict(sites_sha=np.array(digest(r["sites"])), sites_head=r["sites"][:256],
                log_psnr=np.asarray(r["log_psnr"], np.float32), final_psnr=np.float32(r["final_psnr"]),
                counts=np.array([r["n_densify"], r["n_prune"], r["actual_sites"], r["active_estimate"],
                                 r["n_cand_updates"]], np.int64))


def main() -> None:
    from oracle import orc
    orc.lib()
    if sys.argv[1:] == ["fit1024"]:
        orc.set_threads(os.cpu_count() or 1)
        np.savez_compressed(os.path.join(HERE, "fit1024_full.npz"), **fit1024_full(orc))
        print("fit1024_full.npz", os.path.getsize(os.path.join(HERE, "fit1024_full.npz")), "bytes")
        return
    _, sites = fit256_scene()
    w = h = 256
    c0, c1 = orc.init_candidates(len(sites), 0, False, w, h)
    c0 = orc.jfa_seed(c0, sites, w, h)
    f0, f1 = vpt_passes(orc, c0, c1, sites, w, h, 16)
    np.savez_compressed(os.path.join(HERE, "cand_fit256.npz"),
                        init_sha=np.array([digest(c0), digest(c1)]), vpt16_sha=np.array([digest(f0), digest(f1)]),
                        init_rows=np.stack([c0[:32 * w], c1[:32 * w]]), vpt16_rows=np.stack([f0[:32 * w], f1[:32 * w]]))

    s = small_scene()
    c0, c1 = orc.init_candidates(len(s), 7, False, 64, 48)
    c0 = orc.jfa_seed(c0, s, 64, 48)
    f0, f1 = vpt_passes(orc, c0, c1, s, 64, 48, 10)
    np.savez_compressed(os.path.join(HERE, "cand_small.npz"), cand0=f0, cand1=f1)

    sites, mass, energy, keys16 = score_pairs_inputs()
    pairs = orc.densify_pairs(sites, mass, energy, 1.0, 0.7, len(sites))
    np.savez_compressed(os.path.join(HERE, "sort_pairs.npz"), pairs=pairs, sorted=orc.radix_sort(pairs),
                        sorted16=orc.radix_sort(keys16, 1 << 16))

    sites, state, grads = adam_inputs()
    ns, nst = orc.adam(sites, state, grads, ADAM_LRS, 0.9, 0.999, 1e-8, 7, 200, 160)
    np.savez_compressed(os.path.join(HERE, "adam_step.npz"), sites=ns, state=nst)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
