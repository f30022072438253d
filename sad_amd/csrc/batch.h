/*
 * batch.h - device descriptor of one image ("job") of a batched fit.
 *
 * The batched fit advances many images in lock step.  Each per-iteration
 * kernel (VPT search, tiled gradients, tau diffusion, tau + Adam) is one launch
 * over all images: its grid is the concatenation of every job's 16x16 tiles or
 * sites, and a block finds its job by binary search over the jobs' first tile
 * (or site) index.  The descriptors live in device memory; the host changes
 * only `n` (after densify) and `cur` (after a candidate update) with a small
 * kernel, in stream order.
 */
#ifndef SAD_BATCH_H
#define SAD_BATCH_H

#include "sad_device.h"

namespace sadh {

struct BatchJob {
    /* image */
    const float* target;
    const float* mask;
    int width, height;         /* pixels */
    int cw, ch, ds;            /* candidate grid and downscale */
    float iss, inv_scale;      /* 1/max(W,H)^2 and its square root */
    float norm, maxx, maxy;    /* Adam: 1/(W*H), W-1, H-1 */
    /* candidate fields: two plane pairs, ping-ponged by the VPT pass */
    uint32_t* fa0;
    uint32_t* fa1;
    uint32_t* fb0;
    uint32_t* fb1;
    /* sites and their per-site buffers (capacity entries each) */
    SadSite* sites;
    SadAdamState* adam;
    SadSearchSite* search;
    long long* gacc;           /* gradients: fixed-point lines, FIX_GRAD_STRIDE words per site */
    float* tau_raw;
    float* tau_tmp;
    float* tau_tmp2;
    uint32_t* tau_nbr;         /* 8 ids per site: neighbour lists shared by the tau passes */
    long long* removal;        /* prune iterations: removal deltas, one fixed-point (lo, hi) pair per site */
    const float* rendered;     /* SSIM loss: render of the current state, written before the gradient */
    float* render;             /* the same render buffer, written by the batched log renders */
    float ssim_weight;         /* SSIM loss weight (0: L2 only) */
    const uint32_t* h_order;   /* null: no Hilbert probes */
    const uint32_t* h_pos;
    uint32_t n;                /* current site count */
    uint32_t cur;              /* current candidate field: 0 = (fa0, fa1), 1 = (fb0, fb1) */
    uint32_t cap;              /* site capacity (flattened site range) */
    /* offsets of this job in the flattened grids */
    uint32_t tiles_x, ntiles, tile_begin;     /* 16x16 pixel tiles (gradient) */
    uint32_t ctiles_x, cntiles, ctile_begin;  /* 16x16 candidate-cell tiles (VPT) */
    uint32_t site_begin;
    uint32_t pblock_begin;     /* 256-pixel row-major blocks (log PSNR sums) */
    uint32_t sblock_begin;     /* 256-site blocks over the capacity (log active-site counts) */
};

/* Job j of the table with its pointers rebased on the table's own (kernel
 * argument) address: the same addresses, accessed with global rather than
 * generic memory instructions (rebase, sad_device.h). */
__device__ __forceinline__ BatchJob job_global(const BatchJob* __restrict__ jobs, uint32_t j)
{
    BatchJob J = jobs[j];
    J.target = rebase(jobs, J.target);
    J.mask = rebase(jobs, J.mask);
    J.fa0 = rebase(jobs, J.fa0);
    J.fa1 = rebase(jobs, J.fa1);
    J.fb0 = rebase(jobs, J.fb0);
    J.fb1 = rebase(jobs, J.fb1);
    J.sites = rebase(jobs, J.sites);
    J.adam = rebase(jobs, J.adam);
    J.search = rebase(jobs, J.search);
    J.gacc = rebase(jobs, J.gacc);
    J.tau_raw = rebase(jobs, J.tau_raw);
    J.tau_tmp = rebase(jobs, J.tau_tmp);
    J.tau_tmp2 = rebase(jobs, J.tau_tmp2);
    J.tau_nbr = rebase(jobs, J.tau_nbr);
    J.removal = rebase(jobs, J.removal);
    J.rendered = rebase(jobs, J.rendered);
    J.render = rebase(jobs, J.render);
    J.h_order = rebase(jobs, J.h_order);
    J.h_pos = rebase(jobs, J.h_pos);
    return J;
}

/* job owning flattened index f, given each job's first index `begin` (ascending) */
__device__ __forceinline__ uint32_t batch_find(const BatchJob* __restrict__ jobs, uint32_t njobs, uint32_t f,
                                               uint32_t BatchJob::*begin)
{
    uint32_t lo = 0u, hi = njobs;   /* jobs[lo].begin <= f < jobs[hi].begin */
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].*begin <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* per-job scalars of one engine tick (jobs may be at different iterations) */
struct JobStep {
    float lambda;              /* tau diffusion weight at the job's iteration */
    float b1c, b2c;            /* Adam bias corrections 1/(1-beta^t), t = iteration + 1 */
    uint32_t jump_index;       /* the job's VPT pass index */
    uint32_t vpt;              /* 1: the job updates its candidates this tick */
    uint32_t log;              /* 1: the job logs this tick (batched render, PSNR sum and active count) */
    /* batched Hilbert tables: the job's sites occupy [hil_begin, hil_begin + hil_n) of the tick's pair array
     * (hil_n = 0: no table this tick); hil_bits is the job's Hilbert grid (2^bits per side) */
    uint32_t hil_begin, hil_n, hil_bits;
    uint32_t retire;           /* 1: the job's last iteration (its sites are gathered for the copy back) */
};

/* the job whose Hilbert range [hil_begin, hil_begin + hil_n) holds i < the ranges' total; the ranges are
 * consecutive in job order, so it is the last job whose range starts at or before i */
__device__ __forceinline__ uint32_t hil_find(const JobStep* __restrict__ steps, uint32_t njobs, uint32_t i)
{
    uint32_t lo = 0u, hi = njobs;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (steps[mid].hil_begin <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* scalars shared by every job of a batched launch */
struct BatchStep {
    float lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso;
    float beta1, beta2, eps;
    uint32_t vpt_pass;         /* pass p of this tick's update: jump index + p, reads field cur ^ (p & 1) */
    float radius_scale;
    uint32_t radius_probes, inject_count, hil_probes, hil_window;
};

}  // namespace sadh

#endif /* SAD_BATCH_H */
