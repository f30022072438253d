"""Plain device-to-device copy of a known size, used to scale HBM counters.

`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` report KiB per dispatch; running
them over this copy (N bytes read, N bytes written per dispatch) gives the
factor between what the counters read and the bytes a kernel really moved.

    python -m sad_amd.hbm_copy --mib 256 --reps 4
"""
import argparse
import sys


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--reps", type=int, default=4)
    args = ap.parse_args()
    import torch
    n = args.mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(f"copied {args.mib} MiB x {args.reps} (read {args.mib} MiB + write {args.mib} MiB per dispatch) ok={ok}",
          flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
