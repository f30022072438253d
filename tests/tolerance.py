"""Relative tolerance for per-site sums whose terms may cancel.

|gpu - ref| <= 1e-4 |ref| + 1e-5 sum|terms| + atol, where ref is the oracle's
double-precision sum and sum|terms| the sum of the absolute contributions.
A plain fp32 sequential sum of the same terms already deviates by up to
~7e-7 sum|terms|, and the GPU adds in a different order (float partials per
16x16 tile).  atol is the fixed-point resolution of the library's cross-tile
sums: P partials of a site are held to within P * 2^-41 (orc.FIX_HALF_STEP)."""
from __future__ import annotations

import numpy as np


class Rel(float):
    """max of |got - ref| / (1e-4 |ref| + 1e-5 sum|terms|) (<= 1 passes), with
    the worst element's values for the failure message."""

    def __new__(cls, got, ref, ab, atol=0.0):
        err = np.abs(np.asarray(got, np.float64) - ref) / (1e-4 * np.abs(ref) + 1e-5 * ab + atol + 1e-30)
        i = np.unravel_index(int(np.argmax(err)), err.shape)
        obj = super().__new__(cls, float(err[i]))
        obj.info = (f"worst at {tuple(int(v) for v in i)}: gpu {float(np.asarray(got)[i])!r} "
                    f"ref {float(ref[i])!r} sum|terms| {float(ab[i])!r}")
        return obj

    def __repr__(self):
        return f"{float(self):.4g} ({self.info})"
