"""Sites and image files of the fit/render path.

Sites TXT (backends/cuda/src/sad_cuda.cu:277-338 read, :400-420 write):
  "#" lines are comments; a comment containing "image size" gives the width
  and height (the first two integers on the line); every other line holds 7
  (pos.x pos.y r g b log_tau radius) or 10 (+ dir.x dir.y log_aniso) numbers,
  other lines are skipped.  Written with 9 significant digits and the
  reference's three header lines.
Sites JSON (sad_cuda.cu:340-390): {"image_width", "image_height", "sites":
  [{"pos", "color", "log_tau", "radius" | "radius_sq", "log_aniso"?,
  "aniso_dir"?}]}.
Images: 8-bit RGB PNG in, value / 255 as float; out clamped to [0, 1] and
  rounded to 8 bits (sad_cuda.cu:137-246).  Masks: 8-bit grey / 255.
"""
from __future__ import annotations

import json
import os
import re

import numpy as np


_FLOAT = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")
_INT = re.compile(r"\s*([+-]?\d+)")


def _scan(text: str, rx, conv) -> list:
    """numbers read one after another until the first that does not parse,
    as repeated `istream >> value` does"""
    vals, pos = [], 0
    while True:
        m = rx.match(text, pos)
        if not m:
            return vals
        vals.append(conv(m.group(1)))
        pos = m.end()


def _image_size(line: str):
    """parseImageSize (sad_cuda.cu:255-275): every character but digits and '-'
    becomes a blank, then two integers are read"""
    nums = _scan(re.sub(r"[^0-9-]", " ", line), _INT, int)
    return (nums[0], nums[1]) if len(nums) >= 2 else None


def load_sites_txt(path: str):
    """-> (sites float32 [N][10], width or None, height or None)"""
    rows, hw, hh = [], 0, 0
    with open(path) as f:
        for line in f:
            t = line.strip(" \t\r\n")
            if not t:
                continue
            if t[0] == "#":
                if "image size" in t.lower():
                    sz = _image_size(t)
                    if sz:
                        hw, hh = sz
                continue
            vals = _scan(t, _FLOAT, float)
            if len(vals) not in (7, 10):
                continue
            s = [vals[0], vals[1], vals[5], vals[6], vals[2], vals[3], vals[4], 1.0, 0.0, 0.0]
            if len(vals) == 10:
                s[7], s[8], s[9] = vals[7], vals[8], vals[9]
            rows.append(s)
    if not rows:
        raise ValueError(f"no sites found in {path}")
    if hw > 0 and hh > 0:
        return np.asarray(rows, np.float32), hw, hh
    return np.asarray(rows, np.float32), None, None


def load_sites_json(path: str):
    with open(path) as f:
        j = json.load(f)
    rows = []
    for s in j["sites"]:
        r = s.get("radius", s.get("radius_sq", 0.0))
        d = s.get("aniso_dir", [1.0, 0.0])
        dx = d[0] if len(d) > 0 else 1.0
        dy = d[1] if len(d) > 1 else 0.0
        rows.append([s["pos"][0], s["pos"][1], s["log_tau"], r, s["color"][0], s["color"][1], s["color"][2], dx, dy,
                     s.get("log_aniso", 0.0)])
    if not rows:
        raise ValueError(f"no sites found in {path}")
    return np.asarray(rows, np.float32), int(j["image_width"]), int(j["image_height"])


def load_sites(path: str):
    """Sites from a .json or TXT file (loadSites, sad_cuda.cu:392-398)."""
    if os.path.splitext(path)[1].lower() == ".json":
        return load_sites_json(path)
    return load_sites_txt(path)


def active_sites(sites: np.ndarray) -> np.ndarray:
    """Rows with pos.x >= 0 (filterActiveSites)."""
    return sites[sites[:, 0] >= 0.0]


def write_sites_txt(sites: np.ndarray, width: int, height: int, path: str) -> None:
    """writeSitesTXT (sad_cuda.cu:400-420): columns pos.x pos.y r g b log_tau radius dir.x dir.y log_aniso."""
    s = np.asarray(sites, np.float32)
    cols = s[:, [0, 1, 4, 5, 6, 2, 3, 7, 8, 9]]
    with open(path, "w") as f:
        f.write("# SAD Sites (position_x, position_y, color_r, color_g, color_b, log_tau, radius, aniso_dir_x, "
                "aniso_dir_y, log_aniso)\n")
        f.write(f"# Image size: {width} {height}\n")
        f.write(f"# Total sites: {len(s)}\n")
        for row in cols:
            f.write(" ".join(f"{float(v):.9g}" for v in row) + "\n")


def _mitchell(x: np.ndarray) -> np.ndarray:
    """Mitchell-Netravali cubic (B = C = 1/3), support 2"""
    x = np.abs(x)
    return np.where(x < 1.0, (16.0 + x * x * (21.0 * x - 36.0)) / 18.0,
                    np.where(x < 2.0, (32.0 + x * (-60.0 + x * (36.0 - 7.0 * x))) / 18.0, 0.0))


def _catmull_rom(x: np.ndarray) -> np.ndarray:
    """Catmull-Rom cubic (B = 0, C = 1/2), support 2"""
    x = np.abs(x)
    return np.where(x < 1.0, 1.0 - x * x * (2.5 - 1.5 * x),
                    np.where(x < 2.0, 2.0 - x * (4.0 + x * (0.5 * x - 2.5)), 0.0))


def resample_matrix(n_in: int, n_out: int) -> np.ndarray:
    """[n_out][n_in] weights of one axis of the reference's resize (stb_image_resize2's easy API, which loadImage
    and loadMask call at sad_cuda.cu:157, 196): pixel centres at i + 0.5; a downscale (s = n_out / n_in < 1)
    weighs input pixel i for output j by the Mitchell filter at (j + 0.5) - (i + 0.5) s, in output pixels, an
    upscale by the Catmull-Rom filter at (i + 0.5) - (j + 0.5) / s, in input pixels; each output's weights are
    normalised to sum 1 and the taps beyond the image are moved onto its edge pixel (clamp)."""
    m = np.zeros((n_out, n_in), np.float64)
    s = n_out / n_in
    j = np.arange(n_out, dtype=np.float64)[:, None]
    if n_out < n_in:
        r = 2.0 / s
        i = np.arange(int(np.floor(-r)) - 1, n_in + int(np.ceil(r)) + 1)[None, :]
        w = _mitchell((j + 0.5) - (i + 0.5) * s) * s
    else:
        i = np.arange(-3, n_in + 3)[None, :]
        w = _catmull_rom((i + 0.5) - (j + 0.5) / s)
    w = w / w.sum(axis=1, keepdims=True)
    cols = np.clip(i[0], 0, n_in - 1)
    for k, c in enumerate(cols):
        m[:, c] += w[:, k]
    return m


def resize_plane(a: np.ndarray, out_w: int, out_h: int) -> np.ndarray:
    """[H][W](...) float64 planes resized separably with resample_matrix"""
    h, w = a.shape[:2]
    my, mx = resample_matrix(h, out_h), resample_matrix(w, out_w)
    return np.einsum("yh,hw...,xw->yx...", my, a, mx, optimize=True)


def srgb_to_linear(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, np.float64)
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)


def linear_to_srgb(v: np.ndarray) -> np.ndarray:
    v = np.clip(np.asarray(v, np.float64), 0.0, 1.0)
    return np.where(v <= 0.0031308, v * 12.92, 1.055 * v ** (1.0 / 2.4) - 0.055)


def resize_rgb8_srgb(img: np.ndarray, out_w: int, out_h: int) -> np.ndarray:
    """uint8 [H][W][3] -> uint8 [out_h][out_w][3], filtered in linear light as stbir_resize_uint8_srgb"""
    lin = resize_plane(srgb_to_linear(np.asarray(img, np.float64) / 255.0), out_w, out_h)
    return np.floor(linear_to_srgb(lin) * 255.0 + 0.5).astype(np.uint8)


def resize_gray8_linear(img: np.ndarray, out_w: int, out_h: int) -> np.ndarray:
    """uint8 [H][W] -> uint8 [out_h][out_w], filtered on the 8-bit values as stbir_resize_uint8_linear"""
    v = resize_plane(np.asarray(img, np.float64) / 255.0, out_w, out_h)
    return np.floor(np.clip(v, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def load_image(path: str, max_dim: int = 0) -> np.ndarray:
    """float32 [H][W][3] in [0, 1] (loadImage, sad_cuda.cu:137-180).  An image whose larger side exceeds
    max_dim is first scaled to round(side * max_dim / max(W, H)) pixels per side (resize_rgb8_srgb)."""
    from PIL import Image
    a = np.asarray(Image.open(path).convert("RGB"), np.uint8)
    h, w = a.shape[:2]
    if max_dim > 0 and max(w, h) > max_dim:
        sc = np.float32(max_dim) / np.float32(max(w, h))
        # std::round of the float32 product: halves away from zero
        ow = max(1, int(np.floor(np.float64(np.float32(w) * sc) + 0.5)))
        oh = max(1, int(np.floor(np.float64(np.float32(h) * sc) + 0.5)))
        a = resize_rgb8_srgb(a, ow, oh)
    return a.astype(np.float32) / np.float32(255.0)


def load_mask(path: str, width: int, height: int) -> np.ndarray:
    """float32 [H][W] = grey / 255 (loadMask, sad_cuda.cu:182-218), resized to the target size first when it
    differs (resize_gray8_linear)."""
    from PIL import Image
    a = np.asarray(Image.open(path).convert("L"), np.uint8)
    if a.shape != (height, width):
        a = resize_gray8_linear(a, width, height)
    return a.astype(np.float32) / np.float32(255.0)


def to_8bit(image: np.ndarray) -> np.ndarray:
    """clamp to [0, 1], then round(x * 255) with halves away from zero, the
    product taken in float32 (std::round(c * 255.0f))"""
    v = np.clip(np.asarray(image, np.float32), np.float32(0.0), np.float32(1.0)) * np.float32(255.0)
    return np.floor(v.astype(np.float64) + 0.5).astype(np.uint8)


def save_image(path: str, image: np.ndarray) -> None:
    """saveImage (sad_cuda.cu:228-246): 8-bit RGB PNG."""
    from PIL import Image
    Image.fromarray(to_8bit(image), mode="RGB").save(path)
