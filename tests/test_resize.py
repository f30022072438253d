"""Image and mask resizing of the fit's inputs against the reference's own resize.

loadImage scales an image whose larger side exceeds MAX_DIM with
stbir_resize_uint8_srgb, and loadMask fits a mask to the image with
stbir_resize_uint8_linear (sad_cuda.cu:150-160, 193-200).  The reference's
stb_image_resize2 is compiled from the reference tree into oracle/_ref
(`make -C oracle ref`); these tests compare sad_amd.sites_io's resize with it.

Bars: every value within 1 of stb's, and a differing value only where the
exact (float64) value lies close to a rounding boundary: within 0.001 for
masks (8-bit values filtered directly, up- and downscaling: the spread of
stb's float32 sums), within 0.05 for images (filtered in linear light: also
stb's table-driven sRGB encoder)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from sad_amd import sites_io, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libstbref.so")


@pytest.fixture(scope="module")
def stb():
    if not os.path.exists(REF_LIB):
        # builds it where the reference tree is present, does nothing elsewhere
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    if not os.path.exists(REF_LIB):
        pytest.skip("the reference's stb resize is not built (no reference tree here)")
    return C.CDLL(REF_LIB)


def _ref_rgb(lib, a, ow, oh):
    a = np.ascontiguousarray(a, np.uint8)
    out = np.zeros((oh, ow, 3), np.uint8)
    assert lib.ref_resize_rgb8_srgb(a.ctypes.data_as(C.c_void_p), a.shape[1], a.shape[0],
                                    out.ctypes.data_as(C.c_void_p), ow, oh) == 1
    return out


def _ref_gray(lib, a, ow, oh):
    a = np.ascontiguousarray(a, np.uint8)
    out = np.zeros((oh, ow), np.uint8)
    assert lib.ref_resize_gray8_linear(a.ctypes.data_as(C.c_void_p), a.shape[1], a.shape[0],
                                       out.ctypes.data_as(C.c_void_p), ow, oh) == 1
    return out


@pytest.mark.parametrize("w,h,ow,oh", [(30, 20, 60, 40), (64, 48, 32, 24), (50, 50, 73, 29), (17, 9, 100, 63),
                                       (200, 3, 41, 7)])
def test_mask_resize_matches_stb(stb, w, h, ow, oh):
    g = np.random.default_rng(w * 1000 + h).integers(0, 256, (h, w)).astype(np.uint8)
    d = np.abs(sites_io.resize_gray8_linear(g, ow, oh).astype(int) - _ref_gray(stb, g, ow, oh).astype(int))
    assert d.max() <= 1
    exact = np.clip(sites_io.resize_plane(g.astype(np.float64) / 255.0, ow, oh), 0.0, 1.0) * 255.0
    off = np.abs(exact - np.floor(exact) - 0.5)
    assert np.all(off[d > 0] < 0.001), off[d > 0].max()
    assert (d == 0).mean() > 0.999


@pytest.mark.parametrize("w,h,ow,oh", [(257, 190, 100, 74), (300, 300, 97, 97), (1000, 37, 120, 4),
                                       (123, 77, 61, 38)])
def test_image_downscale_matches_stb(stb, w, h, ow, oh):
    rng = np.random.default_rng(w + h)
    imgs = [sites_io.to_8bit(synth.make_target(w, h, 7, n_ellipses=30)),
            rng.integers(0, 256, (h, w, 3)).astype(np.uint8)]
    for a in imgs:
        ref = _ref_rgb(stb, a, ow, oh).astype(int)
        got = sites_io.resize_rgb8_srgb(a, ow, oh).astype(int)
        d = np.abs(ref - got)
        assert d.max() <= 1
        exact = sites_io.linear_to_srgb(sites_io.resize_plane(
            sites_io.srgb_to_linear(a.astype(np.float64) / 255.0), ow, oh)) * 255.0
        off = np.abs(exact - np.floor(exact) - 0.5)
        assert np.all(off[d > 0] < 0.05), off[d > 0].max()
        assert (d == 0).mean() > 0.97


def test_load_image_and_mask_scale_like_the_reference(stb, tmp_path):
    """load_image(path, max_dim) picks the reference's output size (std::round of the float32 product) and
    its pixels; load_mask fits a mask to the image size."""
    from PIL import Image
    a = sites_io.to_8bit(synth.make_target(301, 199, 3, n_ellipses=20))
    Image.fromarray(a, mode="RGB").save(str(tmp_path / "t.png"))
    img = sites_io.load_image(str(tmp_path / "t.png"), 120)
    ow, oh = 120, int(np.floor(np.float64(np.float32(199) * (np.float32(120) / np.float32(301))) + 0.5))
    assert img.shape == (oh, ow, 3)
    ref = _ref_rgb(stb, a, ow, oh).astype(np.float32) / np.float32(255.0)
    assert np.abs(img - ref).max() <= np.float32(1.0 / 255.0) + 1e-7
    g = np.random.default_rng(5).integers(0, 256, (50, 70)).astype(np.uint8)
    Image.fromarray(g, mode="L").save(str(tmp_path / "m.png"))
    m = sites_io.load_mask(str(tmp_path / "m.png"), ow, oh)
    ref = _ref_gray(stb, g, ow, oh).astype(np.float32) / np.float32(255.0)
    assert m.shape == (oh, ow) and np.abs(m - ref).max() <= np.float32(1.0 / 255.0) + 1e-7
