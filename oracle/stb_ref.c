/* Test infrastructure only (never linked into the product): the reference's image
 * resize, stb_image_resize2 as vendored in the reference tree at
 * backends/cuda/third_party/stb_image_resize.h, compiled from there into
 * oracle/_ref/libstbref.so by `make -C oracle ref`.  loadImage and loadMask
 * (sad_cuda.cu:150-160, 193-200) call these two easy-API functions; the
 * tests compare sad_amd.sites_io's resize with them. */
#define STB_IMAGE_RESIZE_IMPLEMENTATION
#include "stb_image_resize.h"

/* stbir_resize_uint8_srgb(..., STBIR_RGB): packed RGB8 rows; 1 on success */
int ref_resize_rgb8_srgb(const unsigned char* in, int w, int h, unsigned char* out, int ow, int oh)
{
    return stbir_resize_uint8_srgb(in, w, h, 0, out, ow, oh, 0, STBIR_RGB) != 0;
}

/* stbir_resize_uint8_linear(..., STBIR_1CHANNEL): one 8-bit channel; 1 on success */
int ref_resize_gray8_linear(const unsigned char* in, int w, int h, unsigned char* out, int ow, int oh)
{
    return stbir_resize_uint8_linear(in, w, h, 0, out, ow, oh, 0, STBIR_1CHANNEL) != 0;
}
