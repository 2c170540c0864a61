// k_input.hip -- input conditioning on the GPU: undistort/rectify remap of the
// uint8 camera image, replacing the CPU cv::remap(INTER_LINEAR,
// BORDER_CONSTANT) of SeedMatrix::input_raw (src/seed_matrix.cu:81-90).
//
// Bilinear interpolation in OpenCV's documented fixed-point form for 8-bit
// images: map coordinates are rounded to 1/32 pixel (round half to even), the
// four weights are (32-a)(32-b)*32, a(32-b)*32, (32-a)b*32, ab*32 (they sum to
// 2^15 exactly), the result is (sum + 2^14) >> 15; taps outside the image read
// the constant border value 0.
#include "qm_device.h"

namespace qm {

__global__ __launch_bounds__(256) void k_remap_u8(const uint8_t* __restrict__ src, int SW, int SH, int spitch,
                                                  int W, int H, const float* __restrict__ mapx,
                                                  const float* __restrict__ mapy, uint8_t* __restrict__ dst) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const float mx = mapx[i], my = mapy[i];
  const float sx = mx * 32.0f, sy = my * 32.0f;
  // saturate to int like OpenCV's saturate_cast<int>(float)
  int ix, iy;
  if (!(sx == sx)) ix = 0;
  else if (sx >= 2147483520.0f) ix = 2147483647;
  else if (sx <= -2147483648.0f) ix = (-2147483647 - 1);
  else ix = (int)rintf(sx);
  if (!(sy == sy)) iy = 0;
  else if (sy >= 2147483520.0f) iy = 2147483647;
  else if (sy <= -2147483648.0f) iy = (-2147483647 - 1);
  else iy = (int)rintf(sy);
  const int x0 = ix >> 5, y0 = iy >> 5;
  const int a = ix & 31, b = iy & 31;
  auto px = [&](int x, int y) -> int {
    return (x >= 0 && x < SW && y >= 0 && y < SH) ? (int)src[(size_t)y * spitch + x] : 0;
  };
  const int v = (32 - a) * (32 - b) * 32 * px(x0, y0) + a * (32 - b) * 32 * px(x0 + 1, y0) +
                (32 - a) * b * 32 * px(x0, y0 + 1) + a * b * 32 * px(x0 + 1, y0 + 1);
  int r = (v + (1 << 14)) >> 15;
  r = r < 0 ? 0 : (r > 255 ? 255 : r);
  dst[i] = (uint8_t)r;
}

void launch_remap_u8(const uint8_t* src, int SW, int SH, int spitch, int W, int H, const float* mapx,
                     const float* mapy, uint8_t* dst, hipStream_t s) {
  const int nb = (W * H + 255) / 256;
  hipLaunchKernelGGL(k_remap_u8, dim3(nb), dim3(256), 0, s, src, SW, SH, spitch, W, H, mapx, mapy, dst);
}

}  // namespace qm
