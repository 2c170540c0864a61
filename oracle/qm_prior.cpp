// qm_prior.cpp -- CPU oracle (test infrastructure only; nothing here is
// shipped or measured) for the reference's prior-driven alternates:
//   quadtree_depth_kernal  src/quadtree.cu:79-154
//   prior_to_cost          src/depth_extract.cu:23-47 (with image_to_cost's and
//                          normalize_the_cost's add_num, :223-244)
// Single-threaded loops in the reference's own terms (tiles, blocks, pixels);
// the sums the reference forms with atomics in no fixed order are formed here
// in one fixed order: a block's children ((top-left + top-right) + bottom-left)
// + bottom-right, a cell's prior pixels in raster order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// levels (H x W) in place: max(level, depth level of the pixel)
void qmo_quadtree_prior(const float* prior_depth, int W, int H, uint8_t* levels) {
  for (int ty = 0; ty < H; ty += 16)
    for (int tx = 0; tx < W; tx += 16) {
      // a tile takes part only when every pixel of it inside the image has depth > 0
      bool all_valid = true;
      float inv[16][16];
      bool inside[16][16];
      for (int ly = 0; ly < 16; ++ly)
        for (int lx = 0; lx < 16; ++lx) {
          const int x = tx + lx, y = ty + ly;
          inside[ly][lx] = x < W && y < H;
          inv[ly][lx] = 0.0f;
          if (!inside[ly][lx]) continue;
          const float d = prior_depth[(size_t)y * W + x];
          if (d > 0.0f)
            inv[ly][lx] = 1.0f / d;
          else
            all_valid = false;
        }
      if (!all_valid) continue;
      int level[16][16] = {};
      // blocks of the previous level: sums and approval, indexed by block
      float sum[16][16];
      bool approved[16][16];
      for (int ly = 0; ly < 16; ++ly)
        for (int lx = 0; lx < 16; ++lx) {
          sum[ly][lx] = inv[ly][lx];
          approved[ly][lx] = inside[ly][lx];
        }
      for (int i = 1; i <= 4; ++i) {
        const int nb = 16 >> i;
        float nsum[16][16];
        bool nappr[16][16];
        for (int by = 0; by < nb; ++by)
          for (int bx = 0; bx < nb; ++bx) {
            const int cy = 2 * by, cx = 2 * bx;  // children at the previous level
            const bool complete = approved[cy][cx] && approved[cy][cx + 1] && approved[cy + 1][cx] &&
                                  approved[cy + 1][cx + 1];
            const float s = ((sum[cy][cx] + sum[cy][cx + 1]) + sum[cy + 1][cx]) + sum[cy + 1][cx + 1];
            nsum[by][bx] = s;
            bool ok = complete;
            if (complete) {
              const int sz = 1 << i;
              const float avg = s / (float)(sz * sz);
              for (int py = 0; py < sz && ok; ++py)
                for (int px = 0; px < sz; ++px)
                  if (std::fabs(inv[by * sz + py][bx * sz + px] - avg) > 0.01f) {
                    ok = false;
                    break;
                  }
              if (ok)
                for (int py = 0; py < sz; ++py)
                  for (int px = 0; px < sz; ++px) level[by * sz + py][bx * sz + px] = i;
            }
            nappr[by][bx] = ok;
          }
        for (int by = 0; by < nb; ++by)
          for (int bx = 0; bx < nb; ++bx) {
            sum[by][bx] = nsum[by][bx];
            approved[by][bx] = nappr[by][bx];
          }
      }
      for (int ly = 0; ly < 16; ++ly)
        for (int lx = 0; lx < 16; ++lx) {
          if (!inside[ly][lx]) continue;
          uint8_t& l = levels[(size_t)(ty + ly) * W + tx + lx];
          if (level[ly][lx] > l) l = (uint8_t)level[ly][lx];
        }
    }
}

// cost (Hc x Wc x N, sample order) and count (nullable) from the cost-volume
// oracle, adjusted in place; prior is H x W x 2 (inverse depth, variance)
void qmo_prior_to_cost(const float* prior, int W, int H, const uint8_t* levels, int N, double imin, double imax,
                       float* cost, int32_t* count) {
  const int Wc = W / 4, Hc = H / 4;
  const double step = (imax - imin) / (double)(N - 1);
  std::vector<float> acc(N);
  for (int cy = 0; cy < Hc; ++cy)
    for (int cx = 0; cx < Wc; ++cx) {
      std::fill(acc.begin(), acc.end(), 0.0f);
      int P = 0;
      for (int py = 0; py < 4; ++py)
        for (int px = 0; px < 4; ++px) {
          const int x = 4 * cx + px, y = 4 * cy + py;
          if (x >= W - 3 || y >= H - 3 || x <= 2 || y <= 2) continue;
          const float p = prior[2 * ((size_t)y * W + x)], v = prior[2 * ((size_t)y * W + x) + 1];
          if (p <= 0 || v <= 0) continue;
          ++P;
          for (int d = 0; d < N; ++d) {
            const float my_invdepth = (float)(step * d + imin);
            float c = (float)(0.1 * (my_invdepth - p) * (my_invdepth - p) / v);
            c = c < 0.5 ? c : 0.5;
            acc[d] += c;
          }
        }
      if (P == 0) continue;
      const int x0 = 4 * cx, y0 = 4 * cy;
      const int qs = 1 << levels[(size_t)y0 * W + x0];
      const int node = !(x0 >= W - 1 || y0 >= H - 1 || x0 <= 0 || y0 <= 0 || x0 % qs != 0 || y0 % qs != 0);
      const int add_num = 64 * (P + node);
      float* c = cost + ((size_t)cy * Wc + cx) * N;
      for (int d = 0; d < N; ++d) c[d] = (acc[d] + c[d] * 64.0f) / (float)add_num;
      if (count) count[(size_t)cy * Wc + cx] = add_num;
    }
}

}  // extern "C"
