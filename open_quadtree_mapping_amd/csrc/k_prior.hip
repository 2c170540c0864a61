// k_prior.hip -- the reference's prior-driven alternates, which use a
// semi-dense prior of the keyframe in the quadtree and in the cost volume:
//   quadtree_depth_kernal (src/quadtree.cu:79-154): raises a pixel's quadtree
//     level where the prior inverse depth is flat over a 2^i block;
//   prior_to_cost (src/depth_extract.cu:23-47, its launch commented out at
//     src/seed_matrix.cu:492-499): adds a truncated quadratic prior cost of
//     every prior pixel of a node's 4 x 4 block to the node's costs.
// Neither is on the reference's live path; both run here when a frame carries
// the prior (FrameDev::prior_depth, FrameDev::prior).
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

// ---------------------------------------------------------------------------
// Depth quadtree. One block of 256 threads per 16 x 16 tile, one pixel each.
// A tile takes part only when every in-image pixel has a prior depth > 0. A
// 2^i block is complete when its four 2^(i-1) children were approved, and
// approved when every pixel's inverse depth lies within 0.01 of the block
// mean; a pixel's depth level is the largest i whose block it approved, and
// its new quadtree level is max(image level, depth level). Blocks that reach
// past the image edge are never complete (the reference's out-of-image
// threads never add their counts). The reference sums a block with shared
// atomics in no fixed order; here the children are summed as
// ((top-left + top-right) + bottom-left) + bottom-right.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_quadtree_prior(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const int tiles_x = (W + 15) >> 4;
  const int t = blockIdx.x - bm.start[f];
  const int ty = t / tiles_x, tx = t - ty * tiles_x;
  const int tid = threadIdx.x;
  const int lx = tid & 15, ly = tid >> 4;
  const int x = tx * 16 + lx, y = ty * 16 + ly;
  const bool in = x < W && y < H;

  __shared__ float sum[256];
  __shared__ int appr[256];  // approved at the previous level (rep positions)
  __shared__ int comp[256];  // complete at the current level
  __shared__ int bad[256];   // some pixel of the current block disapproved

  float inv = 0.0f;
  bool valid = false;
  if (in) {
    const float d = F.prior_depth[(size_t)y * W + x];
    valid = d > 0.0f;
    inv = valid ? 1.0f / d : 0.0f;
  }
  sum[tid] = inv;
  appr[tid] = in ? 1 : 0;
  const int all_valid = __syncthreads_and(!in || valid);
  if (!all_valid) return;  // whole block

  int plevel = 0;
#pragma unroll
  for (int i = 1; i <= 4; ++i) {
    const int sz = 1 << i, h = sz >> 1;
    const int r = (ly - (ly & (sz - 1))) * 16 + (lx - (lx & (sz - 1)));
    if (r == tid) {
      const int c1 = r + h, c2 = r + 16 * h, c3 = r + 16 * h + h;
      comp[r] = appr[r] & appr[c1] & appr[c2] & appr[c3];
      sum[r] = ((sum[r] + sum[c1]) + sum[c2]) + sum[c3];
      bad[r] = 0;
    }
    __syncthreads();
    const int complete = comp[r];
    if (complete) {
      const float avg = sum[r] / (float)(sz * sz);
      if (fabsf(inv - avg) > 0.01f) bad[r] = 1;
    }
    __syncthreads();
    const int ok = complete && !bad[r];
    if (ok) plevel = i;
    if (r == tid) appr[r] = ok;
    __syncthreads();
  }
  if (in) {
    uint8_t* lv = F.levels + (size_t)y * W + x;
    const int c = *lv;
    *lv = (uint8_t)(c > plevel ? c : plevel);
  }
}

void launch_quadtree_prior(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i)
    counts[i] = hframes[i].prior_depth ? ((hframes[i].W + 15) / 16) * ((hframes[i].H + 15) / 16) : 0;
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_quadtree_prior, dim3(nb), dim3(256), 0, s, dframes, bm);
}

// ---------------------------------------------------------------------------
// Prior cost. One wave per coarse cell, lanes over samples. The reference
// atomically adds, per sample, min(0.1 (s_d - p)^2 / v, 0.5) of every prior
// pixel (p, v) of the cell's 4 x 4 block with 3 <= x <= W-4, 3 <= y <= H-4
// and neither p <= 0 nor v <= 0 (a NaN prior passes and adds 0.5), and 1 to
// add_num per sample thread; image_to_cost then adds the node's image cost
// (64 more to add_num) and normalize_the_cost divides by add_num. The cost
// volume already holds the image cost divided by its own add_num of 64, so
// with P prior pixels and I = 1 on a node (0 elsewhere) the cell becomes
//   (sum of prior costs + 64 * cost) / (64 * (P + I)),
// the prior costs summed in raster order (the reference's atomics have no
// fixed order), 64 * cost being the node's raw image cost exactly.
// ---------------------------------------------------------------------------
template <int N, bool ROW>
__global__ __launch_bounds__(256) void k_prior_cost(const FrameDev* __restrict__ frames, BlockMap bm, double imin,
                                                    double istep) {
  constexpr int K = N / 64;
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H, Wc = F.Wc;
  const int cell = (blockIdx.x - bm.start[f]) * 4 + (threadIdx.x >> 6);
  if (cell >= Wc * F.Hc) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int cy = cell / Wc, cx = cell - cy * Wc;
  float sidepth[K], acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    sidepth[k] = (float)(istep * (double)(lane + 64 * k) + imin);
    acc[k] = 0.0f;
  }
  const float2* prior = F.prior;
  int P = 0;
  for (int py = 0; py < 4; ++py) {
    const int y = 4 * cy + py;
    if (y <= 2 || y >= H - 3) continue;
    for (int px = 0; px < 4; ++px) {
      const int x = 4 * cx + px;
      if (x <= 2 || x >= W - 3) continue;
      const float2 pr = prior[(size_t)y * W + x];
      if (pr.x <= 0.0f || pr.y <= 0.0f) continue;
      ++P;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float diff = sidepth[k] - pr.x;
        double c = 0.1 * (double)diff;
        c = c * (double)diff;
        c = c / (double)pr.y;
        const float cf = (float)c;
        acc[k] += cf < 0.5f ? cf : 0.5f;
      }
    }
  }
  if (P == 0) return;
  const int x0 = 4 * cx, y0 = 4 * cy;
  const int qs = 1 << F.levels[(size_t)y0 * W + x0];
  const int node = !(x0 >= W - 1 || y0 >= H - 1 || x0 <= 0 || y0 <= 0 || (x0 % qs) != 0 || (y0 % qs) != 0);
  const float denom = (float)(64 * (P + node));
  float* c = F.cost[0] + (size_t)cell * N;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int d = lane + 64 * k;
    float* slot = c + (ROW ? row_pos<N>(d) : d);
    *slot = (acc[k] + *slot * 64.0f) / denom;
  }
  if (F.count && lane == 0) F.count[cell] = 64 * (P + node);
}

void launch_prior_cost(const FrameDev* dframes, const FrameDev* hframes, int n, int N, bool natural,
                       double idepth_min, double idepth_step, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = hframes[i].prior ? (hframes[i].Wc * hframes[i].Hc + 3) / 4 : 0;
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 128) {
    if (natural)
      hipLaunchKernelGGL((k_prior_cost<128, false>), dim3(nb), dim3(256), 0, s, dframes, bm, idepth_min, idepth_step);
    else
      hipLaunchKernelGGL((k_prior_cost<128, true>), dim3(nb), dim3(256), 0, s, dframes, bm, idepth_min, idepth_step);
  } else {
    if (natural)
      hipLaunchKernelGGL((k_prior_cost<64, false>), dim3(nb), dim3(256), 0, s, dframes, bm, idepth_min, idepth_step);
    else
      hipLaunchKernelGGL((k_prior_cost<64, true>), dim3(nb), dim3(256), 0, s, dframes, bm, idepth_min, idepth_step);
  }
}

}  // namespace qm
