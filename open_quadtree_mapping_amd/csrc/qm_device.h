// qm_device.h -- frame descriptors, work tables and pinned device arithmetic
// shared by the gfx950 kernels of the per-keyframe depth pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qm {

constexpr int kMaxFrames = 128;
constexpr int kWave = 64;

// Everything one keyframe of one stream needs on the device. Lives in device
// memory (an array over the batch); the kernels read it through the scalar
// cache. Level l of the BP pyramid has lw[l] x lh[l] cells.
struct FrameDev {
  int W, H, pitch, Wc, Hc;
  int lw[4], lh[4];
  float fx, fy, cx, cy;
  int row_seg, col_seg;
  int edge_r, edge_c;  // weight of the x=W / y=H edge (1 when W / H is not a multiple of 16)
  const uint8_t* key;
  const uint8_t* ref[4];
  float T[4][12];       // key -> reference r
  float T_fuse[12];     // last fused frame -> this frame
  const float* prefill;  // H x W or null
  const float* sparse_in;  // explicit sparse map (stage API) or null
  uint8_t* levels;      // H x W
  float* cost[4];       // lh*lw*N per level
  float* msg[4];        // lh*lw*4*N per level, [cell][dir][N], dir = up, down, left, right
  float* dcoarse;       // Hc x Wc
  int32_t* label;       // Hc x Wc
  int32_t* count;       // Hc x Wc or null
  float* sparse_out;    // H x W or null
  float4* fwd;          // H x W Thomas forward coefficients
  float2* rowxn;        // H x W row pass result
  float* dense;         // H x W
  float* dense_out;     // H x W copy for the caller or null
  float4* seeds;        // H x W (warped in place)
  float4* seeds_new;    // H x W
  unsigned long long* zbuf;  // H x W
  float* depth_out;     // H x W or null
  int32_t* table_out;   // H x W or null
  // generalised upsample inputs (differentiable path); null in the pipeline
  const float* g_d;
  const float* g_conf;
  const float* g_wrow;
  const float* g_wcol;
  float* row_x_out;
  float* row_n_out;
  // adjoint of the upsample (recomputation buffers and gradients)
  float2* colxn;        // H x W column pass (x, n) or null
  float2* adjxn;        // H x W column-adjoint solution
  const float* g_out;   // dL/d(dense)
  float* grad_d;
  float* grad_conf;
  float* grad_wrow;     // H x (W+1)
  float* grad_wcol;     // (H+1) x W
  // adjoint of the fusion update
  const int32_t* table_in;     // hole-filled warp table
  const float* g_depth_out;    // dL/d(depth_out)
  const float4* g_seeds_out;   // dL/d(seeds_out) or null
  float* grad_depth_in;
};

// blockIdx.x -> frame: frame f owns blocks [start[f], start[f+1]).
struct BlockMap {
  int n;
  int start[kMaxFrames + 1];
};

__device__ __forceinline__ int find_frame(const BlockMap& bm, int block) {
  int lo = 0, hi = bm.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bm.start[mid] <= block) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Dispatch places workgroup b on XCD b % 8. Renumber so that each XCD gets one
// contiguous range of logical blocks (neighbouring cells share its L2).
__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
  const int xcd = b & 7, k = b >> 3;
  const int base = nb >> 3, rem = nb & 7;
  return xcd * base + (xcd < rem ? xcd : rem) + k;
}

// float -> int as the reference's (int)(x + 0.5) with a double literal:
// truncation toward zero, NaN -> 0, saturating.
__device__ __forceinline__ int trunc_i32(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (-2147483647 - 1);
  return (int)v;
}

__device__ __forceinline__ float pinned_expf(float x) { return (float)exp((double)x); }

struct V3 {
  float x, y, z;
};

// normalize(cam2world(x, y)) with the pinned order ((x*x + y*y) + z*z)
__device__ __forceinline__ V3 ray_dir(int x, int y, float fx, float fy, float cx, float cy) {
  V3 v;
  v.x = ((float)x - cx) / fx;
  v.y = ((float)y - cy) / fy;
  v.z = 1.0f;
  const float d = (v.x * v.x + v.y * v.y) + v.z * v.z;
  const float inv = 1.0f / sqrtf(d);
  v.x = v.x * inv;
  v.y = v.y * inv;
  v.z = v.z * inv;
  return v;
}

__device__ __forceinline__ V3 se3_apply(const float* T, const V3& p) {
  V3 q;
  q.x = ((T[0] * p.x + T[1] * p.y) + T[2] * p.z) + T[3];
  q.y = ((T[4] * p.x + T[5] * p.y) + T[6] * p.z) + T[7];
  q.z = ((T[8] * p.x + T[9] * p.y) + T[10] * p.z) + T[11];
  return q;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const float o = __shfl_xor(v, m, kWave);
    v = o < v ? o : v;
  }
  return v;
}

// Pairwise tree sum with offsets 32, 16, ..., 1: lane d < m adds lane d + m.
// The xor butterfly leaves the same bits in every lane (float + is commutative),
// equal to a shared-memory halving tree.
__device__ __forceinline__ float wave_tree_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, kWave);
  return v;
}

__device__ __forceinline__ int wave_and(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v &= __shfl_xor(v, m, kWave);
  return v;
}

}  // namespace qm
