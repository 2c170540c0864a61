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
  double lam;          // smoothness weight lambda of the upsample (per frame)
  const uint8_t* key;
  const uint8_t* ref[4];
  float T[4][12];       // key -> reference r
  float T_fuse[12];     // last fused frame -> this frame
  const float* prefill;  // H x W or null
  const float2* prior;   // H x W semi-dense prior (inverse depth, variance) for the prior cost, or null
  const float* prior_depth;  // H x W prior depth for the depth quadtree, or null
  const float* sparse_in;  // explicit sparse map (stage API) or null
  uint8_t* levels;      // H x W
  float* cost[4];       // lh*lw*N per level
  float* msg[4];        // lh*lw*4*N per level, [cell][dir][N], dir = up, down, left, right
  float* msg2[4];       // a second message buffer per level (the fused sweep pairs read one and write the other)
  float* dcoarse;       // Hc x Wc
  int32_t* label;       // Hc x Wc
  int32_t* count;       // Hc x Wc or null
  float* sparse_out;    // H x W or null
  float4* fwd;          // H x W scratch: upsample chunk checkpoints, adjoint Thomas coefficients
  float* rowx;          // row pass result: X then N, each a W x H plane [x][y]
  float* dense;         // H x W
  float* dense_out;     // H x W copy for the caller or null
  float4* seeds;        // H x W (warped in place unless warp_skip)
  int warp_skip;        // 1: the warp leaves the seeds as they are (the update rebuilds the winner's)
  float4* seeds_new;    // H x W
  unsigned long long* zbuf;  // H x W
  unsigned long long* zbuf_clear;  // H x W: the z-buffer the column pass resets (in a batch the next run's)
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
  float* adj_rowv;      // the row adjoint's solution: X then N, each a W x H plane [x][y], or null
  const float* g_out;   // dL/d(dense)
  float* grad_d;
  float* grad_conf;
  float* grad_wrow;     // H x (W+1)
  float* grad_wcol;     // (H+1) x W
  int grad_w_unscaled;  // 1: grad_w* receive dL/d(lambda w_e) (no lambda factor), for dL/dlambda
  double* lam_part;     // H + W per-line sums of w_e dL/d(lambda w_e) (rows, then columns) or null
  double* grad_lambda;  // dL/dlambda of the frame (one double) or null
  // adjoint of the fusion update
  const int32_t* table_in;     // hole-filled warp table
  const float* g_depth_out;    // dL/d(depth_out)
  const float4* g_seeds_out;   // dL/d(seeds_out) or null
  float* grad_depth_in;
};

// Pointers into the frame buffers, typed as global-memory pointers in device
// code so that loads and stores compile to global_* instructions (flat ones
// also count against the scalar-load counter, so every scalar reload of a
// descriptor field would wait for all outstanding vector loads). Kernels copy
// the descriptor fields they use into locals of this type once.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
#endif
template <typename T>
__device__ __forceinline__ gptr<T> gp(T* p) {
  return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<const T> gp(const T* p) {
  return (gptr<const T>)p;
}

// Streaming (non-temporal) accesses, for data a kernel touches once: the
// caches keep such lines the shortest, so they do not push out data that
// other kernels of the run read again
typedef float F4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load4_nt(gptr<const float4> p) {
  const F4v t = __builtin_nontemporal_load((gptr<const F4v>)p);
  return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void store4_nt(gptr<float4> p, float4 v) {
  const F4v t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, (gptr<F4v>)p);
}

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

// The same rounding for the texel selection of the cost volume, exact in float:
// floor(v) + (v - floor(v) >= 0.5), where v - floor(v) is exact (Sterbenz). It
// equals trunc((double)v + 0.5) for v >= -0.5 and is <= 0 below, where the
// callers reject (centre test p >= 2) or clamp (patch taps, to 0) either way;
// NaN maps far below 0 as well.
__device__ __forceinline__ int round_half_up(float v) {
  v = fminf(fmaxf(v, -1.0e9f), 1.0e9f);
  const float fl = floorf(v);
  return (int)fl + ((v - fl) >= 0.5f ? 1 : 0);
}

__device__ __forceinline__ float pinned_expf(float x) { return (float)exp((double)x); }

struct V3 {
  float x, y, z;
};

// a 3x4 row-major rigid transform passed by value to a kernel
struct Se3 {
  float m[12];
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

// ---- row layout of cost and message rows ----
// Inside the pipeline a row of N samples (one cell's costs, or one direction of
// one cell's messages) keeps sample d = j + 16 s at position S*j + s, with
// S = N / 16, j < 16, s < S. The 16 lanes of a DPP row then hold one row, lane
// j owning the S samples j, j + 16, ..., j + 16 (S - 1) as S consecutive floats:
// one row is one 16-lane vector load, the reference's halving-tree stages
// N/2 .. 16 pair samples inside a lane, and stages 8 .. 1 pair lanes of a row.
template <int N>
__host__ __device__ constexpr int row_pos(int d) {
  return (N / 16) * (d & 15) + (d >> 4);
}
template <int N>
__host__ __device__ constexpr int row_sample(int p) {
  return (p % (N / 16)) * 16 + p / (N / 16);
}

// ---- DPP moves inside rows of 16 lanes (gfx950) ----
// update_dpp(old, v, ctrl): each lane reads v from its source lane; a lane whose
// source lies outside its row keeps `old`. row_shl:k = lane j + k, row_shr:k =
// lane j - k, row_ror:k = lane (j - k) mod 16, row_newbcast:0 = lane 0 of the row.
constexpr int kRowShl1 = 0x101, kRowShr1 = 0x111, kRowRor1 = 0x121, kRowRor15 = 0x12F, kRowBcast0 = 0x150;
template <int CTRL>
__device__ __forceinline__ float dpp_old(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_old_i(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xF, 0xF, false);
}
// row_shl:K for halving trees: lanes j + K >= 16 read 0 (bound_ctrl), which only
// reaches lanes whose partial results the tree never uses
template <int K>
__device__ __forceinline__ float dpp_row_shl(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + K, 0xF, 0xF, true));
}
template <int K>
__device__ __forceinline__ int dpp_row_shl_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x100 + K, 0xF, 0xF, true);
}
// moves whose source lane always exists (rotations, broadcasts): no old value
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_bcast0(float v) { return dpp_old<kRowBcast0>(v, v); }
__device__ __forceinline__ int row_bcast0_i(int v) { return dpp_old_i<kRowBcast0>(v, v); }

// IEEE minimum (v_minimum3_f32, no canonicalising moves): equal to fminf and to
// the reference's compare-select min for the finite, NaN-free values of a cost
// volume and its messages. The sign of a zero minimum never reaches an output:
// minima are only used as min + P1 and min + P2 (P1, P2 > 0) or compared.
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_elementwise_minimum(a, b); }

}  // namespace qm
