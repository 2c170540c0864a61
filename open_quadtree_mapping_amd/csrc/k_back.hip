// k_back.hip -- tridiagonal smoothness upsampling (reference:
// src/depth_upsample.cu:20-302) and Gaussian x Beta fusion (reference:
// src/depth_fusion.cu:13-213) for a batch of keyframes.
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

// Thomas coefficients of A = I + lambda * L(w) for the unknown between couplings
// wl (to the previous unknown) and wr (to the next); lambda is a double as in
// the reference (upsample_lambda 10.0), the couplings are summed in float first.
__device__ __forceinline__ void thomas_coeffs(float wl, float wr, double lam, float& a, float& b, float& c) {
  a = (float)(-lam * (double)wl);
  b = (float)(1.0 + lam * (double)(wl + wr));
  c = (float)(-lam * (double)wr);
}

// ---------------------------------------------------------------------------
// Row pass: one lane per image row. Mode 0 composes the sparse map on the fly
// (BP winner depth at node pixels (4x, 4y), else the semi-dense pre-fill, else
// 0), so the full-resolution sparse map never has to be cleared or written;
// mode 1 reads an explicit sparse map; mode 2 is the generalised form with
// caller-supplied right-hand sides and weights. Also resets the fusion z-buffer
// row when asked (the fusion stage runs after the column pass).
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(64) void k_up_row(const FrameDev* __restrict__ frames, BlockMap bm,
                                               const float* __restrict__ wlut, double lam, int clear_zbuf) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int y = (blockIdx.x - bm.start[f]) * 64 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (y >= H) return;
  const size_t row = (size_t)y * W;
  if (clear_zbuf) {
    for (int x = 0; x < W; ++x) F.zbuf[row + x] = ~0ull;
  }
  const uint8_t* img = F.key + (size_t)y * F.pitch;
  auto weight = [&](int x) -> float {
    if (MODE == 2) return F.g_wrow[(size_t)y * (W + 1) + x];
    if (x == 0) return 0.0f;
    if (x == W) return F.edge_r ? 1.0f : 0.0f;
    return wlut[abs((int)img[x] - (int)img[x - 1])];
  };
  auto rhs = [&](int x, float& rx, float& rn) {
    if (MODE == 2) {
      const float d = F.g_d[row + x];
      if (F.g_conf) {
        const float c = F.g_conf[row + x];
        rx = c * d;
        rn = c;
      } else {
        rx = d;
        rn = d > 0.0f ? 1.0f : 0.0f;
      }
      return;
    }
    float s;
    if (MODE == 0) {
      s = F.prefill ? F.prefill[row + x] : 0.0f;
      if (((x | y) & 3) == 0 && (x >> 2) < F.Wc && (y >> 2) < F.Hc) {
        const int c = (y >> 2) * F.Wc + (x >> 2);
        if (F.label[c] >= 0) s = F.dcoarse[c];
      }
      if (F.sparse_out) F.sparse_out[row + x] = s;
    } else {
      s = F.sparse_in[row + x];
    }
    rx = s;
    rn = s > 0.0f ? 1.0f : 0.0f;
  };
  const int seg = F.row_seg < W ? F.row_seg : W;
  const int bounds[3] = {0, seg, W};
  for (int sg = 0; sg < 2; ++sg) {
    const int s0 = bounds[sg], s1 = bounds[sg + 1];
    if (s1 <= s0) continue;
    float fc = 0.0f, fd = 0.0f, fn = 0.0f;
    float wl = weight(s0);
    for (int i = s0; i < s1; ++i) {
      const float wr = weight(i + 1);
      float a, b, c;
      thomas_coeffs(wl, wr, lam, a, b, c);
      float rx, rn;
      rhs(i, rx, rn);
      const float den = b - fc * a;
      const float ch = c / den;
      const float dh = (rx - fd * a) / den;
      const float nh = (rn - fn * a) / den;
      F.fwd[row + i] = make_float4(ch, dh, nh, 0.0f);
      fc = ch;
      fd = dh;
      fn = nh;
      wl = wr;
    }
    fd = 0.0f;
    fn = 0.0f;
    for (int i = s1 - 1; i >= s0; --i) {
      const float4 t = F.fwd[row + i];
      const float xv = t.y - t.x * fd;
      const float nv = t.z - t.x * fn;
      F.rowxn[row + i] = make_float2(xv, nv);
      if (MODE == 2 && F.row_x_out) {
        F.row_x_out[row + i] = xv;
        F.row_n_out[row + i] = nv;
      }
      fd = xv;
      fn = nv;
    }
  }
}

// ---------------------------------------------------------------------------
// Column pass: one lane per column, so every step reads and writes one
// coalesced row segment across the wave. Output x / n (NaN where n = 0, as the
// reference).
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(64) void k_up_col(const FrameDev* __restrict__ frames, BlockMap bm,
                                               const float* __restrict__ wlut, double lam) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int x = (blockIdx.x - bm.start[f]) * 64 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (x >= W) return;
  auto weight = [&](int y) -> float {
    if (MODE == 2) return F.g_wcol[(size_t)y * W + x];
    if (y == 0) return 0.0f;
    if (y == H) return F.edge_c ? 1.0f : 0.0f;
    return wlut[abs((int)F.key[(size_t)y * F.pitch + x] - (int)F.key[(size_t)(y - 1) * F.pitch + x])];
  };
  const int seg = F.col_seg < H ? F.col_seg : H;
  const int bounds[3] = {0, seg, H};
  for (int sg = 0; sg < 2; ++sg) {
    const int s0 = bounds[sg], s1 = bounds[sg + 1];
    if (s1 <= s0) continue;
    float fc = 0.0f, fd = 0.0f, fn = 0.0f;
    float wl = weight(s0);
    for (int i = s0; i < s1; ++i) {
      const float wr = weight(i + 1);
      float a, b, c;
      thomas_coeffs(wl, wr, lam, a, b, c);
      const float2 r = F.rowxn[(size_t)i * W + x];
      const float den = b - fc * a;
      const float ch = c / den;
      const float dh = (r.x - fd * a) / den;
      const float nh = (r.y - fn * a) / den;
      F.fwd[(size_t)i * W + x] = make_float4(ch, dh, nh, 0.0f);
      fc = ch;
      fd = dh;
      fn = nh;
      wl = wr;
    }
    fd = 0.0f;
    fn = 0.0f;
    for (int i = s1 - 1; i >= s0; --i) {
      const float4 t = F.fwd[(size_t)i * W + x];
      const float xv = t.y - t.x * fd;
      const float nv = t.z - t.x * fn;
      const float out = xv / nv;
      if (F.colxn) F.colxn[(size_t)i * W + x] = make_float2(xv, nv);
      F.dense[(size_t)i * W + x] = out;
      if (F.dense_out) F.dense_out[(size_t)i * W + x] = out;
      fd = xv;
      fn = nv;
    }
  }
}

// ---------------------------------------------------------------------------
// Fusion.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool is_badpoint(float a, float b) {
  return ((double)a < 0.001) || ((double)(a / (a + b)) < 0.45);
}
__device__ __forceinline__ bool is_goodpoint(float a, float b) { return (double)(a / (a + b)) > 0.60; }

__global__ __launch_bounds__(256) void k_fill_zbuf(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int i = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  if (i < F.W * F.H) F.zbuf[i] = ~0ull;
}

// Warp the previous seeds into this frame (src/depth_fusion.cu:48-119). The
// z-buffer entry of a target pixel is the minimum of (float bits of the warped
// depth, source index) over its sources: nearest wins, ties -> lower index.
// Source 0 is the reference's empty sentinel and is never stored.
__global__ __launch_bounds__(256) void k_fuse_warp(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int idx = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (idx >= W * H) return;
  const int y = idx / W, x = idx - y * W;
  float4 s = F.seeds[idx];
  if (is_badpoint(s.x, s.y)) return;
  const V3 dir = ray_dir(x, y, F.fx, F.fy, F.cx, F.cy);
  V3 P;
  P.x = dir.x * s.z;
  P.y = dir.y * s.z;
  P.z = dir.z * s.z;
  const V3 q = se3_apply(F.T_fuse, P);
  const float nd = sqrtf((q.x * q.x + q.y * q.y) + q.z * q.z);
  if ((double)nd <= 0.5) return;
  const float pxf = (F.fx * q.x) / q.z + F.cx;
  const float pyf = (F.fy * q.y) / q.z + F.cy;
  const int pxi = trunc_i32((double)pxf + 0.5);
  const int pyi = trunc_i32((double)pyf + 0.5);
  if (pxi >= W || pxi < 0 || pyi >= H || pyi < 0) return;
  s.z = nd;
  s.w = (float)((double)s.w + (double)nd * 0.01);
  F.seeds[idx] = s;
  if (idx != 0) {
    uint32_t bits = __float_as_uint(nd);
    if (nd != nd) bits = 0x7FC00000u;
    const unsigned long long key = ((unsigned long long)bits << 32) | (unsigned long long)(uint32_t)idx;
    atomicMin(&F.zbuf[(size_t)pyi * W + pxi], key);
  }
}

// Hole filling from the z-buffer snapshot (an empty cell takes the first
// non-empty source in raster order whose position is in [2, W-2] x [2, H-2],
// src/depth_fusion.cu:121-144) and the per-pixel update
// (src/depth_fusion.cu:146-213).
__global__ __launch_bounds__(256) void k_fuse_update(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int idx = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (idx >= W * H) return;
  const int y = idx / W, x = idx - y * W;
  const unsigned long long empty = ~0ull;
  unsigned long long k = F.zbuf[idx];
  int src = k == empty ? 0 : (int)(uint32_t)(k & 0xFFFFFFFFull);
  if (src == 0) {
    for (int sy = y - 1; sy <= y + 1 && src == 0; ++sy)
      for (int sx = x - 1; sx <= x + 1; ++sx) {
        if (sx == x && sy == y) continue;
        if (sx < 2 || sx > W - 2 || sy < 2 || sy > H - 2) continue;
        const unsigned long long kk = F.zbuf[(size_t)sy * W + sx];
        if (kk != empty) {
          src = (int)(uint32_t)(kk & 0xFFFFFFFFull);
          break;
        }
      }
  }
  if (F.table_out) F.table_out[idx] = src;
  const float de = F.dense[idx];
  float unc = fmaxf(0.5f, (float)((double)de * 0.2));
  unc = unc * unc;
  if (de <= 0.0f) unc = 1e9f;
  float a, b, mu, s2;
  if (src > 0) {
    const float4 s = F.seeds[src];
    a = s.x;
    b = s.y;
    mu = s.z;
    s2 = s.w;
  } else {
    a = 10.0f;
    b = 10.0f;
    mu = de;
    s2 = 2500.0f;
  }
  if ((de - mu) * (de - mu) > unc + s2) {
    a = 10.0f;
    b = 10.0f;
    mu = de;
    s2 = 2500.0f;
  }
  const float new_sq = unc * s2 / (unc + s2);
  const float new_mu = (de * s2 + mu * unc) / (unc + s2);
  const float e0 = pinned_expf(((-(de - mu)) * (de - mu)) / (2.0f * (unc + s2)));
  const float r0 = 1.0f / sqrtf((float)(2.0 * 3.14159265358979323846 * (double)(unc + s2)));
  float c1 = (a / (a + b)) * (e0 * r0);
  float c2 = (b / (a + b)) * 1.0f / 50.0f;
  const float norm = c1 + c2;
  c1 = c1 / norm;
  c2 = c2 / norm;
  const float f1 = c1 * ((a + 1.0f) / (a + b + 1.0f)) + c2 * (a / (a + b + 1.0f));
  const float e1 = c1 * (((a + 1.0f) * (a + 2.0f)) / ((a + b + 1.0f) * (a + b + 2.0f))) +
                   c2 * (a * (a + 1.0f) / ((a + b + 1.0f) * (a + b + 2.0f)));
  const float mu_p = c1 * new_mu + c2 * mu;
  const float s_p = c1 * (new_sq + new_mu * new_mu) + c2 * (s2 + mu * mu) - mu_p * mu_p;
  const float a_p = (e1 - f1) / (f1 - e1 / f1);
  const float b_p = a_p * (1.0f - f1) / f1;
  if (F.depth_out) F.depth_out[idx] = is_goodpoint(a, b) ? mu_p : -1.0f;
  F.seeds_new[idx] = make_float4(a_p, b_p, mu_p, s_p);
}

// ---------------------------------------------------------------------------
static int rows_map(BlockMap& bm, const FrameDev* hframes, int n) {
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].H + 63) / 64;
  return make_block_map(bm, counts, n);
}
static int cols_map(BlockMap& bm, const FrameDev* hframes, int n) {
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].W + 63) / 64;
  return make_block_map(bm, counts, n);
}
static int pixels_map(BlockMap& bm, const FrameDev* hframes, int n) {
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].W * hframes[i].H + 255) / 256;
  return make_block_map(bm, counts, n);
}

void launch_upsample_row(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         double lambda, bool clear_zbuf, hipStream_t s) {
  BlockMap bm;
  const int nb = rows_map(bm, hframes, n);
  if (nb <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL(k_up_row<0>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, lambda, (int)clear_zbuf);
  else if (mode == 1)
    hipLaunchKernelGGL(k_up_row<1>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, lambda, (int)clear_zbuf);
  else
    hipLaunchKernelGGL(k_up_row<2>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, lambda, (int)clear_zbuf);
}

void launch_upsample_col(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         double lambda, hipStream_t s) {
  BlockMap bm;
  const int nb = cols_map(bm, hframes, n);
  if (nb <= 0) return;
  if (mode == 2)
    hipLaunchKernelGGL(k_up_col<2>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, lambda);
  else
    hipLaunchKernelGGL(k_up_col<0>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, lambda);
}

void launch_fill_zbuf(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  const int nb = pixels_map(bm, hframes, n);
  if (nb > 0) hipLaunchKernelGGL(k_fill_zbuf, dim3(nb), dim3(256), 0, s, dframes, bm);
}

void launch_fuse_warp(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  const int nb = pixels_map(bm, hframes, n);
  if (nb > 0) hipLaunchKernelGGL(k_fuse_warp, dim3(nb), dim3(256), 0, s, dframes, bm);
}

void launch_fuse_update(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  const int nb = pixels_map(bm, hframes, n);
  if (nb > 0) hipLaunchKernelGGL(k_fuse_update, dim3(nb), dim3(256), 0, s, dframes, bm);
}

void launch_upsample(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                     double lambda, bool clear_zbuf, hipStream_t s) {
  launch_upsample_row(dframes, hframes, n, mode, wlut, lambda, clear_zbuf, s);
  launch_upsample_col(dframes, hframes, n, mode, wlut, lambda, s);
}

void launch_fuse(const FrameDev* dframes, const FrameDev* hframes, int n, bool clear_zbuf, hipStream_t s) {
  if (clear_zbuf) launch_fill_zbuf(dframes, hframes, n, s);
  launch_fuse_warp(dframes, hframes, n, s);
  launch_fuse_update(dframes, hframes, n, s);
}

}  // namespace qm
