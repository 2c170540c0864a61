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

constexpr int kChunk = 32;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Chunked walk of one serial chain [s0, s1) in steps of kChunk, with the
// inputs of the next chunk in flight while the current one is solved: two
// register buffers alternate (A, B). Loads of a chunk past the end re-read the
// last chunk (clamped indices), so every load is issued unconditionally and the
// chain only waits for the chunk it is about to use.
template <typename Buf, typename Load, typename Step>
__device__ __forceinline__ void chain_forward(int s0, int s1, Load load, Step step) {
  Buf A, B;
  int c0 = s0;
  load(A, c0);
  while (true) {
    const int c1 = c0 + kChunk;
    const bool more1 = c1 < s1;
    load(B, more1 ? c1 : c0);
    step(A, c0, min(kChunk, s1 - c0));
    if (!more1) break;
    const int c2 = c1 + kChunk;
    const bool more2 = c2 < s1;
    load(A, more2 ? c2 : c1);
    step(B, c1, min(kChunk, s1 - c1));
    if (!more2) break;
    c0 = c2;
  }
}
template <typename Buf, typename Load, typename Step>
__device__ __forceinline__ void chain_backward(int s0, int s1, Load load, Step step) {
  Buf A, B;
  int c0 = s0 + ((s1 - 1 - s0) / kChunk) * kChunk;
  load(A, c0);
  while (true) {
    const int c1 = c0 - kChunk;
    const bool more1 = c1 >= s0;
    load(B, more1 ? c1 : c0);
    step(A, c0, min(kChunk, s1 - c0));
    if (!more1) break;
    const int c2 = c1 - kChunk;
    const bool more2 = c2 >= s0;
    load(A, more2 ? c2 : c1);
    step(B, c1, kChunk);
    if (!more2) break;
    c0 = c2;
  }
}

struct F4Chunk {
  float4 t[kChunk];
};

// ---------------------------------------------------------------------------
// Row pass: one lane per image row, the serial Thomas chain walked in chunks.
// Scratch (fwd) and the result (rowxn) are column-major, [x][y], so the 64
// lanes of a wave (64 consecutive rows) store one contiguous 1 KB / 512 B row
// segment per step. Mode 0 composes the sparse map on the fly (BP winner
// depth at node pixels (4x, 4y), else the semi-dense pre-fill, else 0), so the
// full-resolution sparse map is never cleared or written unless asked for;
// mode 1 reads an explicit sparse map; mode 2 is the generalised form with
// caller-supplied right-hand sides and weights. Same operations, in the same
// order, as the reference's smooth_row (src/depth_upsample.cu:94-197).
// ---------------------------------------------------------------------------
template <int MODE>
struct RowIn {
  int px[kChunk + 2];   // image at c0 - 1 + k (modes 0, 1)
  float sp[kChunk];     // pre-fill (mode 0), sparse map (mode 1), d (mode 2)
  float cf[kChunk];     // confidence (mode 2)
  float wv[kChunk + 1];  // row weights (mode 2)
  int lb[kChunk / 4];   // BP labels of node cells (mode 0)
  float dv[kChunk / 4];  // BP depths of node cells (mode 0)
};

template <int MODE>
__global__ __launch_bounds__(64) void k_up_row(const FrameDev* __restrict__ frames, BlockMap bm,
                                               const float* __restrict__ wlut) {
  __shared__ float lut[256];
  if (MODE != 2) {
    for (int i = threadIdx.x; i < 256; i += 64) lut[i] = wlut[i];
    __syncthreads();
  }
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int y = (blockIdx.x - bm.start[f]) * 64 + threadIdx.x;
  const int W = F.W, H = F.H, Wc = F.Wc, Hc = F.Hc;
  const double lam = F.lam;
  if (y >= H) return;
  const size_t row = (size_t)y * W;
  const gptr<const uint8_t> img = gp(F.key + (size_t)y * F.pitch);
  const gptr<float4> fwd = gp(F.fwd);
  const gptr<float2> rowxn = gp(F.rowxn);
  const gptr<const float> prefill = gp(F.prefill);
  const gptr<const float> sparse_in = gp(F.sparse_in);
  const gptr<float> sparse_out = gp(F.sparse_out);
  const gptr<const float> g_d = gp(F.g_d);
  const gptr<const float> g_conf = gp(F.g_conf);
  const gptr<const float> g_wrow = gp(F.g_wrow);
  const gptr<float> row_x_out = gp(F.row_x_out);
  const gptr<float> row_n_out = gp(F.row_n_out);
  // node cells of this row (mode 0): rows 4*cy, cy < Hc; other rows read a
  // valid cell row and ignore it
  const bool node_row = (y & 3) == 0 && (y >> 2) < Hc;
  const size_t cb = (size_t)min(y >> 2, Hc - 1) * Wc;
  const gptr<const int32_t> label = gp((const int32_t*)F.label) + cb;
  const gptr<const float> dcoarse = gp((const float*)F.dcoarse) + cb;
  const float edge = F.edge_r ? 1.0f : 0.0f;
  const int seg = F.row_seg < W ? F.row_seg : W;
  auto load = [&](RowIn<MODE>& b, int c0) {
    if (MODE == 2) {
#pragma unroll
      for (int k = 0; k <= kChunk; ++k) b.wv[k] = g_wrow[(size_t)y * (W + 1) + min(c0 + k, W)];
#pragma unroll
      for (int k = 0; k < kChunk; ++k) b.sp[k] = g_d[row + min(c0 + k, W - 1)];
      if (g_conf) {
#pragma unroll
        for (int k = 0; k < kChunk; ++k) b.cf[k] = g_conf[row + min(c0 + k, W - 1)];
      }
    } else {
#pragma unroll
      for (int k = 0; k < kChunk + 2; ++k) b.px[k] = img[clampi(c0 - 1 + k, 0, W - 1)];
      if (MODE == 1) {
#pragma unroll
        for (int k = 0; k < kChunk; ++k) b.sp[k] = sparse_in[row + min(c0 + k, W - 1)];
      } else {
        if (prefill) {
#pragma unroll
          for (int k = 0; k < kChunk; ++k) b.sp[k] = prefill[row + min(c0 + k, W - 1)];
        }
        // node pixels (4x, 4y): chunk starts are multiples of 4 (0 or 16 * (W / 16), plus 32 k)
#pragma unroll
        for (int q = 0; q < kChunk / 4; ++q) {
          const int cx = min((c0 >> 2) + q, Wc - 1);
          b.lb[q] = label[cx];
          b.dv[q] = dcoarse[cx];
        }
      }
    }
  };
  for (int sg = 0; sg < 2; ++sg) {
    const int s0 = sg == 0 ? 0 : seg, s1 = sg == 0 ? seg : W;
    if (s1 <= s0) continue;
    float fc = 0.0f, fd = 0.0f, fn = 0.0f;
    chain_forward<RowIn<MODE>>(s0, s1, load, [&](RowIn<MODE>& b, int c0, int n) {
      float w[kChunk + 1], rx[kChunk], rn[kChunk];
      if (MODE == 2) {
#pragma unroll
        for (int k = 0; k <= kChunk; ++k) w[k] = b.wv[k];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
          if (g_conf) {
            rx[k] = b.cf[k] * b.sp[k];
            rn[k] = b.cf[k];
          } else {
            rx[k] = b.sp[k];
            rn[k] = b.sp[k] > 0.0f ? 1.0f : 0.0f;
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k <= kChunk; ++k) {
          const int x = c0 + k;
          w[k] = (x == 0) ? 0.0f : (x == W ? edge : lut[abs(b.px[k + 1] - b.px[k])]);
        }
        float sv[kChunk];
        if (MODE == 0) {
#pragma unroll
          for (int k = 0; k < kChunk; ++k) sv[k] = prefill ? b.sp[k] : 0.0f;
          if (node_row) {
#pragma unroll
            for (int q = 0; q < kChunk / 4; ++q)
              if (4 * q < n && (c0 >> 2) + q < Wc && b.lb[q] >= 0) sv[4 * q] = b.dv[q];
          }
          if (sparse_out) {
#pragma unroll
            for (int k = 0; k < kChunk; ++k)
              if (k < n) sparse_out[row + c0 + k] = sv[k];
          }
        } else {
#pragma unroll
          for (int k = 0; k < kChunk; ++k) sv[k] = b.sp[k];
        }
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
          rx[k] = sv[k];
          rn[k] = sv[k] > 0.0f ? 1.0f : 0.0f;
        }
      }
#pragma unroll
      for (int k = 0; k < kChunk; ++k) {
        if (k >= n) continue;
        float a, bb, c;
        thomas_coeffs(w[k], w[k + 1], lam, a, bb, c);
        const float den = bb - fc * a;
        const float ch = c / den;
        const float dh = (rx[k] - fd * a) / den;
        const float nh = (rn[k] - fn * a) / den;
        fwd[(size_t)(c0 + k) * H + y] = make_float4(ch, dh, nh, 0.0f);
        fc = ch;
        fd = dh;
        fn = nh;
      }
    });
    fd = 0.0f;
    fn = 0.0f;
    chain_backward<F4Chunk>(
        s0, s1,
        [&](F4Chunk& b, int c0) {
#pragma unroll
          for (int k = 0; k < kChunk; ++k) b.t[k] = fwd[(size_t)min(c0 + k, W - 1) * H + y];
        },
        [&](F4Chunk& b, int c0, int n) {
#pragma unroll
          for (int k = kChunk - 1; k >= 0; --k) {
            if (k >= n) continue;
            const float xv = b.t[k].y - b.t[k].x * fd;
            const float nv = b.t[k].z - b.t[k].x * fn;
            rowxn[(size_t)(c0 + k) * H + y] = make_float2(xv, nv);
            if (MODE == 2 && row_x_out) {
              row_x_out[row + c0 + k] = xv;
              row_n_out[row + c0 + k] = nv;
            }
            fd = xv;
            fn = nv;
          }
        });
  }
}

// ---------------------------------------------------------------------------
// Column pass: one lane per column, the same chunked chain down the rows. The
// row results are read from the lane's own column (column-major rowxn);
// coefficients, outputs and the image are row-major and coalesced across the
// wave. Output x / n (NaN where n = 0, as the reference,
// src/depth_upsample.cu:198-302). Also resets the fusion z-buffer column when
// asked (the fusion stage runs next).
// ---------------------------------------------------------------------------
template <int MODE>
struct ColIn {
  int px[kChunk + 2];   // key image at rows c0 - 1 + k (modes 0, 1)
  float wv[kChunk + 1];  // column weights (mode 2)
  float2 r[kChunk];     // row pass (x, n)
};

template <int MODE>
__global__ __launch_bounds__(64) void k_up_col(const FrameDev* __restrict__ frames, BlockMap bm,
                                               const float* __restrict__ wlut, int clear_zbuf) {
  __shared__ float lut[256];
  if (MODE != 2) {
    for (int i = threadIdx.x; i < 256; i += 64) lut[i] = wlut[i];
    __syncthreads();
  }
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int x = (blockIdx.x - bm.start[f]) * 64 + threadIdx.x;
  const int W = F.W, H = F.H, pitch = F.pitch;
  const double lam = F.lam;
  if (x >= W) return;
  const gptr<const uint8_t> key = gp(F.key) + x;
  const gptr<float4> fwd = gp(F.fwd);
  const gptr<const float2> colin = gp((const float2*)F.rowxn + (size_t)x * H);
  const gptr<float> dense = gp(F.dense);
  const gptr<float> dense_out = gp(F.dense_out);
  const gptr<float2> colxn = gp(F.colxn);
  const gptr<const float> g_wcol = gp(F.g_wcol);
  if (clear_zbuf) {
    const gptr<unsigned long long> zbuf = gp(F.zbuf);
    for (int i = 0; i < H; ++i) zbuf[(size_t)i * W + x] = ~0ull;
  }
  const float edge = F.edge_c ? 1.0f : 0.0f;
  const int seg = F.col_seg < H ? F.col_seg : H;
  auto load = [&](ColIn<MODE>& b, int c0) {
    if (MODE == 2) {
#pragma unroll
      for (int k = 0; k <= kChunk; ++k) b.wv[k] = g_wcol[(size_t)min(c0 + k, H) * W + x];
    } else {
#pragma unroll
      for (int k = 0; k < kChunk + 2; ++k) b.px[k] = key[(size_t)clampi(c0 - 1 + k, 0, H - 1) * pitch];
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) b.r[k] = colin[min(c0 + k, H - 1)];
  };
  for (int sg = 0; sg < 2; ++sg) {
    const int s0 = sg == 0 ? 0 : seg, s1 = sg == 0 ? seg : H;
    if (s1 <= s0) continue;
    float fc = 0.0f, fd = 0.0f, fn = 0.0f;
    chain_forward<ColIn<MODE>>(s0, s1, load, [&](ColIn<MODE>& b, int c0, int n) {
      float w[kChunk + 1];
      if (MODE == 2) {
#pragma unroll
        for (int k = 0; k <= kChunk; ++k) w[k] = b.wv[k];
      } else {
#pragma unroll
        for (int k = 0; k <= kChunk; ++k) {
          const int yy = c0 + k;
          w[k] = (yy == 0) ? 0.0f : (yy == H ? edge : lut[abs(b.px[k + 1] - b.px[k])]);
        }
      }
#pragma unroll
      for (int k = 0; k < kChunk; ++k) {
        if (k >= n) continue;
        float a, bb, c;
        thomas_coeffs(w[k], w[k + 1], lam, a, bb, c);
        const float den = bb - fc * a;
        const float ch = c / den;
        const float dh = (b.r[k].x - fd * a) / den;
        const float nh = (b.r[k].y - fn * a) / den;
        fwd[(size_t)(c0 + k) * W + x] = make_float4(ch, dh, nh, 0.0f);
        fc = ch;
        fd = dh;
        fn = nh;
      }
    });
    fd = 0.0f;
    fn = 0.0f;
    chain_backward<F4Chunk>(
        s0, s1,
        [&](F4Chunk& b, int c0) {
#pragma unroll
          for (int k = 0; k < kChunk; ++k) b.t[k] = fwd[(size_t)min(c0 + k, H - 1) * W + x];
        },
        [&](F4Chunk& b, int c0, int n) {
#pragma unroll
          for (int k = kChunk - 1; k >= 0; --k) {
            if (k >= n) continue;
            const float xv = b.t[k].y - b.t[k].x * fd;
            const float nv = b.t[k].z - b.t[k].x * fn;
            const float out = xv / nv;
            const size_t o = (size_t)(c0 + k) * W + x;
            if (colxn) colxn[o] = make_float2(xv, nv);
            dense[o] = out;
            if (dense_out) dense_out[o] = out;
            fd = xv;
            fn = nv;
          }
        });
  }
}

// ---------------------------------------------------------------------------
// Local (bilateral) upsample, the reference's alternate to the global solve
// (src/depth_upsample.cu:304-370, not called on its live path): each pixel
// averages the sparse depths of the 7 x 7 node grid around it (nodes every 4
// pixels) with weights exp(-|dx|^2 / 100 - dI^2 / 50), skipping non-positive
// depths; 0 / 0 = NaN where no node contributes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_up_local(const float* __restrict__ sparse, const uint8_t* __restrict__ img,
                                                  int W, int H, int pitch, float* __restrict__ dense) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  const float me = (float)img[(size_t)y * pitch + x];
  float acc = 0.0f, wsum = 0.0f;
  for (int a = 0; a < 7; ++a) {
    for (int b = 0; b < 7; ++b) {
      const int cx = (x / 4 + a - 3) * 4, cy = (y / 4 + b - 3) * 4;
      float w = 0.0f, dep = 0.0f;
      if (!(cx < 0 || cx >= W || cy < 0 || cy >= H)) {
        dep = sparse[(size_t)cy * W + cx];
        if ((double)dep <= 0.0) continue;
        const float di = me - (float)img[(size_t)cy * pitch + cx];
        const float sx = (float)(x - cx), sy = (float)(y - cy);
        const float sw = sx * sx + sy * sy;
        const float iw = di * di;
        w = pinned_expf(-sw / 100.0f - iw / 50.0f);
      }
      acc += w * dep;
      wsum += w;
    }
  }
  dense[i] = acc / wsum;
}

void launch_upsample_local(const float* sparse, const uint8_t* img, int W, int H, int pitch, float* dense,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_up_local, dim3((W * H + 255) / 256), dim3(256), 0, s, sparse, img, W, H, pitch, dense);
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
                         hipStream_t s) {
  BlockMap bm;
  const int nb = rows_map(bm, hframes, n);
  if (nb <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL(k_up_row<0>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut);
  else if (mode == 1)
    hipLaunchKernelGGL(k_up_row<1>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut);
  else
    hipLaunchKernelGGL(k_up_row<2>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut);
}

void launch_upsample_col(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         bool clear_zbuf, hipStream_t s) {
  BlockMap bm;
  const int nb = cols_map(bm, hframes, n);
  if (nb <= 0) return;
  if (mode == 2)
    hipLaunchKernelGGL(k_up_col<2>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, (int)clear_zbuf);
  else
    hipLaunchKernelGGL(k_up_col<0>, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, (int)clear_zbuf);
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
                     bool clear_zbuf, hipStream_t s) {
  launch_upsample_row(dframes, hframes, n, mode, wlut, s);
  launch_upsample_col(dframes, hframes, n, mode, wlut, clear_zbuf, s);
}

void launch_fuse(const FrameDev* dframes, const FrameDev* hframes, int n, bool clear_zbuf, hipStream_t s) {
  if (clear_zbuf) launch_fill_zbuf(dframes, hframes, n, s);
  launch_fuse_warp(dframes, hframes, n, s);
  launch_fuse_update(dframes, hframes, n, s);
}

}  // namespace qm
