// k_adjoint.hip -- hand-written adjoints of the generalised upsample solve and
// of the per-pixel fusion update. No counterpart in the reference.
//
// Upsample. Per line, A = I + lambda * sum_e w_e M_e is symmetric tridiagonal
// (M_e = (u_{e-1} - u_e)(u_{e-1} - u_e)^T for an inner edge, u u^T for an edge
// that only sits on the diagonal: the two image borders and the chain split),
// so the adjoint of z = A^{-1} r is the same Thomas solve on the incoming
// gradient: rbar = A^{-1} zbar, and dL/dw_e = -lambda * rbar^T M_e z summed over
// the two right-hand sides (x and n). With out = x_col / n_col the incoming
// gradients are g / n_col and -g * out / n_col. The forward passes are
// recomputed first (the row pass into the rowx planes, the column pass into colxn); the
// adjoint runs the column solve, then the row solve.
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

// The adjoint kernels walk each line in chunks of kAK steps. A chunk's inputs
// are loaded together (indices clamped to the segment, so every load is in
// bounds; the steps past the segment's end are skipped), one chunk ahead of the
// steps that use them: two buffers alternate, so the loads of chunk k + 1 are
// in flight while chunk k runs.
constexpr int kAK = 16;

__device__ __forceinline__ void adj_coeffs(float wl, float wr, double lam, float& a, float& b, float& c) {
  a = (float)(-lam * (double)wl);
  b = (float)(1.0 + lam * (double)(wl + wr));
  c = (float)(-lam * (double)wr);
}

// Column adjoint: lane per column. rhs from g and the recomputed column pass,
// result (xbar, nbar) into adjxn, dL/dw_col into grad_wcol.
__global__ __launch_bounds__(64) void k_adj_col(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int x = (blockIdx.x - bm.start[f]) * 64 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (x >= W) return;
  const double lam = F.lam;
  const float flam = (float)lam;
  // dL/dw_e = lambda dL/d(lambda w_e); the pipeline backward takes the
  // unscaled term so that dL/dlambda = sum_e w_e dL/d(lambda w_e) needs no division
  const float gscale = F.grad_w_unscaled ? 1.0f : flam;
  auto weight = [&](int y) -> float { return gp(F.g_wcol)[(size_t)y * W + x]; };
  const int seg = F.col_seg < H ? F.col_seg : H;
  const int bounds[3] = {0, seg, H};
  float carry = 0.0f;
  for (int sg = 0; sg < 2; ++sg) {
    const int s0 = bounds[sg], s1 = bounds[sg + 1];
    if (s1 <= s0) continue;
    float fc = 0.0f, fd = 0.0f, fn = 0.0f;
    float wl = weight(s0);
    struct FwdIn {
      float w[kAK], g[kAK];
      float2 z[kAK];
    };
    auto load_f = [&](FwdIn& c, int i0) {
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        const int i = min(i0 + j, s1 - 1);
        c.w[j] = weight(i + 1);
        c.z[j] = gp(F.colxn)[(size_t)i * W + x];
        c.g[j] = gp(F.g_out)[(size_t)i * W + x];
      }
    };
    auto run_f = [&](const FwdIn& cin, int i0) {
      const int n = min(kAK, s1 - i0);
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        if (j < n) {
          const int i = i0 + j;
          const float wr = cin.w[j];
          float a, b, c;
          adj_coeffs(wl, wr, lam, a, b, c);
          const float2 z = cin.z[j];
          const float g = cin.g[j];
          float gx = 0.0f, gn = 0.0f;
          if (z.y != 0.0f) {
            gx = g / z.y;
            gn = -gx * (z.x / z.y);
          }
          const float den = b - fc * a;
          const float ch = c / den;
          const float dh = (gx - fd * a) / den;
          const float nh = (gn - fn * a) / den;
          gp(F.fwd)[(size_t)i * W + x] = make_float4(ch, dh, nh, 0.0f);
          fc = ch;
          fd = dh;
          fn = nh;
          wl = wr;
        }
      }
    };
    {
      FwdIn A, B;
      load_f(A, s0);
      for (int i0 = s0; i0 < s1; i0 += 2 * kAK) {
        load_f(B, i0 + kAK);
        run_f(A, i0);
        load_f(A, i0 + 2 * kAK);
        run_f(B, i0 + kAK);
      }
    }
    float vx_next = 0.0f, vn_next = 0.0f;
    float2 z_next = make_float2(0.0f, 0.0f);
    struct BwdIn {
      float4 t[kAK];
      float2 z[kAK];
    };
    auto load_b = [&](BwdIn& c, int i1) {
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        const int i = max(i1 - 1 - j, s0);
        c.t[j] = gp(F.fwd)[(size_t)i * W + x];
        c.z[j] = gp(F.colxn)[(size_t)i * W + x];
      }
    };
    auto run_b = [&](const BwdIn& cin, int i1) {
      const int n = min(kAK, i1 - s0);
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        if (j < n) {
          const int i = i1 - 1 - j;
          const float4 t = cin.t[j];
          const float vx = t.y - t.x * vx_next;
          const float vn = t.z - t.x * vn_next;
          gp(F.adjxn)[(size_t)x * H + i] = make_float2(vx, vn);
          const float2 z = cin.z[j];
          if (i + 1 < s1) {
            gp(F.grad_wcol)[(size_t)(i + 1) * W + x] = -gscale * ((vx - vx_next) * (z.x - z_next.x) + (vn - vn_next) * (z.y - z_next.y));
          } else {
            const float e = -gscale * (vx * z.x + vn * z.y);
            if (s1 == H) gp(F.grad_wcol)[(size_t)H * W + x] = e;
            else carry = e;
          }
          if (i == s0) {
            const float e = -gscale * (vx * z.x + vn * z.y);
            gp(F.grad_wcol)[(size_t)s0 * W + x] = (s0 == 0) ? e : carry + e;
          }
          vx_next = vx;
          vn_next = vn;
          z_next = z;
        }
      }
    };
    {
      BwdIn A, B;
      load_b(A, s1);
      for (int i1 = s1; i1 > s0; i1 -= 2 * kAK) {
        load_b(B, i1 - kAK);
        run_b(A, i1);
        load_b(A, i1 - 2 * kAK);
        run_b(B, i1 - kAK);
      }
    }
  }
}

// Row adjoint: lane per row. rhs = column-adjoint solution; z = recomputed row
// pass. Gives dL/dd, dL/dconf and dL/dw_row.
__global__ __launch_bounds__(64) void k_adj_row(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int y = (blockIdx.x - bm.start[f]) * 64 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (y >= H) return;
  const double lam = F.lam;
  const float flam = (float)lam;
  // dL/dw_e = lambda dL/d(lambda w_e); the pipeline backward takes the
  // unscaled term so that dL/dlambda = sum_e w_e dL/d(lambda w_e) needs no division
  const float gscale = F.grad_w_unscaled ? 1.0f : flam;
  const size_t row = (size_t)y * W;
  const size_t wrow = (size_t)y * (W + 1);
  auto weight = [&](int x) -> float { return gp(F.g_wrow)[wrow + x]; };
  const int seg = F.row_seg < W ? F.row_seg : W;
  const int bounds[3] = {0, seg, W};
  float carry = 0.0f;
  for (int sg = 0; sg < 2; ++sg) {
    const int s0 = bounds[sg], s1 = bounds[sg + 1];
    if (s1 <= s0) continue;
    float fc = 0.0f, fd = 0.0f, fn = 0.0f;
    float wl = weight(s0);
    struct FwdIn {
      float w[kAK];
      float2 r[kAK];
    };
    auto load_f = [&](FwdIn& c, int i0) {
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        const int i = min(i0 + j, s1 - 1);
        c.w[j] = weight(i + 1);
        c.r[j] = gp(F.adjxn)[(size_t)i * H + y];
      }
    };
    auto run_f = [&](const FwdIn& cin, int i0) {
      const int n = min(kAK, s1 - i0);
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        if (j < n) {
          const int i = i0 + j;
          const float wr = cin.w[j];
          float a, b, c;
          adj_coeffs(wl, wr, lam, a, b, c);
          const float2 r = cin.r[j];
          const float den = b - fc * a;
          const float ch = c / den;
          const float dh = (r.x - fd * a) / den;
          const float nh = (r.y - fn * a) / den;
          gp(F.fwd)[(size_t)i * H + y] = make_float4(ch, dh, nh, 0.0f);
          fc = ch;
          fd = dh;
          fn = nh;
          wl = wr;
        }
      }
    };
    {
      FwdIn A, B;
      load_f(A, s0);
      for (int i0 = s0; i0 < s1; i0 += 2 * kAK) {
        load_f(B, i0 + kAK);
        run_f(A, i0);
        load_f(A, i0 + 2 * kAK);
        run_f(B, i0 + kAK);
      }
    }
    float vx_next = 0.0f, vn_next = 0.0f;
    float2 z_next = make_float2(0.0f, 0.0f);
    struct BwdIn {
      float4 t[kAK];
      float2 z[kAK];
      float d[kAK], c[kAK];
    };
    // the confidences' row, or (unused) the depths' when there is none: every
    // load is made, so the prefetch's waits count a fixed number of loads
    const float* confp = F.g_conf ? F.g_conf : F.g_d;
    auto load_b = [&](BwdIn& c, int i1) {
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        const int i = max(i1 - 1 - j, s0);
        c.t[j] = gp(F.fwd)[(size_t)i * H + y];
        c.z[j] = make_float2(gp(F.rowx)[(size_t)i * H + y], gp(F.rowx)[(size_t)W * H + (size_t)i * H + y]);
        c.d[j] = gp(F.g_d)[row + i];
        c.c[j] = gp(confp)[row + i];
      }
    };
    auto run_b = [&](const BwdIn& cin, int i1) {
      const int n = min(kAK, i1 - s0);
#pragma unroll
      for (int j = 0; j < kAK; ++j) {
        if (j < n) {
          const int i = i1 - 1 - j;
          const float4 t = cin.t[j];
          const float vx = t.y - t.x * vx_next;
          const float vn = t.z - t.x * vn_next;
          const float2 z = cin.z[j];
          if (i + 1 < s1) {
            if (F.grad_wrow)
              gp(F.grad_wrow)[wrow + i + 1] = -gscale * ((vx - vx_next) * (z.x - z_next.x) + (vn - vn_next) * (z.y - z_next.y));
          } else {
            const float e = -gscale * (vx * z.x + vn * z.y);
            if (s1 == W) {
              if (F.grad_wrow) gp(F.grad_wrow)[wrow + W] = e;
            } else {
              carry = e;
            }
          }
          if (i == s0 && F.grad_wrow) {
            const float e = -gscale * (vx * z.x + vn * z.y);
            gp(F.grad_wrow)[wrow + s0] = (s0 == 0) ? e : carry + e;
          }
          // rhs_x = conf * d, rhs_n = conf (reference mode: rhs_x = d, rhs_n = 1{d > 0})
          const float d = cin.d[j];
          if (F.g_conf) {
            const float c = cin.c[j];
            if (F.grad_d) gp(F.grad_d)[row + i] = c * vx;
            if (F.grad_conf) gp(F.grad_conf)[row + i] = d * vx + vn;
          } else {
            if (F.grad_d) gp(F.grad_d)[row + i] = vx;
            if (F.grad_conf) gp(F.grad_conf)[row + i] = 0.0f;
          }
          vx_next = vx;
          vn_next = vn;
          z_next = z;
        }
      }
    };
    {
      BwdIn A, B;
      load_b(A, s1);
      for (int i1 = s1; i1 > s0; i1 -= 2 * kAK) {
        load_b(B, i1 - kAK);
        run_b(A, i1);
        load_b(A, i1 - 2 * kAK);
        run_b(B, i1 - kAK);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fusion update adjoint w.r.t. the dense depth input, by forward-mode dual
// numbers through the same operation sequence as k_fuse_update: one input, so
// a single tangent gives every output derivative; the vector-Jacobian product
// with (dL/d depth_out, dL/d seeds_out) is then a dot product per pixel.
// ---------------------------------------------------------------------------
struct Dual {
  float v, d;
};
__device__ __forceinline__ Dual dc(float v) { return {v, 0.0f}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const float q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual dexp(Dual a) {
  const float e = pinned_expf(a.v);
  return {e, e * a.d};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
  const float s = sqrtf(a.v);
  return {s, a.d / (2.0f * s)};
}

__global__ __launch_bounds__(256) void k_fuse_backward(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int idx = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (idx >= W * H) return;
  const int src = F.table_in[idx];
  const Dual de = {F.dense[idx], 1.0f};
  // uncertainty: max(0.5, 0.2 de)^2, or 1e9 for non-positive depth
  const float m02 = (float)((double)de.v * 0.2);
  Dual unc = (m02 > 0.5f) ? Dual{m02, 0.2f} : dc(0.5f);
  unc = unc * unc;
  if (de.v <= 0.0f) unc = dc(1e9f);
  Dual a, b, mu, s2;
  if (src > 0) {
    const float4 s = F.seeds[src];
    a = dc(s.x);
    b = dc(s.y);
    mu = dc(s.z);
    s2 = dc(s.w);
  } else {
    a = dc(10.0f);
    b = dc(10.0f);
    mu = de;
    s2 = dc(2500.0f);
  }
  if ((de.v - mu.v) * (de.v - mu.v) > unc.v + s2.v) {
    a = dc(10.0f);
    b = dc(10.0f);
    mu = de;
    s2 = dc(2500.0f);
  }
  const bool good = (double)(a.v / (a.v + b.v)) > 0.60;
  const Dual one = dc(1.0f), two = dc(2.0f);
  const Dual new_sq = unc * s2 / (unc + s2);
  const Dual new_mu = (de * s2 + mu * unc) / (unc + s2);
  const Dual ss = unc + s2;
  const Dual e0 = dexp(((-(de - mu)) * (de - mu)) / (two * ss));
  const Dual r0 = one / dsqrt(dc((float)(2.0 * 3.14159265358979323846)) * ss);
  Dual c1 = (a / (a + b)) * (e0 * r0);
  Dual c2 = (b / (a + b)) * one / dc(50.0f);
  const Dual norm = c1 + c2;
  c1 = c1 / norm;
  c2 = c2 / norm;
  const Dual f1 = c1 * ((a + one) / (a + b + one)) + c2 * (a / (a + b + one));
  const Dual e1 = c1 * (((a + one) * (a + two)) / ((a + b + one) * (a + b + two))) +
                  c2 * (a * (a + one) / ((a + b + one) * (a + b + two)));
  const Dual mu_p = c1 * new_mu + c2 * mu;
  const Dual s_p = c1 * (new_sq + new_mu * new_mu) + c2 * (s2 + mu * mu) - mu_p * mu_p;
  const Dual a_p = (e1 - f1) / (f1 - e1 / f1);
  const Dual b_p = a_p * (one - f1) / f1;
  float gsum = 0.0f;
  if (F.g_depth_out && good) gsum += F.g_depth_out[idx] * mu_p.d;
  if (F.g_seeds_out) {
    const float4 g = F.g_seeds_out[idx];
    gsum += g.x * a_p.d + g.y * b_p.d + g.z * mu_p.d + g.w * s_p.d;
  }
  F.grad_depth_in[idx] = gsum;
}

void launch_upsample_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  // recompute the forward passes (row pass -> rowx planes, column pass -> colxn)
  launch_upsample_row(dframes, hframes, n, 2, nullptr, s);
  launch_upsample_col(dframes, hframes, n, 2, nullptr, false, s);
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].W + 63) / 64;
  int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_adj_col, dim3(nb), dim3(64), 0, s, dframes, bm);
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].H + 63) / 64;
  nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_adj_row, dim3(nb), dim3(64), 0, s, dframes, bm);
}

// ---------------------------------------------------------------------------
// Pipeline backward (per-stream lambda). The keyframe upsample (mode 0) is the
// generalised solve with d = the composed sparse map, no confidence, and the
// edge weights of the keyframe, so its adjoint reuses the generalised kernels
// on those three arrays, materialised here with exactly the values the
// forward pass used: row edge x of row y is 0 at x = 0, edge_r at x = W and
// lut(|I(x) - I(x-1)|) inside; column edges likewise.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_compose_inputs(const FrameDev* __restrict__ frames, BlockMap bm,
                                                        const float* __restrict__ wlut) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int i = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (i >= (W + 1) * (H + 1)) return;
  const int y = i / (W + 1), x = i - y * (W + 1);  // x in [0, W], y in [0, H]
  const uint8_t* key = F.key;
  // the generalised inputs of this frame's descriptor are the scratch arrays
  // being filled here
  float* wrow = const_cast<float*>(F.g_wrow);
  float* wcol = const_cast<float*>(F.g_wcol);
  float* sparse = const_cast<float*>(F.g_d);
  if (y < H) {
    float w = 0.0f;
    if (x == W) w = F.edge_r ? 1.0f : 0.0f;
    else if (x > 0) w = wlut[abs((int)key[(size_t)y * F.pitch + x] - (int)key[(size_t)y * F.pitch + x - 1])];
    wrow[(size_t)y * (W + 1) + x] = w;
  }
  if (x < W) {
    float w = 0.0f;
    if (y == H) w = F.edge_c ? 1.0f : 0.0f;
    else if (y > 0) w = wlut[abs((int)key[(size_t)y * F.pitch + x] - (int)key[(size_t)(y - 1) * F.pitch + x])];
    wcol[(size_t)y * W + x] = w;
  }
  if (x < W && y < H) {
    float v = F.prefill ? F.prefill[(size_t)y * W + x] : 0.0f;
    if ((x & 3) == 0 && (y & 3) == 0 && (x >> 2) < F.Wc && (y >> 2) < F.Hc) {
      const size_t c = (size_t)(y >> 2) * F.Wc + (x >> 2);
      if (F.label[c] >= 0) v = F.dcoarse[c];
    }
    sparse[(size_t)y * W + x] = v;
  }
}

// dL/dlambda = sum over edges of w_e * dL/d(lambda w_e) (A depends on lambda
// and w only through lambda * w; the adjoint kernels wrote the unscaled
// dL/d(lambda w_e) for this, so lambda = 0 needs no division). Fixed-order
// block sums in double, one partial per block; the host adds a frame's
// partials in block order.
__global__ __launch_bounds__(256) void k_lambda_partials(const FrameDev* __restrict__ frames, BlockMap bm,
                                                         double* __restrict__ partial) {
  __shared__ double red[256];
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const size_t nrow = (size_t)H * (W + 1), ncol = (size_t)(H + 1) * W;
  const size_t i = (size_t)(blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  double v = 0.0;
  if (i < nrow) v = (double)F.g_wrow[i] * (double)F.grad_wrow[i];
  else if (i < nrow + ncol) v = (double)F.g_wcol[i - nrow] * (double)F.grad_wcol[i - nrow];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int m = 128; m >= 1; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

int lambda_blocks(const FrameDev& F) {
  return (int)((((size_t)F.H * (F.W + 1) + (size_t)(F.H + 1) * F.W) + 255) / 256);
}

void launch_compose_inputs(const FrameDev* dframes, const FrameDev* hframes, int n, const float* wlut,
                           hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = ((hframes[i].W + 1) * (hframes[i].H + 1) + 255) / 256;
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_compose_inputs, dim3(nb), dim3(256), 0, s, dframes, bm, wlut);
}

int launch_lambda_partials(const FrameDev* dframes, const FrameDev* hframes, int n, double* partial,
                           int* first_block, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = lambda_blocks(hframes[i]);
  const int nb = make_block_map(bm, counts, n);
  for (int i = 0; i <= n; ++i) first_block[i] = bm.start[i];
  if (nb > 0) hipLaunchKernelGGL(k_lambda_partials, dim3(nb), dim3(256), 0, s, dframes, bm, partial);
  return nb;
}

void launch_fuse_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].W * hframes[i].H + 255) / 256;
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_fuse_backward, dim3(nb), dim3(256), 0, s, dframes, bm);
}

}  // namespace qm
