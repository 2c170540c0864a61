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
// recomputed first (the row pass into the rowx planes, the column pass into
// colxn); the adjoint solves then run in the forward passes' wave pipeline
// (k_up_col / k_up_row mode 3, k_back.hip): columns, then rows.
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

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

// dL/dlambda of each frame with a grad_lambda output: the sum of its H + W
// per-line partials (rows, then columns), in one fixed order: thread t adds
// lines t, t + 256, ... in turn, then a fixed pairwise tree over the threads.
__global__ __launch_bounds__(256) void k_lambda_reduce(const FrameDev* __restrict__ frames, BlockMap bm) {
  __shared__ double red[256];
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int n = F.H + F.W;
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += F.lam_part[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int m = 128; m >= 1; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) *F.grad_lambda = red[0];
}

void launch_upsample_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s,
                              bool recompute) {
  // recompute the forward passes unless the forward kept them (row pass ->
  // rowx planes, column pass -> colxn), then the adjoint solves (the same wave
  // pipeline, mode 3): columns (-> adjxn, dL/dw_col), then rows (-> dL/dd,
  // dL/dconf, dL/dw_row), each with its per-line lambda partials
  if (recompute) {
    launch_upsample_row(dframes, hframes, n, 2, nullptr, s);
    launch_upsample_col(dframes, hframes, n, 2, nullptr, false, s);
  }
  launch_upsample_col(dframes, hframes, n, 3, nullptr, false, s);
  launch_upsample_row(dframes, hframes, n, 3, nullptr, s);
  launch_adj_dconf(dframes, hframes, n, s);  // dL/dd, dL/dconf from the row adjoint's solution planes
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].grad_lambda && hframes[i].lam_part) ? 1 : 0;
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_lambda_reduce, dim3(nb), dim3(256), 0, s, dframes, bm);
}

// ---------------------------------------------------------------------------
// Pipeline backward (per-stream lambda). The keyframe upsample (mode 0) is the
// generalised solve with d = the composed sparse map, no confidence, and the
// edge weights of the keyframe, so its adjoint reuses the generalised kernels
// on those three arrays, materialised here with exactly the values the
// forward pass used: row edge x of row y is 0 at x = 0, edge_r at x = W and
// lut(|I(x) - I(x-1)|) inside; column edges likewise.
// ---------------------------------------------------------------------------
// One block per (row y in [0, H], 256 columns): a thread forms 4 consecutive
// columns (row edges x in [0, W], column edges of row y, sparse values);
// want_sparse = 0 leaves the sparse map out (an adjoint that neither
// recomputes the forward nor returns dL/dd or dL/dconf never reads it).
__global__ __launch_bounds__(64) void k_compose_inputs(const FrameDev* __restrict__ frames, BlockMap bm,
                                                       const float* __restrict__ wlut, int want_sparse) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const int per_row = (W + 256) / 256;  // blocks per row: columns 0 .. W
  const int b = blockIdx.x - bm.start[f];
  const int y = b / per_row, x0 = (b - y * per_row) * 256 + 4 * (int)threadIdx.x;
  const gptr<const uint8_t> key = gp(F.key);
  // the generalised inputs of this frame's descriptor are the scratch arrays
  // being filled here
  const gptr<float> wrow = gp(const_cast<float*>(F.g_wrow));
  const gptr<float> wcol = gp(const_cast<float*>(F.g_wcol));
  const gptr<float> sparse = gp(const_cast<float*>(F.g_d));
  const gptr<const float> prefill = gp(F.prefill);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int x = x0 + k;
    if (x > W) break;
    const int cur = x < W && y < H ? (int)key[(size_t)y * F.pitch + x] : 0;
    if (y < H) {
      float w = 0.0f;
      if (x == W) w = F.edge_r ? 1.0f : 0.0f;
      else if (x > 0) w = wlut[abs(cur - (int)key[(size_t)y * F.pitch + x - 1])];
      wrow[(size_t)y * (W + 1) + x] = w;
    }
    if (x < W) {
      float w = 0.0f;
      if (y == H) w = F.edge_c ? 1.0f : 0.0f;
      else if (y > 0) w = wlut[abs(cur - (int)key[(size_t)(y - 1) * F.pitch + x])];
      wcol[(size_t)y * W + x] = w;
    }
    if (want_sparse && x < W && y < H) {
      float v = prefill ? prefill[(size_t)y * W + x] : 0.0f;
      if ((x & 3) == 0 && (y & 3) == 0 && (x >> 2) < F.Wc && (y >> 2) < F.Hc) {
        const size_t c = (size_t)(y >> 2) * F.Wc + (x >> 2);
        if (gp(F.label)[c] >= 0) v = gp(F.dcoarse)[c];
      }
      sparse[(size_t)y * W + x] = v;
    }
  }
}

void launch_compose_inputs(const FrameDev* dframes, const FrameDev* hframes, int n, const float* wlut,
                           hipStream_t s, bool want_sparse) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].H + 1) * ((hframes[i].W + 256) / 256);
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_compose_inputs, dim3(nb), dim3(64), 0, s, dframes, bm, wlut, (int)want_sparse);
}

void launch_fuse_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].W * hframes[i].H + 255) / 256;
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_fuse_backward, dim3(nb), dim3(256), 0, s, dframes, bm);
}

}  // namespace qm
