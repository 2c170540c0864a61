// k_bp.hip -- hierarchical min-sum belief propagation and winner extraction
// (reference: src/bpextract.cu:37-395) for a batch of keyframes.
//
// One wave updates one cell: lane = inverse-depth sample (two per lane for
// N = 128, holding d and d + 64). A cell's four outgoing messages are stored
// together, [cell][dir][N] with dir = up, down, left, right, so the four 256 B
// rows are one contiguous 1 KB write. The message upsample between levels is
// folded into the first half-sweep of each level: that sweep visits every cell,
// reads neighbour messages straight from the parent level, and copies the
// parent's messages into cells it does not update.
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

template <int N>
struct Lanes {
  float v[N / 64];
};

// Outgoing message for one direction from h (this direction's belief without
// the target's message): min(h[d], h[d-1]+P1, h[d+1]+P1, min h + P2), minus the
// tree mean (src/bpextract.cu:272-302).
template <int N>
__device__ __forceinline__ Lanes<N> message_from(const Lanes<N>& h, float p1, float p2, int lane) {
  Lanes<N> out;
  if constexpr (N == 64) {
    const float hmin = tree_min64(h.v[0]);
    const float hl = __shfl_up(h.v[0], 1, 64);
    const float hr = __shfl_down(h.v[0], 1, 64);
    float m = h.v[0];
    if (lane > 0) m = fminf(m, hl + p1);
    if (lane < 63) m = fminf(m, hr + p1);
    m = fminf(m, hmin + p2);
    const float s = tree_sum64(m);
    out.v[0] = m - s / 64.0f;
  } else {
    const float h0 = h.v[0], h1 = h.v[1];
    const float hmin = tree_min64(h0 < h1 ? h0 : h1);
    const float l0 = __shfl_up(h0, 1, 64);
    const float r0 = __shfl_down(h0, 1, 64);
    const float l1 = __shfl_up(h1, 1, 64);
    const float r1 = __shfl_down(h1, 1, 64);
    const float h1_first = __shfl(h1, 0, 64);   // d = 64, right neighbour of d = 63
    const float h0_last = __shfl(h0, 63, 64);   // d = 63, left neighbour of d = 64
    float m0 = h0, m1 = h1;
    if (lane > 0) m0 = fminf(m0, l0 + p1);
    m0 = fminf(m0, (lane < 63 ? r0 : h1_first) + p1);
    m1 = fminf(m1, (lane > 0 ? l1 : h0_last) + p1);
    if (lane < 63) m1 = fminf(m1, r1 + p1);
    m0 = fminf(m0, hmin + p2);
    m1 = fminf(m1, hmin + p2);
    const float s = tree_sum64(m0 + m1);  // first tree stage: d + (d + 64)
    out.v[0] = m0 - s / 128.0f;
    out.v[1] = m1 - s / 128.0f;
  }
  return out;
}

__device__ __forceinline__ bool cell_skipped(const FrameDev& F, int l, int bx, int y) {
  const int lev = l + 2;
  if (lev > 4) return false;
  // the lookup uses the half-row index bx, not the cell's x (src/bpextract.cu:209-228)
  int X = bx << lev, Y = y << lev;
  X = X > F.W - 1 ? F.W - 1 : X;
  Y = Y > F.H - 1 ? F.H - 1 : Y;
  return (int)F.levels[(size_t)Y * F.W + X] > lev;
}

template <int N>
__global__ __launch_bounds__(256) void k_bp_sweep(const FrameDev* __restrict__ frames, BlockMap bm, int l, int A,
                                                  int first, int active_ok, float p1, float p2) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int lane = threadIdx.x & 63;
  const int item = (blockIdx.x - bm.start[f]) * 4 + (threadIdx.x >> 6);
  const int w = F.lw[l], h = F.lh[l];
  int x, y;
  bool active;
  if (first) {
    if (item >= w * h) return;
    y = item / w;
    x = item - y * w;
    const bool mine = A ? ((x & 1) == (y & 1)) : ((x & 1) != (y & 1));
    active = active_ok && mine && !cell_skipped(F, l, x >> 1, y);
  } else {
    const int hw = (w + 1) >> 1;
    if (item >= hw * h) return;
    y = item / hw;
    const int bx = item - y * hw;
    if (cell_skipped(F, l, bx, y)) return;
    x = 2 * bx + (A ? (y & 1) : ((y + 1) & 1));
    if (x >= w) return;
    active = true;
  }
  const gptr<float> M = gp(F.msg[l]);
  const gptr<float> dst = M + ((size_t)y * w + x) * 4 * N;
  const gptr<const float> P = gp((const float*)((first && l < 3) ? F.msg[l + 1] : nullptr));
  const int wp = (l < 3) ? F.lw[l + 1] : 0;
  if (!active) {
    // first sweep only: copy the parent's messages (zeros at the coarsest level)
#pragma unroll
    for (int dir = 0; dir < 4; ++dir)
#pragma unroll
      for (int k = 0; k < N / 64; ++k) {
        const int d = lane + 64 * k;
        dst[dir * N + d] = P ? P[(((size_t)(y >> 1) * wp + (x >> 1)) * 4 + dir) * N + d] : 0.0f;
      }
    return;
  }
  // incoming messages: from the up neighbour its "down" message, from the down
  // neighbour its "up" message, from the left its "right", from the right its "left"
  auto incoming = [&](int nx, int ny, int dir, int d) -> float {
    if (first) {
      if (!P) return 0.0f;
      return P[(((size_t)(ny >> 1) * wp + (nx >> 1)) * 4 + dir) * N + d];
    }
    return M[(((size_t)ny * w + nx) * 4 + dir) * N + d];
  };
  const gptr<const float> data = gp((const float*)F.cost[l]) + ((size_t)y * w + x) * N;
  Lanes<N> hd[4];
#pragma unroll
  for (int k = 0; k < N / 64; ++k) {
    const int d = lane + 64 * k;
    const float dv = data[d];
    const bool has_up = y > 0, has_down = y < h - 1, has_left = x > 0, has_right = x < w - 1;
    const float fu = has_up ? incoming(x, y - 1, 1, d) : 0.0f;
    const float fd = has_down ? incoming(x, y + 1, 0, d) : 0.0f;
    const float fl = has_left ? incoming(x - 1, y, 3, d) : 0.0f;
    const float fr = has_right ? incoming(x + 1, y, 2, d) : 0.0f;
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) {
      float v = dv;
      if (dir != 0 && has_up) v += fu;
      if (dir != 1 && has_down) v += fd;
      if (dir != 2 && has_left) v += fl;
      if (dir != 3 && has_right) v += fr;
      hd[dir].v[k] = v;
    }
  }
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    const Lanes<N> o = message_from<N>(hd[dir], p1, p2, lane);
#pragma unroll
    for (int k = 0; k < N / 64; ++k) dst[dir * N + lane + 64 * k] = o.v[k];
  }
}

// Winner extraction (src/bpextract.cu:314-372): belief = data + messages from
// the four neighbours, tree argmin (ties keep the lower index at every stage),
// parabola sub-sample refinement, depth = 1 / (step * d + min) in double.
template <int N>
__global__ __launch_bounds__(256) void k_extract(const FrameDev* __restrict__ frames, BlockMap bm, double imin,
                                                 double istep) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int lane = threadIdx.x & 63;
  const int cell = (blockIdx.x - bm.start[f]) * 4 + (threadIdx.x >> 6);
  const int w = F.lw[0], h = F.lh[0];
  if (cell >= w * h) return;
  const int y = cell / w, x = cell - y * w;
  const int qs = 1 << F.levels[(size_t)(4 * y) * F.W + 4 * x];
  if (((4 * x) % qs) != 0 || ((4 * y) % qs) != 0) {
    if (lane == 0) {
      F.dcoarse[cell] = 0.0f;
      F.label[cell] = -1;
    }
    return;
  }
  const gptr<const float> M = gp((const float*)F.msg[0]);
  const gptr<const float> data = gp((const float*)F.cost[0]) + (size_t)cell * N;
  float b[N / 64];
#pragma unroll
  for (int k = 0; k < N / 64; ++k) {
    const int d = lane + 64 * k;
    float v = data[d];
    if (x != 0) v += M[(((size_t)cell - 1) * 4 + 3) * N + d];
    if (x != w - 1) v += M[(((size_t)cell + 1) * 4 + 2) * N + d];
    if (y != 0) v += M[(((size_t)cell - w) * 4 + 1) * N + d];
    if (y != h - 1) v += M[(((size_t)cell + w) * 4 + 0) * N + d];
    b[k] = v;
  }
  float v;
  int id;
  if constexpr (N == 64) {
    v = b[0];
    id = lane;
  } else {
    // first tree stage (offset 64) inside the lane
    if (b[1] < b[0]) {
      v = b[1];
      id = lane + 64;
    } else {
      v = b[0];
      id = lane;
    }
  }
  tree_argmin64(v, id);
  // neighbours of the winner
  const int dm1 = id > 0 ? id - 1 : 0, dp1 = id < N - 1 ? id + 1 : N - 1;
  float pre, post;
  if constexpr (N == 64) {
    pre = __shfl(b[0], dm1, 64);
    post = __shfl(b[0], dp1, 64);
  } else {
    const float pre0 = __shfl(b[0], dm1 & 63, 64), pre1 = __shfl(b[1], dm1 & 63, 64);
    const float post0 = __shfl(b[0], dp1 & 63, 64), post1 = __shfl(b[1], dp1 & 63, 64);
    pre = (dm1 >> 6) ? pre1 : pre0;
    post = (dp1 >> 6) ? post1 : post0;
  }
  if (lane == 0) {
    const gptr<float> dco = gp(F.dcoarse);
    const gptr<int32_t> lab = gp(F.label);
    float disp = (float)id;
    if (id > 0 && id < N - 1) {
      const float a = (pre - 2.0f * v) + post;
      const float bb = (-pre) + post;
      const float ba = bb / a;
      if (isfinite(ba)) disp = (float)id - ba / 2.0f;
    }
    dco[cell] = (float)(1.0 / (istep * (double)disp + imin));
    lab[cell] = id;
  }
}

// [cell][dir][N] -> [dir][cell][N] for callers that want the reference layout
template <int N>
__global__ __launch_bounds__(256) void k_copy_messages(const FrameDev* __restrict__ frames, float* out) {
  const FrameDev& F = frames[0];
  const int ncell = F.lw[0] * F.lh[0];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)ncell * 4 * N) return;
  const int d = (int)(i % N);
  const size_t t = i / N;
  const int dir = (int)(t % 4);
  const size_t cell = t / 4;
  out[((size_t)dir * ncell + cell) * N + d] = F.msg[0][i];
}

void launch_bp_sweep(const FrameDev* dframes, const FrameDev* hframes, int n, int N, int l, bool A, bool first,
                     bool active, const BpLaunchParams& p, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) {
    const int w = hframes[i].lw[l], h = hframes[i].lh[l];
    const int items = first ? w * h : ((w + 1) / 2) * h;
    counts[i] = (items + 3) / 4;
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 64)
    hipLaunchKernelGGL(k_bp_sweep<64>, dim3(nb), dim3(256), 0, s, dframes, bm, l, (int)A, (int)first, (int)active,
                       p.p1, p.p2);
  else
    hipLaunchKernelGGL(k_bp_sweep<128>, dim3(nb), dim3(256), 0, s, dframes, bm, l, (int)A, (int)first, (int)active,
                       p.p1, p.p2);
}

void launch_extract(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const BpLaunchParams& p,
                    hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].lw[0] * hframes[i].lh[0] + 3) / 4;
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 64)
    hipLaunchKernelGGL(k_extract<64>, dim3(nb), dim3(256), 0, s, dframes, bm, p.idepth_min, p.idepth_step);
  else
    hipLaunchKernelGGL(k_extract<128>, dim3(nb), dim3(256), 0, s, dframes, bm, p.idepth_min, p.idepth_step);
}

void launch_copy_messages(const FrameDev* dframes, const FrameDev* hframes, int n, int N, float* out,
                          hipStream_t s) {
  (void)n;
  const size_t total = (size_t)hframes[0].lw[0] * hframes[0].lh[0] * 4 * N;
  const int nb = (int)((total + 255) / 256);
  if (N == 64)
    hipLaunchKernelGGL(k_copy_messages<64>, dim3(nb), dim3(256), 0, s, dframes, out);
  else
    hipLaunchKernelGGL(k_copy_messages<128>, dim3(nb), dim3(256), 0, s, dframes, out);
}

}  // namespace qm
