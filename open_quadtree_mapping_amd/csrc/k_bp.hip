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

// ---- cross-lane moves of the sweep (gfx950) ----
// DPP wave shifts: wave_shr:1 hands lane d the value of lane d - 1, wave_shl:1
// that of lane d + 1, wave_ror:1 / wave_rol:1 the same with wrap-around. A
// lane without a source (lane 0 / 63 of the shifts) keeps `old`.
constexpr int kWaveShl1 = 0x130, kWaveRol1 = 0x134, kWaveShr1 = 0x138, kWaveRor1 = 0x13C;
template <int CTRL>
__device__ __forceinline__ float dpp_old(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// IEEE minimum (v_minimum3_f32): equal to fminf and to the reference's
// compare-select min for the finite, NaN-free values of a cost volume (the
// sign of a zero minimum never reaches an output: it is only used as
// min + P2 and min + P1 with P1, P2 > 0, or compared).
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_elementwise_minimum(a, b); }

struct Min {
  __device__ float operator()(float a, float b) const { return vmin(a, b); }
};
struct Add {
  __device__ float operator()(float a, float b) const { return a + b; }
};

// Four lane-0 halving trees at once (stage m: element d < m combines element
// d + m, m = 32, 16, ..., 1), one per direction. permlane32_swap(a, b) yields
// [a(0..31), b(0..31)] and [a(32..63), b(32..63)], so one combine does stage 32
// of a in lanes 0-31 and of b in lanes 32-63; permlane16_swap(P, Q) yields
// [P.r0, Q.r0, P.r2, Q.r2] and [P.r1, Q.r1, P.r3, Q.r3] (rows of 16 lanes), so
// one combine does stage 16 of all four, leaving rows (a0, a2, a1, a3); DPP
// row_shl:8/4/2/1 finishes each row. Every element meets exactly the partner
// of the reference's shared-memory tree (src/bpextract.cu:274-302), and the
// combine is commutative, so the four results are bit-identical to it.
template <typename Op>
__device__ __forceinline__ void tree4(float a0, float a1, float a2, float a3, Op op, float (&r)[4]) {
  const auto s01 = __builtin_amdgcn_permlane32_swap(__float_as_uint(a0), __float_as_uint(a1), false, false);
  const auto s23 = __builtin_amdgcn_permlane32_swap(__float_as_uint(a2), __float_as_uint(a3), false, false);
  const float P = op(__uint_as_float(s01[0]), __uint_as_float(s01[1]));
  const float Q = op(__uint_as_float(s23[0]), __uint_as_float(s23[1]));
  const auto pq = __builtin_amdgcn_permlane16_swap(__float_as_uint(P), __float_as_uint(Q), false, false);
  float R = op(__uint_as_float(pq[0]), __uint_as_float(pq[1]));
  R = op(R, dpp_row_shl<8>(R));
  R = op(R, dpp_row_shl<4>(R));
  R = op(R, dpp_row_shl<2>(R));
  R = op(R, dpp_row_shl<1>(R));
  r[0] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R), 0));
  r[1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R), 32));
  r[2] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R), 16));
  r[3] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R), 48));
}

__device__ __forceinline__ bool cell_skipped(const FrameDev& F, int l, int bx, int y) {
  const int lev = l + 2;
  if (lev > 4) return false;
  // the lookup uses the half-row index bx, not the cell's x (src/bpextract.cu:209-228)
  int X = bx << lev, Y = y << lev;
  X = X > F.W - 1 ? F.W - 1 : X;
  Y = Y > F.H - 1 ? F.H - 1 : Y;
  return (int)gp((const uint8_t*)F.levels)[(size_t)Y * F.W + X] > lev;
}

// One half-sweep (src/bpextract.cu:200-312) for a batch. A wave owns one cell;
// the cell index and every address are wave-uniform (scalar registers), lanes
// index the inverse-depth samples (d = lane + 64 k). For each direction the
// outgoing message is min(h[d], h[d-1]+P1, h[d+1]+P1, min h + P2) minus the
// tree mean, h being the data term plus the incoming messages other than the
// target's, summed in the reference's order (up, down, left, right).
template <int N>
__global__ __launch_bounds__(256) void k_bp_sweep(const FrameDev* __restrict__ frames, BlockMap bm, int l, int A,
                                                  int first, int active_ok, float p1, float p2) {
  constexpr int K = N / 64;
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int lane = threadIdx.x & 63;
  const int item = (blockIdx.x - bm.start[f]) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
  const gptr<float> dst = gp(F.msg[l]) + ((size_t)y * w + x) * 4 * N;
  const gptr<const float> P = gp((const float*)((first && l < 3) ? F.msg[l + 1] : nullptr));
  const int wp = (l < 3) ? F.lw[l + 1] : 0;
  if (!active) {
    // first sweep only: copy the parent's four messages (zeros at the coarsest level)
    const gptr<float4> d4 = (gptr<float4>)dst;
    if (P) {
      const gptr<const float4> s4 = (gptr<const float4>)(P + ((size_t)(y >> 1) * wp + (x >> 1)) * 4 * N);
      float4 v[K];
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = s4[lane + 64 * k];
#pragma unroll
      for (int k = 0; k < K; ++k) d4[lane + 64 * k] = v[k];
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) d4[lane + 64 * k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return;
  }
  // incoming messages: from the up neighbour its "down" message, from the down
  // neighbour its "up" message, from the left its "right", from the right its
  // "left"; the first sweep of a level reads them from the parent cell
  const bool has_up = y > 0, has_down = y < h - 1, has_left = x > 0, has_right = x < w - 1;
  const gptr<const float> M = gp((const float*)F.msg[l]);
  auto src = [&](int nx, int ny, int dir) -> gptr<const float> {
    if (first) return P + (((size_t)(ny >> 1) * wp + (nx >> 1)) * 4 + dir) * N;
    return M + (((size_t)ny * w + nx) * 4 + dir) * N;
  };
  const bool msgs = !first || P;
  const gptr<const float> data = gp((const float*)F.cost[l]) + ((size_t)y * w + x) * N;
  const gptr<const float> su = src(x, y - 1, 1), sd = src(x, y + 1, 0), sl = src(x - 1, y, 3), sr = src(x + 1, y, 2);
  float dv[K], fu[K], fd[K], fl[K], fr[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int d = lane + 64 * k;
    dv[k] = data[d];
    fu[k] = (msgs && has_up) ? su[d] : 0.0f;
    fd[k] = (msgs && has_down) ? sd[d] : 0.0f;
    fl[k] = (msgs && has_left) ? sl[d] : 0.0f;
    fr[k] = (msgs && has_right) ? sr[d] : 0.0f;
  }
  float hv[4][K];
  if (has_up && has_down && has_left && has_right) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      hv[0][k] = ((dv[k] + fd[k]) + fl[k]) + fr[k];
      hv[1][k] = ((dv[k] + fu[k]) + fl[k]) + fr[k];
      hv[2][k] = ((dv[k] + fu[k]) + fd[k]) + fr[k];
      hv[3][k] = ((dv[k] + fu[k]) + fd[k]) + fl[k];
    }
  } else {
#pragma unroll
    for (int dir = 0; dir < 4; ++dir)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        float v = dv[k];
        if (dir != 0 && has_up) v += fu[k];
        if (dir != 1 && has_down) v += fd[k];
        if (dir != 2 && has_left) v += fl[k];
        if (dir != 3 && has_right) v += fr[k];
        hv[dir][k] = v;
      }
  }
  // min over d of each direction (for N = 128 the first stage, d with d + 64, is in-lane)
  float hmin[4];
  if constexpr (K == 1) {
    tree4(hv[0][0], hv[1][0], hv[2][0], hv[3][0], Min(), hmin);
  } else {
    tree4(vmin(hv[0][0], hv[0][1]), vmin(hv[1][0], hv[1][1]), vmin(hv[2][0], hv[2][1]), vmin(hv[3][0], hv[3][1]),
          Min(), hmin);
  }
  float m[4][K];
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    const float c2 = hmin[dir] + p2;
    if constexpr (K == 1) {
      const float v = hv[dir][0];
      const float hl = dpp_old<kWaveShr1>(v, v);  // lane 0: own value, h + P1 >= h never wins
      const float hr = dpp_old<kWaveShl1>(v, v);  // lane 63: likewise
      m[dir][0] = vmin(vmin(v, hl + p1), vmin(hr + p1, c2));
    } else {
      const float v0 = hv[dir][0], v1 = hv[dir][1];
      const float hl0 = dpp_old<kWaveShr1>(v0, v0);                          // d - 1, none at d = 0
      const float hl1 = dpp_old<kWaveShr1>(dpp_old<kWaveRor1>(v1, v0), v1);  // d = 64 takes d = 63
      const float hr0 = dpp_old<kWaveShl1>(dpp_old<kWaveRol1>(v0, v1), v0);  // d = 63 takes d = 64
      const float hr1 = dpp_old<kWaveShl1>(v1, v1);                          // none at d = 127
      m[dir][0] = vmin(vmin(v0, hl0 + p1), vmin(hr0 + p1, c2));
      m[dir][1] = vmin(vmin(v1, hl1 + p1), vmin(hr1 + p1, c2));
    }
  }
  float msum[4];
  if constexpr (K == 1) {
    tree4(m[0][0], m[1][0], m[2][0], m[3][0], Add(), msum);
  } else {
    tree4(m[0][0] + m[0][1], m[1][0] + m[1][1], m[2][0] + m[2][1], m[3][0] + m[3][1], Add(), msum);
  }
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    const float mean = msum[dir] / (float)N;
#pragma unroll
    for (int k = 0; k < K; ++k) dst[dir * N + lane + 64 * k] = m[dir][k] - mean;
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
  const int cell = (blockIdx.x - bm.start[f]) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = F.lw[0], h = F.lh[0];
  if (cell >= w * h) return;
  const int y = cell / w, x = cell - y * w;
  const int qs = 1 << gp((const uint8_t*)F.levels)[(size_t)(4 * y) * F.W + 4 * x];
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
