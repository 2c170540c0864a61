// k_bp.hip -- hierarchical min-sum belief propagation and winner extraction
// (reference: src/bpextract.cu:37-395) for a batch of keyframes.
//
// One 16-lane DPP row updates one cell, four cells per wave. Cost rows and
// message rows use the row layout of qm_device.h (sample j + 16 s at position
// S*j + s), so lane j holds S = N/16 samples of its cell and a row is one
// 16-lane float4 load. A cell's four outgoing messages are stored together,
// [cell][dir][N] with dir = up, down, left, right: one contiguous 1 KB
// (N = 64) per cell. The message upsample between levels is folded into the
// first half-sweep of each level: that sweep visits every cell, reads
// neighbour messages straight from the parent level, and copies the parent's
// messages into the cells it does not update that will need them: cells the
// quadtree skips at this level, and every cell when the level has fewer than
// two half-sweeps.
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

template <int S>
struct Row {
  float v[S];
};

template <int S>
__device__ __forceinline__ Row<S> load_row(gptr<const float> p, int j) {
  Row<S> r;
  const gptr<const float4> q = (gptr<const float4>)(p + S * j);
#pragma unroll
  for (int k = 0; k < S / 4; ++k) {
    const float4 t = q[k];
    r.v[4 * k] = t.x;
    r.v[4 * k + 1] = t.y;
    r.v[4 * k + 2] = t.z;
    r.v[4 * k + 3] = t.w;
  }
  return r;
}

template <int S>
__device__ __forceinline__ void store_row(gptr<float> p, int j, const Row<S>& r) {
  const gptr<float4> q = (gptr<float4>)(p + S * j);
#pragma unroll
  for (int k = 0; k < S / 4; ++k) q[k] = make_float4(r.v[4 * k], r.v[4 * k + 1], r.v[4 * k + 2], r.v[4 * k + 3]);
}

// The reference's shared-memory halving tree over a row (stage m: sample
// d < m combines sample d + m, m = N/2, ..., 1; src/bpextract.cu:274-302):
// stages N/2 .. 16 pair slot s with slot s + m/16 inside each lane, stages
// 8 .. 1 pair lane j with lane j + m of the row (DPP row_shl). Lane 0 of the
// row ends with the result, which row_newbcast hands to the whole row. Each
// element meets exactly the partner it meets in the reference, and both
// combines are commutative, so the result is bit-identical.
template <int S>
__device__ __forceinline__ float row_tree_min(Row<S> a) {
#pragma unroll
  for (int m = S / 2; m >= 1; m >>= 1)
#pragma unroll
    for (int s = 0; s < m; ++s) a.v[s] = vmin(a.v[s], a.v[s + m]);
  float t = a.v[0];
  t = vmin(t, dpp_row_shl<8>(t));
  t = vmin(t, dpp_row_shl<4>(t));
  t = vmin(t, dpp_row_shl<2>(t));
  t = vmin(t, dpp_row_shl<1>(t));
  return row_bcast0(t);
}

template <int S>
__device__ __forceinline__ float row_tree_sum(Row<S> a) {
#pragma unroll
  for (int m = S / 2; m >= 1; m >>= 1)
#pragma unroll
    for (int s = 0; s < m; ++s) a.v[s] = a.v[s] + a.v[s + m];
  float t = a.v[0];
  t = t + dpp_row_shl<8>(t);
  t = t + dpp_row_shl<4>(t);
  t = t + dpp_row_shl<2>(t);
  t = t + dpp_row_shl<1>(t);
  return row_bcast0(t);
}

// Outgoing message for one direction from h (the direction's belief without
// the target's message): min(h[d], h[d-1]+P1, h[d+1]+P1, min h + P2), minus the
// tree mean (src/bpextract.cu:272-302). Neighbour d - 1 of sample j + 16 s is
// lane j - 1 (row_shr:1), or for j = 0 lane 15 of slot s - 1 (row_ror:1);
// neighbour d + 1 is lane j + 1 (row_shl:1), or for j = 15 lane 0 of slot
// s + 1 (row_ror:15). Samples 0 and N - 1 lack one neighbour and see their own
// value there, and h + P1 > h never wins the min.
template <int S>
__device__ __forceinline__ Row<S> message_from(const Row<S>& h, float p1, float p2, float inv_n) {
  const float c2 = row_tree_min(h) + p2;
  Row<S> m;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const float lwrap = s > 0 ? dpp_mov<kRowRor1>(h.v[s - 1]) : h.v[s];
    const float rwrap = s < S - 1 ? dpp_mov<kRowRor15>(h.v[s + 1]) : h.v[s];
    const float hl = dpp_old<kRowShr1>(lwrap, h.v[s]);
    const float hr = dpp_old<kRowShl1>(rwrap, h.v[s]);
    m.v[s] = vmin(vmin(h.v[s], hl + p1), vmin(hr + p1, c2));
  }
  const float mean = row_tree_sum(m) * inv_n;  // sum / N, exact (N is a power of 2)
#pragma unroll
  for (int s = 0; s < S; ++s) m.v[s] = m.v[s] - mean;
  return m;
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

// Winner extraction of one cell (src/bpextract.cu:314-372), one 16-lane row:
// b = the cell's belief (data + the messages from its four neighbours, added
// in the reference's order). Argmin by the reference's tree (stage m: sample
// d < m takes sample d + m only if strictly smaller, so ties keep the lower
// index at every stage), parabola sub-sample refinement, depth = 1 / (step * d
// + min) in double. Cells that are not quadtree nodes get depth 0, label -1.
template <int N>
__device__ __forceinline__ bool is_node(const FrameDev& F, int x, int y) {
  const int qs = 1 << gp((const uint8_t*)F.levels)[(size_t)(4 * y) * F.W + 4 * x];
  return ((4 * x) % qs) == 0 && ((4 * y) % qs) == 0;
}

template <int N, int S>
__device__ __forceinline__ void extract_winner(const FrameDev& F, int cell, int j, int lane, const Row<S>& b,
                                               double imin, double istep, bool node = true) {
  // argmin tree: in-lane stages N/2 .. 16, then row_shl 8 .. 1
  float v[S];
  int id[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    v[s] = b.v[s];
    id[s] = j + 16 * s;
  }
#pragma unroll
  for (int m = S / 2; m >= 1; m >>= 1)
#pragma unroll
    for (int s = 0; s < m; ++s)
      if (v[s + m] < v[s]) {
        v[s] = v[s + m];
        id[s] = id[s + m];
      }
  float bv = v[0];
  int bi = id[0];
  {
    float ov;
    int oi;
    ov = dpp_row_shl<8>(bv); oi = dpp_row_shl_i<8>(bi); if (ov < bv) { bv = ov; bi = oi; }
    ov = dpp_row_shl<4>(bv); oi = dpp_row_shl_i<4>(bi); if (ov < bv) { bv = ov; bi = oi; }
    ov = dpp_row_shl<2>(bv); oi = dpp_row_shl_i<2>(bi); if (ov < bv) { bv = ov; bi = oi; }
    ov = dpp_row_shl<1>(bv); oi = dpp_row_shl_i<1>(bi); if (ov < bv) { bv = ov; bi = oi; }
  }
  bv = row_bcast0(bv);
  bi = row_bcast0_i(bi);
  // beliefs of the winner's neighbours (clamped to the range as the reference)
  const int dm1 = bi > 0 ? bi - 1 : 0, dp1 = bi < N - 1 ? bi + 1 : N - 1;
  const int rowbase = lane & 48;
  float pre = 0.0f, post = 0.0f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const float a = __shfl(b.v[s], rowbase + (dm1 & 15), 64);
    const float c = __shfl(b.v[s], rowbase + (dp1 & 15), 64);
    if ((dm1 >> 4) == s) pre = a;
    if ((dp1 >> 4) == s) post = c;
  }
  if (j == 0) {
    float disp = (float)bi;
    if (bi > 0 && bi < N - 1) {
      const float a = (pre - 2.0f * bv) + post;
      const float bb = (-pre) + post;
      const float ba = bb / a;
      if (isfinite(ba)) disp = (float)bi - ba / 2.0f;
    }
    // a cell that is not a node gets 0 and label -1
    const float depth = (float)(1.0 / (istep * (double)disp + imin));
    gp(F.dcoarse)[cell] = node ? depth : 0.0f;
    gp(F.label)[cell] = node ? bi : -1;
  }
}

// One half-sweep (src/bpextract.cu:200-312) for a batch: ROWS work items per
// block, one per 16-lane row. h for direction dir is the data term plus the
// incoming messages other than the target's, added in the reference's order
// (up, down, left, right).
template <int N, int ROWS>
__global__ __launch_bounds__(16 * ROWS) void k_bp_sweep(const FrameDev* __restrict__ frames, BlockMap bm, int l,
                                                        int A, int first, int active_ok, int copy_all, float p1,
                                                        float p2, int ext, double imin, double istep) {
  constexpr int S = N / 16;
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int j = threadIdx.x & 15;
  const int item = (blockIdx.x - bm.start[f]) * ROWS + (threadIdx.x >> 4);
  const int w = F.lw[l], h = F.lh[l];
  int x = 0, y = 0;
  bool upd = false, copy = false, ex = false;
  if (first) {
    if (item < w * h) {
      y = item / w;
      x = item - y * w;
      const bool mine = A ? ((x & 1) == (y & 1)) : ((x & 1) != (y & 1));
      const bool skipped = cell_skipped(F, l, x >> 1, y);
      upd = active_ok && mine && !skipped;
      // A cell of the other colour that the level's second half-sweep will
      // update needs no copy: only this sweep reads its messages, and it reads
      // them from the parent. Skipped cells keep the parent's messages.
      copy = !upd && (copy_all || skipped);
    }
  } else {
    const int hw = (w + 1) >> 1;
    if (item < hw * h) {
      y = item / hw;
      const int bx = item - y * hw;
      x = 2 * bx + (A ? (y & 1) : ((y + 1) & 1));
      upd = x < w && !cell_skipped(F, l, bx, y);
      // the level-0 sweep that ends BP also extracts the winners of its
      // colour's cells: their incoming messages are final
      ex = ext && x < w;
    }
  }
  const gptr<float> dst = gp(F.msg[l]) + ((size_t)y * w + x) * 4 * N;
  const gptr<const float> P = gp((const float*)((first && l < 3) ? F.msg[l + 1] : nullptr));
  const int wp = (l < 3) ? F.lw[l + 1] : 0;
  if (copy) {
    // first sweep only: the parent's four messages (zeros at the coarsest level)
    Row<S> r[4];
    if (P) {
      const gptr<const float> src = P + ((size_t)(y >> 1) * wp + (x >> 1)) * 4 * N;
#pragma unroll
      for (int dir = 0; dir < 4; ++dir) r[dir] = load_row<S>(src + dir * N, j);
    } else {
#pragma unroll
      for (int dir = 0; dir < 4; ++dir)
#pragma unroll
        for (int s = 0; s < S; ++s) r[dir].v[s] = 0.0f;
    }
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) store_row<S>(dst + dir * N, j, r[dir]);
  }
  if (ex && !is_node<N>(F, x, y)) {
    if (j == 0) {
      gp(F.dcoarse)[(size_t)y * w + x] = 0.0f;
      gp(F.label)[(size_t)y * w + x] = -1;
    }
    ex = false;
  }
  if (!upd && !ex) return;
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
  const Row<S> dv = load_row<S>(gp((const float*)F.cost[l]) + ((size_t)y * w + x) * N, j);
  Row<S> fu, fd, fl, fr;
#pragma unroll
  for (int s = 0; s < S; ++s) fu.v[s] = fd.v[s] = fl.v[s] = fr.v[s] = 0.0f;
  if (msgs && has_up) fu = load_row<S>(src(x, y - 1, 1), j);
  if (msgs && has_down) fd = load_row<S>(src(x, y + 1, 0), j);
  if (msgs && has_left) fl = load_row<S>(src(x - 1, y, 3), j);
  if (msgs && has_right) fr = load_row<S>(src(x + 1, y, 2), j);
  if (ex) {
    // belief in the order of src/bpextract.cu:330-337: data, left, right, up, down
    Row<S> b = dv;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (has_left) b.v[s] += fl.v[s];
      if (has_right) b.v[s] += fr.v[s];
      if (has_up) b.v[s] += fu.v[s];
      if (has_down) b.v[s] += fd.v[s];
    }
    extract_winner<N, S>(F, y * w + x, j, threadIdx.x & 63, b, imin, istep);
  }
  if (!upd) return;
  Row<S> hv[4];
  if (__all(has_up && has_down && has_left && has_right)) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      hv[0].v[s] = ((dv.v[s] + fd.v[s]) + fl.v[s]) + fr.v[s];
      hv[1].v[s] = ((dv.v[s] + fu.v[s]) + fl.v[s]) + fr.v[s];
      hv[2].v[s] = ((dv.v[s] + fu.v[s]) + fd.v[s]) + fr.v[s];
      hv[3].v[s] = ((dv.v[s] + fu.v[s]) + fd.v[s]) + fl.v[s];
    }
  } else {
#pragma unroll
    for (int dir = 0; dir < 4; ++dir)
#pragma unroll
      for (int s = 0; s < S; ++s) {
        float v = dv.v[s];
        if (dir != 0 && has_up) v += fu.v[s];
        if (dir != 1 && has_down) v += fd.v[s];
        if (dir != 2 && has_left) v += fl.v[s];
        if (dir != 3 && has_right) v += fr.v[s];
        hv[dir].v[s] = v;
      }
  }
  constexpr float inv_n = 1.0f / (float)N;
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) store_row<S>(dst + dir * N, j, message_from<S>(hv[dir], p1, p2, inv_n));
}

// Winner extraction over a batch, one 16-lane row per cell (extract_winner).
// colour < 0: every cell; else only the cells of checkerboard colour colour
// (0: x + y even, 1: odd), the others having been extracted by the last sweep.
template <int N, int ROWS>
__global__ __launch_bounds__(16 * ROWS) void k_extract(const FrameDev* __restrict__ frames, BlockMap bm, double imin,
                                                       double istep, int colour) {
  constexpr int S = N / 16;
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int lane = threadIdx.x & 63, j = lane & 15;
  const int item = (blockIdx.x - bm.start[f]) * ROWS + (threadIdx.x >> 4);
  const int w = F.lw[0], h = F.lh[0];
  int x, y;
  if (colour < 0) {
    if (item >= w * h) return;
    y = item / w;
    x = item - y * w;
  } else {
    const int hw = (w + 1) >> 1;
    if (item >= hw * h) return;
    y = item / hw;
    const int bx = item - y * hw;
    x = 2 * bx + (((y & 1) ^ colour) & 1);
    if (x >= w) return;
  }
  const int cell = y * w + x;
  // the quadtree level and all five rows are loaded at once (a missing
  // neighbour's row is replaced by a valid one and not added; a cell that is
  // not a node runs the extraction too and stores 0 and -1), so their
  // latencies overlap
  const int qs = 1 << gp((const uint8_t*)F.levels)[(size_t)(4 * y) * F.W + 4 * x];
  const gptr<const float> M = gp((const float*)F.msg[0]);
  const bool hl = x != 0, hr = x != w - 1, hu = y != 0, hd = y != h - 1;
  Row<S> b = load_row<S>(gp((const float*)F.cost[0]) + (size_t)cell * N, j);
  const Row<S> tl = load_row<S>(M + (((size_t)(hl ? cell - 1 : cell)) * 4 + 3) * N, j);
  const Row<S> tr = load_row<S>(M + (((size_t)(hr ? cell + 1 : cell)) * 4 + 2) * N, j);
  const Row<S> tu = load_row<S>(M + (((size_t)(hu ? cell - w : cell)) * 4 + 1) * N, j);
  const Row<S> td = load_row<S>(M + (((size_t)(hd ? cell + w : cell)) * 4 + 0) * N, j);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    b.v[s] = hl ? b.v[s] + tl.v[s] : b.v[s];
    b.v[s] = hr ? b.v[s] + tr.v[s] : b.v[s];
    b.v[s] = hu ? b.v[s] + tu.v[s] : b.v[s];
    b.v[s] = hd ? b.v[s] + td.v[s] : b.v[s];
  }
  const bool node = ((4 * x) % qs) == 0 && ((4 * y) % qs) == 0;  // is_node
  extract_winner<N, S>(F, cell, j, lane, b, imin, istep, node);
}

// [cell][dir][row layout] -> [dir][cell][N] in sample order, for callers that
// want the reference's message maps
template <int N>
__global__ __launch_bounds__(256) void k_copy_messages(const FrameDev* __restrict__ frames, float* out) {
  const FrameDev& F = frames[0];
  const int ncell = F.lw[0] * F.lh[0];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)ncell * 4 * N) return;
  const int p = (int)(i % N);
  const size_t t = i / N;
  const int dir = (int)(t % 4);
  const size_t cell = t / 4;
  out[((size_t)dir * ncell + cell) * N + row_sample<N>(p)] = F.msg[0][i];
}

// caller cost volume [cell][N] in sample order -> row layout
template <int N>
__global__ __launch_bounds__(256) void k_cost_to_rows(const float* __restrict__ in, float* __restrict__ out,
                                                      size_t ncell) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= ncell * N) return;
  const int d = (int)(i % N);
  out[i - d + row_pos<N>(d)] = in[i];
}

template <int N, int ROWS>
static void bp_sweep_shape(int nb, hipStream_t s, const FrameDev* dframes, const BlockMap& bm, int l, bool A,
                           bool first, bool active, bool copy_all, bool ext, const BpLaunchParams& p) {
  hipLaunchKernelGGL((k_bp_sweep<N, ROWS>), dim3(nb), dim3(16 * ROWS), 0, s, dframes, bm, l, (int)A, (int)first,
                     (int)active, (int)copy_all, p.p1, p.p2, (int)ext, p.idepth_min, p.idepth_step);
}

void launch_bp_sweep(const FrameDev* dframes, const FrameDev* hframes, int n, int N, int l, bool A, bool first,
                     bool active, const BpLaunchParams& p, const Tuning& t, hipStream_t s, bool copy_all,
                     bool ext) {
  const int rows = (t.bp_rows == 4 || t.bp_rows == 8) ? t.bp_rows : 16;
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) {
    const int w = hframes[i].lw[l], h = hframes[i].lh[l];
    const int items = first ? w * h : ((w + 1) / 2) * h;
    counts[i] = (items + rows - 1) / rows;
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 64) {
    if (rows == 4) bp_sweep_shape<64, 4>(nb, s, dframes, bm, l, A, first, active, copy_all, ext && l == 0 && !first, p);
    else if (rows == 8) bp_sweep_shape<64, 8>(nb, s, dframes, bm, l, A, first, active, copy_all, ext && l == 0 && !first, p);
    else bp_sweep_shape<64, 16>(nb, s, dframes, bm, l, A, first, active, copy_all, ext && l == 0 && !first, p);
  } else {
    if (rows == 4) bp_sweep_shape<128, 4>(nb, s, dframes, bm, l, A, first, active, copy_all, ext && l == 0 && !first, p);
    else if (rows == 8) bp_sweep_shape<128, 8>(nb, s, dframes, bm, l, A, first, active, copy_all, ext && l == 0 && !first, p);
    else bp_sweep_shape<128, 16>(nb, s, dframes, bm, l, A, first, active, copy_all, ext && l == 0 && !first, p);
  }
}

template <int N, int ROWS>
static void extract_shape(int nb, hipStream_t s, const FrameDev* dframes, const BlockMap& bm,
                          const BpLaunchParams& p, int colour) {
  hipLaunchKernelGGL((k_extract<N, ROWS>), dim3(nb), dim3(16 * ROWS), 0, s, dframes, bm, p.idepth_min,
                     p.idepth_step, colour);
}

void launch_extract(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const BpLaunchParams& p,
                    const Tuning& t, hipStream_t s, int colour) {
  const int rows = (t.extract_rows == 4 || t.extract_rows == 8) ? t.extract_rows : 16;
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) {
    const int w = hframes[i].lw[0], h = hframes[i].lh[0];
    const int items = colour < 0 ? w * h : ((w + 1) / 2) * h;
    counts[i] = (items + rows - 1) / rows;
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 64) {
    if (rows == 4) extract_shape<64, 4>(nb, s, dframes, bm, p, colour);
    else if (rows == 8) extract_shape<64, 8>(nb, s, dframes, bm, p, colour);
    else extract_shape<64, 16>(nb, s, dframes, bm, p, colour);
  } else {
    if (rows == 4) extract_shape<128, 4>(nb, s, dframes, bm, p, colour);
    else if (rows == 8) extract_shape<128, 8>(nb, s, dframes, bm, p, colour);
    else extract_shape<128, 16>(nb, s, dframes, bm, p, colour);
  }
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

void launch_cost_to_rows(const float* in, float* out, size_t ncell, int N, hipStream_t s) {
  const int nb = (int)((ncell * N + 255) / 256);
  if (nb <= 0) return;
  if (N == 64)
    hipLaunchKernelGGL(k_cost_to_rows<64>, dim3(nb), dim3(256), 0, s, in, out, ncell);
  else
    hipLaunchKernelGGL(k_cost_to_rows<128>, dim3(nb), dim3(256), 0, s, in, out, ncell);
}

}  // namespace qm
