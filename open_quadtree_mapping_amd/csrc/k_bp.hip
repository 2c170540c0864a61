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
// two half-sweeps. Consecutive half-sweeps of a level run in pairs, one launch
// each (k_bp_pair): the first colour's messages pass to the second through LDS.
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

// Non-temporal forms, for rows a kernel touches once: the accesses carry the
// streaming hint (the caches keep such lines the shortest), which lets a
// gather of scattered 256 B rows run near the streaming rate
template <int S>
__device__ __forceinline__ Row<S> load_row_nt(gptr<const float> p, int j) {
  Row<S> r;
  const gptr<const F4v> q = (gptr<const F4v>)(p + S * j);
#pragma unroll
  for (int k = 0; k < S / 4; ++k) {
    const F4v t = __builtin_nontemporal_load(q + k);
    r.v[4 * k] = t.x;
    r.v[4 * k + 1] = t.y;
    r.v[4 * k + 2] = t.z;
    r.v[4 * k + 3] = t.w;
  }
  return r;
}

template <int S>
__device__ __forceinline__ void store_row_nt(gptr<float> p, int j, const Row<S>& r) {
  const gptr<F4v> q = (gptr<F4v>)(p + S * j);
#pragma unroll
  for (int k = 0; k < S / 4; ++k) {
    F4v t = {r.v[4 * k], r.v[4 * k + 1], r.v[4 * k + 2], r.v[4 * k + 3]};
    __builtin_nontemporal_store(t, q + k);
  }
}

// message buffer b (0: msg, 1: msg2) of level l
__device__ __forceinline__ float* msg_buf(const FrameDev& F, int l, int b) { return b ? F.msg2[l] : F.msg[l]; }

// The reference's shared-memory halving tree over a row (stage m: sample
// d < m combines sample d + m, m = N/2, ..., 1; src/bpextract.cu:274-302):
// stages N/2 .. 16 pair slot s with slot s + m/16 inside each lane, stages
// 8 .. 1 pair lane j with lane j + m of the row (DPP row_shl). Lane 0 of the
// row ends with the result, which row_newbcast hands to the whole row. Each
// element meets exactly the partner it meets in the reference, and both
// combines are commutative, so the result is bit-identical.
template <int S>
__device__ __forceinline__ float row_tree_min(Row<S> a) {
  // the minimum of a set is one of its values whatever the order: the lane's
  // S values fold by three-operand minima
  float t = a.v[0];
#pragma unroll
  for (int s = 1; s < S; s += 2) t = s + 1 < S ? vmin(t, vmin(a.v[s], a.v[s + 1])) : vmin(t, a.v[s]);
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
// lane j - 1 of slot s, or for j = 0 lane 15 of slot s - 1: one row_ror:1 of a
// per-lane choice (lane 15 offers slot s - 1, the others slot s), which the
// compiler folds into the add of P1; neighbour d + 1 likewise by row_ror:15
// (lane 0 offers slot s + 1). Samples 0 and N - 1 lack one neighbour and see
// their own value there (row_shr:1 / row_shl:1 keeping the lane's own), and
// h + P1 > h never wins the min. The minimum is exact, so its operand order
// does not change the result.
template <int S>
__device__ __forceinline__ Row<S> message_from(const Row<S>& h, float p1, float p2, float inv_n) {
  const float c2 = row_tree_min(h) + p2;
  const int j = __lane_id() & 15;
  // P1 in a vector register: a VOP2 add with a DPP operand takes its other
  // operand from a VGPR
  float p1v = p1;
  asm volatile("" : "+v"(p1v));
  Row<S> m;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const float hl = s > 0 ? dpp_mov<kRowRor1>(j == 15 ? h.v[s - 1] : h.v[s]) : dpp_old<kRowShr1>(h.v[0], h.v[0]);
    const float hr =
        s < S - 1 ? dpp_mov<kRowRor15>(j == 0 ? h.v[s + 1] : h.v[s]) : dpp_old<kRowShl1>(h.v[s], h.v[s]);
    m.v[s] = vmin(vmin(hl + p1v, hr + p1v), vmin(h.v[s], c2));
  }
  const float mean = row_tree_sum(m) * inv_n;  // sum / N, exact (N is a power of 2)
#pragma unroll
  for (int s = 0; s < S; ++s) m.v[s] = m.v[s] - mean;
  return m;
}

// h of the four directions (the data term plus the incoming messages other
// than the target's, each added only where its neighbour exists), in the
// reference's order up, down, left, right
template <int S>
__device__ __forceinline__ void directional_h(const Row<S>& dv, const Row<S>& fu, const Row<S>& fd, const Row<S>& fl,
                                              const Row<S>& fr, bool has_up, bool has_down, bool has_left,
                                              bool has_right, Row<S> (&hv)[4]) {
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
}

// winner extraction's belief in the order of src/bpextract.cu:330-337: data,
// left, right, up, down
template <int S>
__device__ __forceinline__ Row<S> belief(const Row<S>& dv, const Row<S>& fu, const Row<S>& fd, const Row<S>& fl,
                                         const Row<S>& fr, bool has_up, bool has_down, bool has_left,
                                         bool has_right) {
  Row<S> b = dv;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (has_left) b.v[s] += fl.v[s];
    if (has_right) b.v[s] += fr.v[s];
    if (has_up) b.v[s] += fu.v[s];
    if (has_down) b.v[s] += fd.v[s];
  }
  return b;
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

// Every lane of the row computes the refinement and the depth, without a
// branch, so that a caller can interleave this chain with other arithmetic.
template <int N, int S>
__device__ __forceinline__ void winner_of(const Row<S>& b, int j, int lane, double imin, double istep, float& depth,
                                          int& label) {
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
  // bi, dm1 and dp1 are the same in the row's 16 lanes: each lane picks its
  // sample of dm1's (dp1's) slot, and one permute reads lane dm1 & 15's
  const int rowbase = lane & 48;
  float am = b.v[0], ap = b.v[0];
#pragma unroll
  for (int s = 1; s < S; ++s) {
    am = (dm1 >> 4) == s ? b.v[s] : am;
    ap = (dp1 >> 4) == s ? b.v[s] : ap;
  }
  const float pre = __shfl(am, rowbase + (dm1 & 15), 64);
  const float post = __shfl(ap, rowbase + (dp1 & 15), 64);
  // parabola refinement, used only for an interior winner with a finite step
  const float a = (pre - 2.0f * bv) + post;
  const float bb = (-pre) + post;
  const float ba = bb / a;
  const float disp = (bi > 0 && bi < N - 1 && isfinite(ba)) ? (float)bi - ba / 2.0f : (float)bi;
  depth = (float)(1.0 / (istep * (double)disp + imin));
  label = bi;
}

template <int N, int S>
__device__ __forceinline__ void extract_winner(const FrameDev& F, int cell, int j, int lane, const Row<S>& b,
                                               double imin, double istep, bool node = true) {
  float depth;
  int label;
  winner_of<N, S>(b, j, lane, imin, istep, depth, label);
  if (j == 0) {
    // a cell that is not a node gets 0 and label -1
    gp(F.dcoarse)[cell] = node ? depth : 0.0f;
    gp(F.label)[cell] = node ? label : -1;
  }
}

// One half-sweep (src/bpextract.cu:200-312) for a batch: ROWS work items per
// block, one per 16-lane row. h for direction dir is the data term plus the
// incoming messages other than the target's, added in the reference's order
// (up, down, left, right).
template <int N, int ROWS>
__global__ __launch_bounds__(16 * ROWS) void k_bp_sweep(const FrameDev* __restrict__ frames, BlockMap bm, int l,
                                                        int A, int first, int active_ok, int copy_all, float p1,
                                                        float p2, int ext, double imin, double istep, int buf,
                                                        int pbuf) {
  constexpr int S = N / 16;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int f = find_frame(bm, lb);
  const FrameDev& F = frames[f];
  const int j = threadIdx.x & 15;
  const int item = (lb - bm.start[f]) * ROWS + (threadIdx.x >> 4);
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
  const gptr<float> dst = gp(msg_buf(F, l, buf)) + ((size_t)y * w + x) * 4 * N;
  const gptr<const float> P = gp((const float*)((first && l < 3) ? msg_buf(F, l + 1, pbuf) : nullptr));
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
  const gptr<const float> M = gp((const float*)msg_buf(F, l, buf));
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
  if (ex)
    extract_winner<N, S>(F, y * w + x, j, threadIdx.x & 63,
                         belief<S>(dv, fu, fd, fl, fr, has_up, has_down, has_left, has_right), imin, istep);
  if (!upd) return;
  Row<S> hv[4];
  directional_h<S>(dv, fu, fd, fl, fr, has_up, has_down, has_left, has_right, hv);
  constexpr float inv_n = 1.0f / (float)N;
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) store_row<S>(dst + dir * N, j, message_from<S>(hv[dir], p1, p2, inv_n));
}

// ---------------------------------------------------------------------------
// Two consecutive half-sweeps of one level in one launch. The first sweep
// updates the cells of parity ca = (x + y) & 1 (the "A" cells), the second the
// others ("B" cells), which read only A cells' messages. A block owns a strip
// of TW = 2 (GA - 1) columns and seg rows and walks down its rows: in step t
// its A groups update A row y0 - 1 + t (the strip's A cells and the one A cell
// just outside the strip that borders a B cell of the row) into LDS, while its
// B groups update B row y0 + t - 3 from the A rows above, at and below it
// (written in steps t - 3, t - 2, t - 1). Each message direction has a ring
// only as long as its read trails its write (2, 3 and 4 rows), so a block
// holds 24 KB of LDS at N = 64 and five blocks fit a CU. An A cell's messages go to HBM only
// when a later launch reads them (write_first: the level's last pair); the B
// cells' messages go to the other message buffer (wr), so that no block reads
// a message another block of the launch rewrites. A cells on the halo (one per
// row, and the rows y0 - 1 and y0 + seg) are updated by every block that
// borders them, from the same inputs with the same arithmetic, so each message
// equals the one of two separate half-sweeps, bit for bit. Cells the quadtree
// skips keep their messages: both buffers hold them. SRC > 0: the level's
// first two sweeps: A cells read their neighbours' messages from the parent
// level (SRC 1, buffer pbuf; SRC 2 at the coarsest level: zeros), and skipped
// cells of both colours get the parent's messages in both buffers. SRC 0:
// later sweeps, A cells read buffer rd. EXT: the second sweep ends BP (level 0)
// and also extracts the winners of its B cells, computed beside their messages.
// ---------------------------------------------------------------------------
constexpr int kPairMaxSeg = 64;

template <int N, int GA, int SRC, bool EXT>
__global__ __launch_bounds__(32 * GA) void k_bp_pair(const FrameDev* __restrict__ frames, BlockMap bm, int l, int ca,
                                                     int rd, int wr, int pbuf, int write_first, int seg, float p1,
                                                     float p2, double imin, double istep) {
  constexpr bool first = SRC > 0;
  constexpr int S = N / 16, Q = S / 4;  // floats and float4 per lane of one row
  constexpr int TW = 2 * (GA - 1);
  // the A rows' messages, one ring per direction, each as long as its reads
  // trail its writes: the down message (dir 0) is read one step after the A
  // row is updated, the left and right ones (dirs 2, 3) two steps after, the
  // up message (dir 1) three
  __shared__ float4 L0[2][GA][16 * Q];
  __shared__ float4 L23[3][GA][2][16 * Q];
  __shared__ float4 L1[4][GA][16 * Q];
  // per cell of rows y0 - 1 .. y0 + R, columns x0 - 1 .. x0 + TW: bit 0 the
  // quadtree skips it, bit 1 it is a quadtree node (level 0 only)
  __shared__ uint8_t Sk[kPairMaxSeg + 2][TW + 2];
  // neighbouring strips and segments share halo cells: keep them on one XCD's L2
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int f = find_frame(bm, lb);
  const FrameDev& F = frames[f];
  const int w = F.lw[l], h = F.lh[l];
  const int ns = (w + TW - 1) / TW;
  const int bi = lb - bm.start[f];
  const int x0 = (bi % ns) * TW, y0 = (bi / ns) * seg;
  const int R = min(seg, h - y0);
  const int T = 2 * ((R + 4) / 2);  // R + 3 steps, rounded up to a multiple of 2
  const int grp = threadIdx.x >> 4, j = threadIdx.x & 15;
  for (int i = threadIdx.x; i < (R + 2) * (TW + 2); i += 32 * GA) {
    const int r = i / (TW + 2), c = i - r * (TW + 2);
    const int y = y0 - 1 + r, x = x0 - 1 + c;
    const bool in = x >= 0 && x < w && y >= 0 && y < h;
    Sk[r][c] = (uint8_t)((in && cell_skipped(F, l, x >> 1, y) ? 1 : 0) | (in && l == 0 && is_node<N>(F, x, y) ? 2 : 0));
  }
  __syncthreads();
  auto skipped = [&](int x, int y) __attribute__((always_inline)) { return (Sk[y - y0 + 1][x - x0 + 1] & 1) != 0; };
  auto node = [&](int x, int y) __attribute__((always_inline)) { return (Sk[y - y0 + 1][x - x0 + 1] & 2) != 0; };
  const gptr<const float> cost = gp((const float*)F.cost[l]);
  const gptr<const float> Mrd = gp((const float*)msg_buf(F, l, rd));
  const gptr<float> Mwr = gp(msg_buf(F, l, wr));
  const gptr<float> M0 = gp(F.msg[l]), M1 = gp(F.msg2[l]);
  const gptr<const float> P = gp((const float*)(SRC == 1 ? msg_buf(F, l + 1, pbuf) : F.msg[l]));
  const gptr<const float> Min = SRC == 1 ? P : Mrd;  // where the A cells' incoming messages are read
  const int wp = l < 3 ? F.lw[l + 1] : 1;
  constexpr float inv_n = 1.0f / (float)N;
  auto cell_msgs = [&](int x, int y) __attribute__((always_inline)) { return ((size_t)y * w + x) * 4 * N; };
  auto parent_msgs = [&](int x, int y) __attribute__((always_inline)) {
    return ((size_t)(y >> 1) * wp + (x >> 1)) * 4 * N;
  };
  // a pair's stored messages are read only by a later launch: streaming stores
  // (its loads stay plain, since neighbouring blocks re-read halo rows)
  auto store4 = [&](gptr<float> dst, const Row<S> (&m)[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) store_row_nt<S>(dst + dir * N, j, m[dir]);
  };
  auto zero = [&](Row<S>& r) __attribute__((always_inline)) {
#pragma unroll
    for (int s = 0; s < S; ++s) r.v[s] = 0.0f;
  };

  if (grp < GA) {
    // ---------------- A groups: A cell x0 - sh + 2 g of row y0 - 1 + t ----------------
    const int g = grp;
    auto cell_at = [&](int u, int& x, int& y) __attribute__((always_inline)) {
      y = y0 - 1 + u;
      x = x0 - ((((x0 + y) & 1) != ca) ? 1 : 0) + 2 * g;
      return u <= R + 1 && y >= 0 && y < h && x >= 0 && x < w;
    };
    struct In {
      Row<S> dv, m[4];  // data row; incoming messages (up, down, left, right), or the cell's own when skipped
    };
    // step u's inputs, loaded two steps ahead: every load unconditional (an
    // invalid cell or a missing neighbour reads a valid row, not used)
    auto load_in = [&](int u, In& in) __attribute__((always_inline)) {
      int x, y;
      const bool ok = cell_at(u, x, y);
      if (!ok) x = y = 0;
      const bool sk = ok && skipped(x, y);
      in.dv = load_row<S>(cost + ((size_t)y * w + x) * N, j);
      // neighbour and direction of incoming message d: the up neighbour's
      // down message, the down neighbour's up, the left's right, the right's left
      const int nx[4] = {x, x, x > 0 ? x - 1 : x, x < w - 1 ? x + 1 : x};
      const int ny[4] = {y > 0 ? y - 1 : y, y < h - 1 ? y + 1 : y, y, y};
      const int md[4] = {1, 0, 3, 2};
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        if (SRC == 2) {
          zero(in.m[d]);
        } else {
          const int cx = sk ? x : nx[d], cy = sk ? y : ny[d], dd = sk ? d : md[d];
          in.m[d] = load_row<S>(Min + (SRC == 1 ? parent_msgs(cx, cy) : cell_msgs(cx, cy)) + dd * N, j);
        }
      }
    };
    auto step = [&](int u, const In& in) __attribute__((always_inline)) {
      int x, y;
      if (!cell_at(u, x, y)) return;
      const bool interior = x >= x0 && x < x0 + TW && y >= y0 && y < y0 + R;
      Row<S> out[4];
      if (!skipped(x, y)) {
        Row<S> hv[4];
        directional_h<S>(in.dv, in.m[0], in.m[1], in.m[2], in.m[3], y > 0, y < h - 1, x > 0, x < w - 1, hv);
#pragma unroll
        for (int dir = 0; dir < 4; ++dir) out[dir] = message_from<S>(hv[dir], p1, p2, inv_n);
        if (interior && write_first) store4(Mwr + cell_msgs(x, y), out);
      } else {
#pragma unroll
        for (int d = 0; d < 4; ++d) out[d] = in.m[d];
        if (first && interior) {
          store4(M0 + cell_msgs(x, y), out);
          store4(M1 + cell_msgs(x, y), out);
        }
      }
      float4* dst[4] = {&L0[u & 1][g][0], &L1[u & 3][g][0], &L23[u % 3][g][0][0], &L23[u % 3][g][1][0]};
#pragma unroll
      for (int dir = 0; dir < 4; ++dir)
#pragma unroll
        for (int q = 0; q < Q; ++q)
          dst[dir][j * Q + q] =
              make_float4(out[dir].v[4 * q], out[dir].v[4 * q + 1], out[dir].v[4 * q + 2], out[dir].v[4 * q + 3]);
    };
    // inputs are loaded a step ahead into two register sets, which take
    // turns: the loop is unrolled by two, without conditions, so that no set
    // is copied (the steps past R + 2 have no cells)
    In i0, i1;
    load_in(0, i0);
    for (int t = 0; t < T; t += 2) {
      load_in(t + 1, i1);
      step(t, i0);
      __syncthreads();
      load_in(t + 2, i0);
      step(t + 1, i1);
      __syncthreads();
    }
  } else {
    // ---------------- B groups: B cell x0 + sh + 2 gb of row y0 + t - 3 ----------------
    const int gb = grp - GA;
    auto cell_at = [&](int u, int& x, int& y) __attribute__((always_inline)) {
      y = y0 + u - 3;
      x = x0 + ((((x0 + y) & 1) == ca) ? 1 : 0) + 2 * gb;
      return u >= 3 && y < y0 + R && x < w && gb < GA - 1;
    };
    auto load_dv = [&](int u, Row<S>& r) __attribute__((always_inline)) {
      int x, y;
      const bool ok = cell_at(u, x, y);
      r = load_row<S>(cost + (ok ? (size_t)y * w + x : (size_t)0) * N, j);
    };
    auto lds_row = [&](const float4* row) __attribute__((always_inline)) {
      Row<S> r;
      const float4* src = row + j * Q;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const float4 v = src[q];
        r.v[4 * q] = v.x;
        r.v[4 * q + 1] = v.y;
        r.v[4 * q + 2] = v.z;
        r.v[4 * q + 3] = v.w;
      }
      return r;
    };
    auto step = [&](int u, const Row<S>& dv) __attribute__((always_inline)) {
      int x, y;
      if (!cell_at(u, x, y)) return;
      const bool sk = skipped(x, y);
      if (first && sk) {
        // the level's first sweep gives a skipped cell its parent's messages
        Row<S> m[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          if (SRC == 1) m[d] = load_row<S>(P + parent_msgs(x, y) + d * N, j);
          else zero(m[d]);
        }
        store4(M0 + cell_msgs(x, y), m);
        store4(M1 + cell_msgs(x, y), m);
      }
      // a cell that is not a quadtree node gets depth 0 and label -1
      const bool ex = EXT && node(x, y);
      if (sk && !ex) {
        if (EXT && j == 0) {
          gp(F.dcoarse)[(size_t)y * w + x] = 0.0f;
          gp(F.label)[(size_t)y * w + x] = -1;
        }
        return;
      }
      const bool has_up = y > 0, has_down = y < h - 1, has_left = x > 0, has_right = x < w - 1;
      const int gv = gb + ((((x0 + y) & 1) == ca) ? 1 : 0);  // the A group of the cells above and below
      Row<S> fu, fd, fl, fr;
      zero(fu);
      zero(fd);
      zero(fl);
      zero(fr);
      if (has_up) fu = lds_row(&L1[(u - 3) & 3][gv][0]);
      if (has_down) fd = lds_row(&L0[(u - 1) & 1][gv][0]);
      if (has_left) fl = lds_row(&L23[(u - 2) % 3][gb][1][0]);
      if (has_right) fr = lds_row(&L23[(u - 2) % 3][gb + 1][0][0]);
      // the messages and the winner in one branch-free stretch (a skipped
      // node computes messages it does not store), the stores after it
      Row<S> hv[4], out[4];
      directional_h<S>(dv, fu, fd, fl, fr, has_up, has_down, has_left, has_right, hv);
#pragma unroll
      for (int dir = 0; dir < 4; ++dir) out[dir] = message_from<S>(hv[dir], p1, p2, inv_n);
      float depth = 0.0f;
      int label = -1;
      if (EXT)
        winner_of<N, S>(belief<S>(dv, fu, fd, fl, fr, has_up, has_down, has_left, has_right), j,
                        threadIdx.x & 63, imin, istep, depth, label);
      if (!sk) store4(Mwr + cell_msgs(x, y), out);
      if (EXT && j == 0) {
        gp(F.dcoarse)[(size_t)y * w + x] = ex ? depth : 0.0f;
        gp(F.label)[(size_t)y * w + x] = ex ? label : -1;
      }
    };
    Row<S> d0, d1;
    load_dv(0, d0);
    for (int t = 0; t < T; t += 2) {
      load_dv(t + 1, d1);
      step(t, d0);
      __syncthreads();
      load_dv(t + 2, d0);
      step(t + 1, d1);
      __syncthreads();
    }
  }
}

// Winner extraction over a batch, one 16-lane row per cell (extract_winner).
// colour < 0: every cell; else only the cells of checkerboard colour colour
// (0: x + y even, 1: odd), the others having been extracted by the last sweep.
template <int N, int ROWS>
__global__ __launch_bounds__(16 * ROWS) void k_extract(const FrameDev* __restrict__ frames, BlockMap bm, double imin,
                                                       double istep, int colour, int buf) {
  constexpr int S = N / 16;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int f = find_frame(bm, lb);
  const FrameDev& F = frames[f];
  const int lane = threadIdx.x & 63, j = lane & 15;
  const int item = (lb - bm.start[f]) * ROWS + (threadIdx.x >> 4);
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
  // latencies overlap; each row is read once, with the streaming hint
  const int qs = 1 << gp((const uint8_t*)F.levels)[(size_t)(4 * y) * F.W + 4 * x];
  const gptr<const float> M = gp((const float*)msg_buf(F, 0, buf));
  const bool hl = x != 0, hr = x != w - 1, hu = y != 0, hd = y != h - 1;
  Row<S> b = load_row_nt<S>(gp((const float*)F.cost[0]) + (size_t)cell * N, j);
  const Row<S> tl = load_row_nt<S>(M + (((size_t)(hl ? cell - 1 : cell)) * 4 + 3) * N, j);
  const Row<S> tr = load_row_nt<S>(M + (((size_t)(hr ? cell + 1 : cell)) * 4 + 2) * N, j);
  const Row<S> tu = load_row_nt<S>(M + (((size_t)(hu ? cell - w : cell)) * 4 + 1) * N, j);
  const Row<S> td = load_row_nt<S>(M + (((size_t)(hd ? cell + w : cell)) * 4 + 0) * N, j);
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
__global__ __launch_bounds__(256) void k_copy_messages(const FrameDev* __restrict__ frames, float* out, int buf) {
  const FrameDev& F = frames[0];
  const int ncell = F.lw[0] * F.lh[0];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)ncell * 4 * N) return;
  const int p = (int)(i % N);
  const size_t t = i / N;
  const int dir = (int)(t % 4);
  const size_t cell = t / 4;
  out[((size_t)dir * ncell + cell) * N + row_sample<N>(p)] = msg_buf(F, 0, buf)[i];
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
                           bool first, bool active, bool copy_all, bool ext, const BpLaunchParams& p, int buf,
                           int pbuf) {
  hipLaunchKernelGGL((k_bp_sweep<N, ROWS>), dim3(nb), dim3(16 * ROWS), 0, s, dframes, bm, l, (int)A, (int)first,
                     (int)active, (int)copy_all, p.p1, p.p2, (int)ext, p.idepth_min, p.idepth_step, buf, pbuf);
}

void launch_bp_sweep(const FrameDev* dframes, const FrameDev* hframes, int n, int N, int l, bool A, bool first,
                     bool active, const BpLaunchParams& p, const Tuning& t, hipStream_t s, bool copy_all,
                     bool ext, int buf, int pbuf) {
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
  const bool e = ext && l == 0 && !first;
  if (N == 64) {
    if (rows == 4) bp_sweep_shape<64, 4>(nb, s, dframes, bm, l, A, first, active, copy_all, e, p, buf, pbuf);
    else if (rows == 8) bp_sweep_shape<64, 8>(nb, s, dframes, bm, l, A, first, active, copy_all, e, p, buf, pbuf);
    else bp_sweep_shape<64, 16>(nb, s, dframes, bm, l, A, first, active, copy_all, e, p, buf, pbuf);
  } else {
    if (rows == 4) bp_sweep_shape<128, 4>(nb, s, dframes, bm, l, A, first, active, copy_all, e, p, buf, pbuf);
    else if (rows == 8) bp_sweep_shape<128, 8>(nb, s, dframes, bm, l, A, first, active, copy_all, e, p, buf, pbuf);
    else bp_sweep_shape<128, 16>(nb, s, dframes, bm, l, A, first, active, copy_all, e, p, buf, pbuf);
  }
}

template <int N, int GA>
static void bp_pair_shape(int nb, hipStream_t s, const FrameDev* dframes, const BlockMap& bm, int l, bool first,
                          int rd, int wr, int pbuf, bool write_first, int seg, bool ext, const BpLaunchParams& p) {
  const int src = !first ? 0 : l < 3 ? 1 : 2;
  // extraction ends BP at level 0, so only SRC 0 and 1 have an extracting form
  auto k = src == 0   ? (ext ? k_bp_pair<N, GA, 0, true> : k_bp_pair<N, GA, 0, false>)
           : src == 1 ? (ext ? k_bp_pair<N, GA, 1, true> : k_bp_pair<N, GA, 1, false>)
                      : k_bp_pair<N, GA, 2, false>;
  hipLaunchKernelGGL(k, dim3(nb), dim3(32 * GA), 0, s, dframes, bm, l, 0, rd, wr, pbuf, (int)write_first, seg, p.p1,
                     p.p2, p.idepth_min, p.idepth_step);
}

// sweeps it and it + 1 (it even: the first updates the x + y even cells) of level l
static void launch_bp_pair(const FrameDev* dframes, const FrameDev* hframes, int n, int N, int l, bool first, int rd,
                           int wr, int pbuf, bool write_first, bool ext, const BpLaunchParams& p, const Tuning& t,
                           hipStream_t s) {
  const int ga = t.bp_pair == 16 ? 16 : 8;
  const int seg = (t.bp_seg == 8 || t.bp_seg == 16 || t.bp_seg == 64) ? t.bp_seg : 32;
  const int tw = 2 * (ga - 1);
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) {
    const int w = hframes[i].lw[l], h = hframes[i].lh[l];
    counts[i] = ((w + tw - 1) / tw) * ((h + seg - 1) / seg);
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 64) {
    if (ga == 16) bp_pair_shape<64, 16>(nb, s, dframes, bm, l, first, rd, wr, pbuf, write_first, seg, ext, p);
    else bp_pair_shape<64, 8>(nb, s, dframes, bm, l, first, rd, wr, pbuf, write_first, seg, ext, p);
  } else {
    if (ga == 16) bp_pair_shape<128, 16>(nb, s, dframes, bm, l, first, rd, wr, pbuf, write_first, seg, ext, p);
    else bp_pair_shape<128, 8>(nb, s, dframes, bm, l, first, rd, wr, pbuf, write_first, seg, ext, p);
  }
}

bool bp_level_pairs(int l, int iters, const Tuning& t) {
  const int levels = (t.bp_pair_levels >= 1 && t.bp_pair_levels <= 4) ? t.bp_pair_levels : 2;
  return t.bp_pair != 0 && l < levels && iters >= 2;
}

int bp_final_buf(int l, int iters, const Tuning& t) {
  return bp_level_pairs(l, iters, t) ? (iters / 2 - 1) & 1 : 0;
}

void launch_bp_levels(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const int iters[4],
                      const BpLaunchParams& p, const Tuning& t, hipStream_t s, bool ext_last, bool want_msgs) {
  for (int l = 3; l >= 0; --l) {
    const int it_l = iters[l];
    const int pbuf = l < 3 ? bp_final_buf(l + 1, iters[l + 1], t) : 0;
    const bool ext = ext_last && l == 0;
    if (it_l == 0) {
      launch_bp_sweep(dframes, hframes, n, N, l, true, true, false, p, t, s, true, false, 0, pbuf);
      continue;
    }
    if (!bp_level_pairs(l, it_l, t)) {
      for (int it = 0; it < it_l; ++it)
        launch_bp_sweep(dframes, hframes, n, N, l, (it % 2) == 0, it == 0, true, p, t, s, it_l < 2,
                        ext && it_l >= 2 && it == it_l - 1, 0, pbuf);
      continue;
    }
    // pairs (0, 1), (2, 3), ...: the first writes buffer 0, each later one
    // reads the buffer the previous pair wrote and writes the other; an odd
    // last sweep runs alone, in place. The last pair writes its A cells' messages
    // unless no later launch reads them (level 0 when its B cells' winners are
    // extracted there and the A cells' by the extraction stage, which reads only
    // B messages).
    int cur = 0;
    for (int it = 0; it + 1 < it_l; it += 2) {
      const bool first = it == 0, last = it + 2 == it_l;
      const int rd = first ? 0 : cur, wr = first ? 0 : cur ^ 1;
      const bool write_first = last && (l > 0 || want_msgs || !ext);
      launch_bp_pair(dframes, hframes, n, N, l, first, rd, wr, pbuf, write_first, ext && last, p, t, s);
      cur = wr;
    }
    if (it_l % 2) launch_bp_sweep(dframes, hframes, n, N, l, true, false, true, p, t, s, false, ext, cur, pbuf);
  }
}

template <int N, int ROWS>
static void extract_shape(int nb, hipStream_t s, const FrameDev* dframes, const BlockMap& bm,
                          const BpLaunchParams& p, int colour, int buf) {
  hipLaunchKernelGGL((k_extract<N, ROWS>), dim3(nb), dim3(16 * ROWS), 0, s, dframes, bm, p.idepth_min,
                     p.idepth_step, colour, buf);
}

void launch_extract(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const BpLaunchParams& p,
                    const Tuning& t, hipStream_t s, int colour, int buf) {
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
    if (rows == 4) extract_shape<64, 4>(nb, s, dframes, bm, p, colour, buf);
    else if (rows == 8) extract_shape<64, 8>(nb, s, dframes, bm, p, colour, buf);
    else extract_shape<64, 16>(nb, s, dframes, bm, p, colour, buf);
  } else {
    if (rows == 4) extract_shape<128, 4>(nb, s, dframes, bm, p, colour, buf);
    else if (rows == 8) extract_shape<128, 8>(nb, s, dframes, bm, p, colour, buf);
    else extract_shape<128, 16>(nb, s, dframes, bm, p, colour, buf);
  }
}

void launch_copy_messages(const FrameDev* dframes, const FrameDev* hframes, int n, int N, float* out,
                          hipStream_t s, int buf) {
  (void)n;
  const size_t total = (size_t)hframes[0].lw[0] * hframes[0].lh[0] * 4 * N;
  const int nb = (int)((total + 255) / 256);
  if (N == 64)
    hipLaunchKernelGGL(k_copy_messages<64>, dim3(nb), dim3(256), 0, s, dframes, out, buf);
  else
    hipLaunchKernelGGL(k_copy_messages<128>, dim3(nb), dim3(256), 0, s, dframes, out, buf);
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
