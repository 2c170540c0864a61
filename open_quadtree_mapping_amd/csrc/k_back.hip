// k_back.hip -- tridiagonal smoothness upsampling (reference:
// src/depth_upsample.cu:20-302) and Gaussian x Beta fusion (reference:
// src/depth_fusion.cu:13-213) for a batch of keyframes.
#include <type_traits>

#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

// ---------------------------------------------------------------------------
// Upsample solve (reference: smooth_row / smooth_col, src/depth_upsample.cu:
// 94-302). Per line (a row, then a column) the Thomas algorithm on
// A = I + lambda L(w) with two right-hand sides (x, n), in the reference's
// operations and order:
//   forward   den_k = b_k - ch_{k-1} a_k,   ch_k = c_k / den_k,
//             dh_k = (x_k - dh_{k-1} a_k) / den_k,   nh_k likewise,
//   backward  X_k = dh_k - ch_k X_{k+1},   N_k likewise,
// a_k = float(-lambda w(k)), c_k = float(-lambda w(k+1)) (double products) and
// b_k = float(1 + lambda (w(k) + w(k+1))), w(k) the coupling between steps
// k - 1 and k of the line (0 at the line start, the edge weight at its end).
// The chain state is 0 at the start of each of the line's two segments.
//
// A workgroup owns 64 lines, one per lane, and splits a step's work over waves
// that run side by side on the CU's SIMDs, one role each:
//   producers two waves, each for half a chunk of kUK steps: load it and write
//             its a (which is also the c of the step before), b and
//             right-hand sides to LDS;
//   ch        the coefficient chain (it depends on the weights only);
//   x, n      the two right-hand-side chains, reading ch_{k-1} from LDS;
//   out       (columns) X / N and the column pass's stores.
// The waves form a pipeline over chunks: in round r the producers fill chunk
// r, the ch wave runs chunk r - 1 and the x / n waves chunk r - 2, a barrier
// between rounds and rings of LDS slots between the roles. The forward values
// of a line (ch, dh, nh per step) are not kept: each chain wave stores its
// state at every chunk start (one float per chunk) and the backward sweep
// walks the chunks from the end, recomputing a chunk's forward values from its
// checkpoint (the same operations on the same inputs, so the same bits) into
// registers: a backward round of an x / n wave sweeps back through the chunk
// it recomputed in the round before, then recomputes the next one.
// ---------------------------------------------------------------------------
constexpr int kUK = 16;       // steps per chunk
constexpr int kUS = kUK + 1;  // LDS stride of one line's chunk (odd: lane-per-line access is conflict-free)
constexpr int kSlots = 3;     // coefficient slots: producer, ch and x / n work on three chunks at once

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Chunks of a line of len steps split at seg: segment 0 is [0, seg), segment
// 1 is [seg, len); chunk j covers steps [c0, c0 + cnt) of one segment.
struct Chunks {
  int seg, len, n0, n;
  __device__ Chunks(int seg_in, int len_in) {
    len = len_in;
    seg = seg_in < len_in ? seg_in : len_in;
    n0 = (seg + kUK - 1) / kUK;
    n = n0 + (len - seg + kUK - 1) / kUK;
  }
  __device__ void get(int j, int& c0, int& cnt, bool& first, bool& last) const {
    const int s0 = j < n0 ? 0 : seg, s1 = j < n0 ? seg : len;
    c0 = s0 + kUK * (j < n0 ? j : j - n0);
    cnt = min(kUK, s1 - c0);
    first = c0 == s0;
    last = c0 + cnt == s1;
  }
};

// Two producer waves share a chunk: wave HALF writes steps [8 HALF, 8 HALF + 8)
// of the slots (a and b of those steps, and the chunk's last c, a[kUK], for
// HALF = 1). Couplings of the half from image bytes px[j] = I(c0 - 1 +
// 8 HALF + j) (clamped to the line): w(c0 + 8 HALF + j) and a = float(-lambda
// w) for j = 0..8, then b.
constexpr int kHalf = kUK / 2;

template <int HALF>
__device__ __forceinline__ void half_couplings_lut(const uint8_t (&px)[kHalf + 2], int c0, int len, float edge,
                                                   double lam, const float* lw, const float* lm, float* Ms,
                                                   float* Bs) {
  float w[kHalf + 1], m[kHalf + 1];
#pragma unroll
  for (int j = 0; j <= kHalf; ++j) {
    const unsigned i = __builtin_amdgcn_sad_u8((unsigned)px[j + 1], (unsigned)px[j], 0u);  // |difference|
    w[j] = lw[i];
    m[j] = lm[i];
  }
  const int x0 = c0 + kHalf * HALF;
  if (HALF == 0 && c0 == 0) {
    w[0] = 0.0f;
    m[0] = (float)(-lam * 0.0);
  }
  if (x0 + kHalf >= len) {
#pragma unroll
    for (int j = 0; j <= kHalf; ++j)
      if (x0 + j == len) {
        w[j] = edge;
        m[j] = (float)(-lam * (double)edge);
      }
  }
#pragma unroll
  for (int j = 0; j < kHalf; ++j) Ms[x0 - c0 + j] = m[j];
  if (HALF == 1) Ms[kUK] = m[kHalf];
#pragma unroll
  for (int j = 0; j < kHalf; ++j) Bs[x0 - c0 + j] = (float)(1.0 + lam * (double)(w[j] + w[j + 1]));
}

// The same from explicit couplings w[j] = w(c0 + 8 HALF + j).
template <int HALF>
__device__ __forceinline__ void half_couplings(const float (&w)[kHalf + 1], double lam, float* Ms, float* Bs) {
#pragma unroll
  for (int j = 0; j < kHalf; ++j) Ms[kHalf * HALF + j] = (float)(-lam * (double)w[j]);
  if (HALF == 1) Ms[kUK] = (float)(-lam * (double)w[kHalf]);
#pragma unroll
  for (int j = 0; j < kHalf; ++j) Bs[kHalf * HALF + j] = (float)(1.0 + lam * (double)(w[j] + w[j + 1]));
}

// The chain waves always run a chunk's kUK steps. Only the last chunk of a
// segment can be shorter (cnt < kUK steps); its steps past cnt run on whatever
// the slots hold and only feed state that is discarded: the forward state is
// reset at the next segment's start, the backward sweep restarts from 0 at the
// segment's last step, and the stores of those steps go to a sink.

// ch wave, one chunk: den_k = b_k - ch_{k-1} a_k, ch_k = c_k / den_k. Writes
// the ring Cr[0] = ch before the chunk, Cr[k + 1] = ch_k; returns the state.
__device__ __forceinline__ float ch_chunk(const float* Ms, const float* Bs, float* Cr, float st) {
  float m[kUK + 1], b[kUK];
#pragma unroll
  for (int k = 0; k <= kUK; ++k) m[k] = Ms[k];
#pragma unroll
  for (int k = 0; k < kUK; ++k) b[k] = Bs[k];
  Cr[0] = st;
#pragma unroll
  for (int k = 0; k < kUK; ++k) {
    const float den = b[k] - st * m[k];
    st = m[k + 1] / den;
    Cr[k + 1] = st;
  }
  return st;
}

// x / n wave, one chunk forward: v_k = (r_k - v_{k-1} a_k) / den_k with den_k
// recomputed from ch_{k-1} = Cr[k] exactly as the ch wave formed it. DERIVE:
// the right-hand side is n_k = (x_k > 0) from the x slot.
template <bool DERIVE>
__device__ __forceinline__ float rhs_chunk(const float* Ms, const float* Bs, const float* Rs, const float* Cr,
                                           float st, float (&v)[kUK]) {
  float m[kUK], b[kUK], r[kUK], c[kUK];
#pragma unroll
  for (int k = 0; k < kUK; ++k) {
    m[k] = Ms[k];
    b[k] = Bs[k];
    r[k] = Rs[k];
    c[k] = Cr[k];
  }
#pragma unroll
  for (int k = 0; k < kUK; ++k) {
    const float rk = DERIVE ? (r[k] > 0.0f ? 1.0f : 0.0f) : r[k];
    const float den = b[k] - c[k] * m[k];
    st = (rk - st * m[k]) / den;
    v[k] = st;
  }
  return st;
}

// Backward round of an x / n wave: sweeps back through chunk q - 1 (vb in
// place, its ch values cb): X_k = v_k - ch_k X_{k+1}, restarting from 0 at step
// stop, then recomputes the forward values v of chunk q from its checkpoint st
// (as rhs_chunk). Also returns chunk q's ch values (Cr[k + 1]) in c1 for the
// next round's sweep.
template <bool DERIVE>
__device__ __forceinline__ float recompute_back(const float* Ms, const float* Bs, const float* Rs, const float* Cr,
                                                float st, float (&v)[kUK], float (&c1)[kUK], int stop, float nxt,
                                                float (&vb)[kUK], const float (&cb)[kUK]) {
  float m[kUK], b[kUK], r[kUK], c[kUK];
#pragma unroll
  for (int k = 0; k < kUK; ++k) {
    m[k] = Ms[k];
    b[k] = Bs[k];
    r[k] = Rs[k];
    c[k] = Cr[k];
    c1[k] = Cr[k + 1];
  }
  if (stop < 0) {  // no segment end in the chunk (the usual case)
#pragma unroll
    for (int kb = kUK - 1; kb >= 0; --kb) {
      nxt = vb[kb] - cb[kb] * nxt;
      vb[kb] = nxt;
    }
  } else {
#pragma unroll
    for (int kb = kUK - 1; kb >= 0; --kb) {
      nxt = kb == stop ? 0.0f : nxt;
      nxt = vb[kb] - cb[kb] * nxt;
      vb[kb] = nxt;
    }
  }
#pragma unroll
  for (int k = 0; k < kUK; ++k) {
    const float rk = DERIVE ? (r[k] > 0.0f ? 1.0f : 0.0f) : r[k];
    const float den = b[k] - c[k] * m[k];
    st = (rk - st * m[k]) / den;
    v[k] = st;
  }
  return nxt;
}

// ---------------------------------------------------------------------------
// Loads of the producers. A wave's loads are counted in order, and the
// compiler's waits for a chunk prepared one round earlier assume the fewest
// loads any path could have issued since; so every load below is issued
// unconditionally and in a fixed number (no data-dependent or optional loads,
// pointer selects instead), and lanes past the image edge load, compute and
// store exactly what the last line's lane does. A half chunk's 8 floats of
// one line are read as two dword-aligned float4 from min(c0 + 8 HALF, len - 8)
// (lines are at least 16 long); the end of a line whose length is not a
// multiple of 8 then sits shifted by sh = c0 + 8 HALF - base in the loaded
// values and is written to LDS with that shift.
// ---------------------------------------------------------------------------
typedef float4 f4_dword_aligned __attribute__((aligned(4)));

// A half's 8 floats of one line, read as two dword-aligned float4 from
// min(c0 + 8 HALF, len - 8); dst[8 HALF + i - sh] = v[i] for i >= sh.
__device__ __forceinline__ void load8(gptr<const float> line, int base, float (&v)[kHalf]) {
  const gptr<const f4_dword_aligned> p = (gptr<const f4_dword_aligned>)(line + base);
#pragma unroll
  for (int q = 0; q < kHalf / 4; ++q) {
    const float4 t = p[q];
    v[4 * q] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}
template <int HALF>
__device__ __forceinline__ void store8(float* dst, int sh, const float (&v)[kHalf]) {
  if (sh == 0) {
#pragma unroll
    for (int i = 0; i < kHalf; ++i) dst[kHalf * HALF + i] = v[i];
  } else {
#pragma unroll
    for (int i = 0; i < kHalf; ++i)
      if (i >= sh) dst[kHalf * HALF + i - sh] = v[i];
  }
}

// ---------------------------------------------------------------------------
// Row pass. Lines are image rows, lanes rows y0 .. y0 + 63. Results X and N
// go to the planes rowx[x][y] and rowx[W H + x][y] (column-major, so a wave's
// store of one step is 64 consecutive floats). Mode 0 composes the sparse map
// on the fly (BP winner depth at node pixels (4x, 4y), else the semi-dense
// pre-fill, else 0) and writes it to sparse_out when asked; mode 1 reads an
// explicit sparse map; mode 2 is the generalised form with caller-supplied
// right-hand sides, confidences and weights (and writes the row results,
// row-major, to row_x_out / row_n_out when asked).
// ---------------------------------------------------------------------------
struct UpRowLds {
  float lw[256], lm[256];
  float M[kSlots][64 * kUS], B[kSlots][64 * kUS], R[kSlots][64 * kUS], N[kSlots][64 * kUS];
  float C[2][64 * kUS];
};
// the adjoint row pass (mode 3) also hands the x / n chains' solutions to an
// out wave
struct UpRowLdsAdj : UpRowLds {
  float XO[2][64 * kUS], NO[2][64 * kUS];
};
__device__ __forceinline__ float* adj_xo(UpRowLdsAdj& L) { return &L.XO[0][0]; }
__device__ __forceinline__ float* adj_no(UpRowLdsAdj& L) { return &L.NO[0][0]; }
__device__ __forceinline__ float* adj_xo(UpRowLds&) { return nullptr; }  // modes 0-2: no out wave
__device__ __forceinline__ float* adj_no(UpRowLds&) { return nullptr; }

// The adjoint out wave's inputs for one chunk (loaded a round ahead): the
// forward result z at the chunk's steps and the coupling w(i + 1) of each
// step's upper edge; the row adjoint also reads the depths and confidences.
struct AdjOutIn {
  float zx[kUK], zy[kUK], w[kUK], d[kUK], c[kUK];
};

// The adjoint out wave's running state along a line: the solutions and forward
// results of the step after the current one, the split edge's share from
// segment 1, the line's lambda partial, and the gradient of edge 0 (once).
struct AdjEdgeState {
  float vxn = 0.0f, vnn = 0.0f, zxn = 0.0f, zyn = 0.0f, carry = 0.0f;
  double part = 0.0;  // sum over the line's edges of w dL/d(lambda w)
  float e_first = 0.0f;
  bool has_first = false;
};

// dL/dw of edges c0 + 1 .. c0 + cnt of one chunk (ge[k] for edge c0 + k + 1)
// from the adjoint solutions v (LDS) and forward results z of its steps, k from
// cnt - 1 down to 0: -scale ((vx_i - vx_{i+1}) (zx_i - zx_{i+1}) + (vn_i -
// vn_{i+1}) (zy_i - zy_{i+1})) inside the segment [s0, s1), -scale (vx zx + vn
// zy) at its end (plus segment 1's share of the split edge), and at its first
// step the edge before it (edge 0: e_first; the split edge: carry). The lambda
// partial adds w dL/d(lambda w) in the same order, each term an exact double
// product (24-bit significands), so one fma rounds as the add after it would.
// The chunk's LDS reads are issued before its steps; the segment tests are
// wave-uniform branches around the rare end and start steps.
template <bool WHOLE>
__device__ __forceinline__ void adj_edges(const float* XO, const float* NO, const AdjOutIn& b, int c0, int cnt, int s0,
                                          int s1, int len, float gscale, float w_s0, AdjEdgeState& st,
                                          float (&ge)[kUK]) {
  float vx[kUK], vn[kUK];
#pragma unroll
  for (int k = 0; k < kUK; ++k) {
    vx[k] = XO[k];
    vn[k] = NO[k];
  }
#pragma unroll
  for (int k = kUK - 1; k >= 0; --k) {
    ge[k] = 0.0f;
    if (WHOLE || k < cnt) {
      const int i = c0 + k;
      const float zx = b.zx[k], zy = b.zy[k];
      float t;
      if (i + 1 < s1) {
        t = (vx[k] - st.vxn) * (zx - st.zxn) + (vn[k] - st.vnn) * (zy - st.zyn);
        ge[k] = -gscale * t;
      } else {
        t = vx[k] * zx + vn[k] * zy;
        const float e = -gscale * t;
        ge[k] = s1 == len ? e : e + st.carry;
      }
      st.part = fma((double)b.w[k], -(double)t, st.part);
      if (i == s0) {
        const float t0 = vx[k] * zx + vn[k] * zy;
        const float e0 = -gscale * t0;
        st.part = fma((double)w_s0, -(double)t0, st.part);
        if (s0 == 0) {
          st.e_first = e0;
          st.has_first = true;
        } else {
          st.carry = e0;  // segment 1 is swept first: its share of the split edge
        }
      }
      st.vxn = vx[k];
      st.vnn = vn[k];
      st.zxn = zx;
      st.zyn = zy;
    }
  }
}

template <int MODE>
struct RowBuf {                // one producer wave's half of a chunk
  uint8_t px[kHalf + 2];       // image at c0 - 1 + 8 HALF + j (modes 0, 1)
  float sp[kHalf];             // pre-fill (mode 0), sparse map (mode 1), d (mode 2), from base
  float cf[kHalf];             // confidence (mode 2), from base
  float wv[kHalf + 1];         // row couplings (mode 2)
  int lb[kUK / 8];             // BP labels of the half's node cells (mode 0)
  float dv[kUK / 8];           // BP depths of the half's node cells (mode 0)
  float ax[kHalf], an[kHalf];  // the column adjoint's solution (X, N) at c0 + 8 HALF + j (mode 3)
};

template <int MODE>
__global__ __launch_bounds__(MODE == 3 ? 384 : 320) void k_up_row(const FrameDev* __restrict__ frames, BlockMap bm,
                                                                  const float* __restrict__ wlut) {
  __shared__ typename std::conditional<MODE == 3, UpRowLdsAdj, UpRowLds>::type L;
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H, Wc = F.Wc, Hc = F.Hc;
  const double lam = F.lam;
  // waves: 0 ch, 1 and 2 producers (half a chunk each), 3 x, 4 n; waves 0 and 4
  // (the two chains with the most idle issue slots) share a SIMD; mode 3
  // (the adjoint) adds wave 5, the out wave
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;  // wave: uniform
  // the adjoint's out wave trails the x / n chains by one more round
  constexpr int kTail = MODE == 3 ? 1 : 0;
  if (MODE < 2 && threadIdx.x < 256) {
    const float w = wlut[threadIdx.x];
    L.lw[threadIdx.x] = w;
    L.lm[threadIdx.x] = (float)(-lam * (double)w);
  }
  const int y = min((blockIdx.x - bm.start[f]) * 64 + lane, H - 1);  // lanes past the image repeat row H - 1
  const size_t row = (size_t)y * W;
  const Chunks ch(F.row_seg, W);
  const int nch = ch.n;
  const float edge = F.edge_r ? 1.0f : 0.0f;
  const gptr<float> ck = gp((float*)F.fwd);  // checkpoints [role][chunk][y]
  const int lo = lane * kUS;
  __syncthreads();

  auto producer = [&](auto half_tag) __attribute__((always_inline)) {
    constexpr int HALF = decltype(half_tag)::value;
    constexpr int h0 = kHalf * HALF;  // the half's first step
    const gptr<const uint8_t> img = gp(F.key + (size_t)y * F.pitch);
    const bool have_prefill = F.prefill != nullptr;
    // mode 0 without a pre-fill reads (and ignores) the dense buffer instead
    const gptr<const float> src =
        gp(MODE == 2 ? F.g_d : (MODE == 1 ? F.sparse_in : (have_prefill ? F.prefill : (const float*)F.dense)));
    const bool have_conf = F.g_conf != nullptr;
    const gptr<const float> conf = gp(have_conf ? F.g_conf : F.g_d);
    const gptr<const float> g_wrow = gp(F.g_wrow);
    const bool node_row = (y & 3) == 0 && (y >> 2) < Hc;
    const size_t cb = (size_t)min(y >> 2, Hc - 1) * Wc;
    const gptr<const int32_t> label = gp((const int32_t*)F.label) + cb;
    const gptr<const float> dcoarse = gp((const float*)F.dcoarse) + cb;
    const gptr<const float2> adjr = gp((const float2*)F.adjxn) + y;
    auto load = [&](RowBuf<MODE>& b, int j) __attribute__((always_inline)) {
      int c0, cnt;
      bool first, last;
      ch.get(j, c0, cnt, first, last);
      const int base = min(c0 + h0, W - kHalf);
      if (MODE == 3) {
        // every load unconditional, at columns clamped to the line
#pragma unroll
        for (int k = 0; k <= kHalf; ++k) b.wv[k] = g_wrow[(size_t)y * (W + 1) + min(c0 + h0 + k, W)];
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
          const float2 r = adjr[(size_t)min(c0 + h0 + k, W - 1) * H];
          b.ax[k] = r.x;
          b.an[k] = r.y;
        }
        return;
      }
      if (MODE == 2) {
#pragma unroll
        for (int k = 0; k <= kHalf; ++k) b.wv[k] = g_wrow[(size_t)y * (W + 1) + min(c0 + h0 + k, W)];
        load8(src + row, base, b.sp);
        load8(conf + row, base, b.cf);
      } else {
        const int x0 = c0 + h0 - 1;  // first image byte of the half
        if (x0 >= 0 && x0 + kHalf + 2 <= W) {
          const gptr<const uint8_t> p = img + x0;
#pragma unroll
          for (int k = 0; k < kHalf + 2; ++k) b.px[k] = p[k];
        } else {
#pragma unroll
          for (int k = 0; k < kHalf + 2; ++k) b.px[k] = img[clampi(x0 + k, 0, W - 1)];
        }
        load8(src + row, base, b.sp);
        if (MODE == 0) {
          // node pixels (4 cx, 4 cy); chunk starts are multiples of 4 here
#pragma unroll
          for (int q = 0; q < kUK / 8; ++q) {
            const int cx = min((c0 >> 2) + 2 * HALF + q, Wc - 1);
            b.lb[q] = label[cx];
            b.dv[q] = dcoarse[cx];
          }
        }
      }
    };
    auto produce = [&](RowBuf<MODE>& b, int j, int slot) __attribute__((always_inline)) {
      int c0, cnt;
      bool first, last;
      ch.get(j, c0, cnt, first, last);
      const int sh = c0 + h0 - min(c0 + h0, W - kHalf);
      float* Ms = L.M[slot] + lo;
      float* Bs = L.B[slot] + lo;
      float* Rs = L.R[slot] + lo;
      float* Ns = L.N[slot] + lo;
      if (MODE == 3) {
        half_couplings<HALF>(b.wv, lam, Ms, Bs);
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
          Rs[h0 + k] = b.ax[k];
          Ns[h0 + k] = b.an[k];
        }
        return;
      }
      if (MODE == 2) {
        half_couplings<HALF>(b.wv, lam, Ms, Bs);
        float rx[kHalf], rn[kHalf];
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
          if (have_conf) {
            rx[k] = b.cf[k] * b.sp[k];
            rn[k] = b.cf[k];
          } else {
            rx[k] = b.sp[k];
            rn[k] = b.sp[k] > 0.0f ? 1.0f : 0.0f;
          }
        }
        store8<HALF>(Rs, sh, rx);
        store8<HALF>(Ns, sh, rn);
      } else {
        half_couplings_lut<HALF>(b.px, c0, W, edge, lam, L.lw, L.lm, Ms, Bs);
        float sv[kHalf];
#pragma unroll
        for (int k = 0; k < kHalf; ++k) sv[k] = (MODE == 1 || have_prefill) ? b.sp[k] : 0.0f;
        store8<HALF>(Rs, sh, sv);
        if (MODE == 0 && node_row) {
#pragma unroll
          for (int q = 0; q < kUK / 8; ++q) {
            const int k = h0 + 4 * q;
            if (k < cnt && (c0 >> 2) + 2 * HALF + q < Wc && b.lb[q] >= 0) Rs[k] = b.dv[q];
          }
        }
      }
    };
    // forward: chunk r in round r; backward: chunk nch - 1 - r in round r. The
    // next round's inputs are loaded while this round's chunk is produced (the
    // two buffers alternate; past the last chunk the loads repeat it).
    RowBuf<MODE> A, Bf;
    for (int pass = 0; pass < 2; ++pass) {
      auto chunk_at = [&](int r) __attribute__((always_inline)) { return pass == 0 ? r : nch - 1 - r; };
      const int rounds = nch + 2 + (1 + kTail) * pass;  // the backward sweep's stores trail by one round
      load(A, chunk_at(0));
      for (int r = 0; r < rounds; r += 2) {
        load(Bf, chunk_at(min(r + 1, nch - 1)));
        if (r < nch) produce(A, chunk_at(r), r % kSlots);
        __syncthreads();
        if (r + 1 < rounds) {
          load(A, chunk_at(min(r + 2, nch - 1)));
          if (r + 1 < nch) produce(Bf, chunk_at(r + 1), (r + 1) % kSlots);
          __syncthreads();
        }
      }
    }
  };

  if (wave == 1) {
    producer(std::integral_constant<int, 0>{});
  } else if (wave == 2) {
    producer(std::integral_constant<int, 1>{});
  } else if (wave == 0) {
    // ---------------- ch chain ----------------
    float st = 0.0f;
    for (int r = 0; r < nch + 2; ++r) {
      const int j = r - 1;
      if (j >= 0 && j < nch) {
        int c0, cnt;
        bool first, last;
        ch.get(j, c0, cnt, first, last);
        if (first) st = 0.0f;
        ck[(size_t)j * H + y] = st;
        st = ch_chunk(L.M[j % kSlots] + lo, L.B[j % kSlots] + lo, L.C[j & 1] + lo, st);
      }
      __syncthreads();
    }
    // backward: the checkpoint of round r + 1 is loaded in round r (two
    // registers alternate)
    auto ck_at = [&](int q) __attribute__((always_inline)) { return ck[(size_t)(nch - 1 - clampi(q, 0, nch - 1)) * H + y]; };
    auto back_round = [&](int q, float s0) __attribute__((always_inline)) {
      if (q >= 0 && q < nch) {
        int c0, cnt;
        bool first, last;
        ch.get(nch - 1 - q, c0, cnt, first, last);
        ch_chunk(L.M[q % kSlots] + lo, L.B[q % kSlots] + lo, L.C[q & 1] + lo, s0);
      }
    };
    const int rounds = nch + 3 + kTail;
    float cA = ck_at(-1), cB;
    for (int r = 0; r < rounds; r += 2) {
      cB = ck_at(r);
      back_round(r - 1, cA);
      __syncthreads();
      if (r + 1 < rounds) {
        cA = ck_at(r + 1);
        back_round(r, cB);
        __syncthreads();
      }
    }
  } else if (wave == 3 || wave == 4) {
    // ---------------- x (wave 3) and n (wave 4) chains ----------------
    const int role = wave == 3 ? 1 : 2;  // checkpoint plane 1 (x) or 2 (n)
    // mode 3 writes the row adjoint's solution planes (when asked) where the
    // other modes write the row results
    const gptr<float> out = gp(MODE == 3 ? F.adj_rowv : F.rowx) + (role == 2 ? (size_t)W * H : 0);
    const bool store_planes = MODE != 3 || F.adj_rowv != nullptr;
    const gptr<float> out_rm = gp(role == 2 ? F.row_n_out : F.row_x_out);
    const gptr<float> sparse_out = gp(F.sparse_out);
    const float* Rsrc = (role == 2 && MODE >= 2) ? &L.N[0][0] : &L.R[0][0];
    const bool derive = role == 2 && MODE < 2;
    const size_t ckb = (size_t)role * nch * H;
    float st = 0.0f, v[kUK];
    for (int r = 0; r < nch + 2; ++r) {
      const int j = r - 2;
      if (j >= 0) {
        int c0, cnt;
        bool first, last;
        ch.get(j, c0, cnt, first, last);
        if (first) st = 0.0f;
        ck[ckb + (size_t)j * H + y] = st;
        const int s = j % kSlots;
        const float* Rs = Rsrc + s * 64 * kUS + lo;
        if (derive)
          st = rhs_chunk<true>(L.M[s] + lo, L.B[s] + lo, Rs, L.C[j & 1] + lo, st, v);
        else
          st = rhs_chunk<false>(L.M[s] + lo, L.B[s] + lo, Rs, L.C[j & 1] + lo, st, v);
        // the composed sparse map, once (forward sweep, x wave)
        if (MODE == 0 && role == 1 && sparse_out) {
#pragma unroll
          for (int k = 0; k < kUK; ++k)
            if (k < cnt) sparse_out[row + c0 + k] = Rs[k];
        }
      }
      __syncthreads();
    }
    // backward: round r recomputes chunk q = r - 2 and sweeps back through
    // chunk q - 1 (kept in registers from round r - 1), then stores chunk
    // q - 1; checkpoints are loaded a round ahead. Register sets alternate
    // between rounds. A partial chunk's steps past its end store to a sink
    // slot past the checkpoints, so every round issues the same stores.
    float nxt = 0.0f;
    const gptr<float> sink = ck + (size_t)3 * nch * H + y;
    auto ck_at = [&](int q) __attribute__((always_inline)) {
      return ck[ckb + (size_t)(nch - 1 - clampi(q, 0, nch - 1)) * H + y];
    };
    float va[kUK], ca[kUK], vb[kUK], cb[kUK];
#pragma unroll
    for (int k = 0; k < kUK; ++k) va[k] = ca[k] = vb[k] = cb[k] = 0.0f;
    auto back_round = [&](int q, float s0, float (&vn)[kUK], float (&cn)[kUK], float (&vp)[kUK],
                          const float (&cp)[kUK]) __attribute__((always_inline)) {
      if (q >= 0 && q <= nch) {
        const int qc = min(q, nch - 1);  // past the last chunk: recompute a valid chunk, unused
        const int s = qc % kSlots;
        int c0, cnt;
        bool first, last;
        ch.get(nch - q, c0, cnt, first, last);  // chunk q - 1 (valid when q >= 1)
        const int stop = last ? cnt - 1 : -1;
        if (derive)
          nxt = recompute_back<true>(L.M[s] + lo, L.B[s] + lo, Rsrc + s * 64 * kUS + lo, L.C[qc & 1] + lo, s0, vn,
                                     cn, stop, nxt, vp, cp);
        else
          nxt = recompute_back<false>(L.M[s] + lo, L.B[s] + lo, Rsrc + s * 64 * kUS + lo, L.C[qc & 1] + lo, s0, vn,
                                      cn, stop, nxt, vp, cp);
        if (q >= 1 && MODE == 3) {
          // the adjoint hands chunk q - 1 to the out wave
          float* O = (role == 2 ? adj_no(L) : adj_xo(L)) + ((q - 1) & 1) * 64 * kUS + lo;
#pragma unroll
          for (int k = 0; k < kUK; ++k) O[k] = vp[k];
        }
        if (q >= 1 && store_planes) {
          if (cnt == kUK) {
            // 32-bit offsets from the plane's base (W H < 2^28 floats)
            const unsigned o0 = (unsigned)c0 * (unsigned)H + (unsigned)y;
#pragma unroll
            for (int k = 0; k < kUK; ++k) out[o0 + (unsigned)k * (unsigned)H] = vp[k];
          } else {
#pragma unroll
            for (int k = 0; k < kUK; ++k) *(k < cnt ? out + (size_t)(c0 + k) * H + y : sink) = vp[k];
          }
          if (MODE == 2 && out_rm) {
#pragma unroll
            for (int k = 0; k < kUK; ++k)
              if (k < cnt) out_rm[row + c0 + k] = vp[k];
          }
        }
      }
    };
    const int rounds = nch + 3 + kTail;
    float sA = ck_at(-2), sB;
    for (int r = 0; r < rounds; r += 2) {
      sB = ck_at(r - 1);
      back_round(r - 2, sA, va, ca, vb, cb);
      __syncthreads();
      if (r + 1 < rounds) {
        sA = ck_at(r);
        back_round(r - 1, sB, vb, cb, va, ca);
        __syncthreads();
      }
    }
  } else if (MODE == 3) {
    // ---------------- adjoint out (wave 5) ----------------
    // From the row adjoint's solutions v (x / n waves) and the forward row
    // results z (the rowx planes): dL/dw of every edge as in the column pass's
    // adjoint out wave, and the per-row lambda partial (dL/dd and dL/dconf
    // follow from the solution planes the x / n waves store, k_adj_dconf).
    // Chunks come from the row's end; a round's inputs are loaded a round
    // ahead.
    const gptr<const float> zxp = gp((const float*)F.rowx) + y;
    const gptr<const float> zyp = zxp + (size_t)W * H;
    const size_t wrow = (size_t)y * (W + 1);
    const gptr<const float> wr = gp(F.g_wrow) + wrow;
    const gptr<float> gwr = gp(F.grad_wrow);
    const float gscale = F.grad_w_unscaled ? 1.0f : (float)lam;
    for (int r = 0; r < nch + 2; ++r) __syncthreads();
    AdjOutIn A, B;
    // z from the column-major planes (a wave's load of one step is 64
    // consecutive floats); w(i + 1) as the lane's own 16 consecutive floats
    // (four dword-aligned float4 from c0), or one by one for a chunk that ends
    // within 16 of the row's end. All of a chunk's loads are issued a round
    // ahead, so that no round waits for the next one's.
    auto load_o = [&](AdjOutIn& b, int q) __attribute__((always_inline)) {
      int c0, cnt;
      bool first, last;
      ch.get(nch - 1 - clampi(q, 0, nch - 1), c0, cnt, first, last);
#pragma unroll
      for (int k = 0; k < kUK; ++k) {
        const int i = min(c0 + k, W - 1);
        b.zx[k] = zxp[(size_t)i * H];
        b.zy[k] = zyp[(size_t)i * H];
      }
      if (c0 <= W - kUK) {
        const gptr<const f4_dword_aligned> q4 = (gptr<const f4_dword_aligned>)(wr + c0 + 1);
#pragma unroll
        for (int k = 0; k < kUK / 4; ++k) {
          const float4 t = q4[k];
          b.w[4 * k] = t.x;
          b.w[4 * k + 1] = t.y;
          b.w[4 * k + 2] = t.z;
          b.w[4 * k + 3] = t.w;
        }
      } else {  // the row's last chunk (once per row)
#pragma unroll
        for (int k = 0; k < kUK; ++k) b.w[k] = wr[min(c0 + k + 1, W)];
      }
    };
    // the weights of the edges that start the two segments
    const float w_seg0 = wr[0], w_seg1 = wr[min(ch.seg, W)];
    AdjEdgeState st;
    auto out_round = [&](const AdjOutIn& b, int q) __attribute__((always_inline)) {
      if (q < 0 || q >= nch) return;
      const int j = nch - 1 - q;
      int c0, cnt;
      bool first, last;
      ch.get(j, c0, cnt, first, last);
      const bool seg1 = j >= ch.n0;
      const int s0 = seg1 ? ch.seg : 0, s1 = seg1 ? ch.len : ch.seg;
      const float* XO = adj_xo(L) + (q & 1) * 64 * kUS + lo;
      const float* NO = adj_no(L) + (q & 1) * 64 * kUS + lo;
      float ge[kUK];  // edge c0 + k + 1
      if (cnt == kUK)
        adj_edges<true>(XO, NO, b, c0, cnt, s0, s1, W, gscale, seg1 ? w_seg1 : w_seg0, st, ge);
      else
        adj_edges<false>(XO, NO, b, c0, cnt, s0, s1, W, gscale, seg1 ? w_seg1 : w_seg0, st, ge);
      // the lane's 16 consecutive edge gradients as four float4 (a whole chunk), else one by one
      if (gwr) {
        const gptr<float> p = gwr + wrow + c0 + 1;
        if (cnt == kUK) {
          const gptr<f4_dword_aligned> q4 = (gptr<f4_dword_aligned>)p;
#pragma unroll
          for (int k = 0; k < kUK / 4; ++k) q4[k] = make_float4(ge[4 * k], ge[4 * k + 1], ge[4 * k + 2], ge[4 * k + 3]);
        } else {
#pragma unroll
          for (int k = 0; k < kUK; ++k)
            if (k < cnt) p[k] = ge[k];
        }
      }
    };
    // round r forms chunk r - 4, whose inputs were loaded two rounds ahead:
    // three buffers take turns (the loop is unrolled by three without
    // conditions; its last rounds follow it)
    const int rounds = nch + 4;
    AdjOutIn C;
    load_o(A, -4);
    load_o(B, -3);
    int r = 0;
    for (; r + 2 < rounds; r += 3) {
      load_o(C, r - 2);
      out_round(A, r - 4);
      __syncthreads();
      load_o(A, r - 1);
      out_round(B, r - 3);
      __syncthreads();
      load_o(B, r);
      out_round(C, r - 2);
      __syncthreads();
    }
    if (r < rounds) {
      out_round(A, r - 4);
      __syncthreads();
      ++r;
    }
    if (r < rounds) {
      out_round(B, r - 4);
      __syncthreads();
    }
    if (gwr && st.has_first) gwr[wrow] = st.e_first;
    if (F.lam_part) F.lam_part[y] = st.part;
  }
}

// dL/dd and dL/dconf of the generalised upsample from the row adjoint's
// solution planes v ([x][y]): rhs_x = conf d and rhs_n = conf give
// dL/dd = conf vx, dL/dconf = d vx + vn (without confidences, rhs_x = d and
// rhs_n = 1{d > 0}: dL/dd = vx, no dL/dconf). A 64 x 64 tile per block,
// transposed through LDS so that both the planes and the row-major outputs are
// read and written along their rows.
__global__ __launch_bounds__(256) void k_adj_dconf(const FrameDev* __restrict__ frames, BlockMap bm) {
  __shared__ float tx[64][65], tn[64][65];
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const int tiles_x = (W + 63) / 64;
  const int b = blockIdx.x - bm.start[f];
  const int x0 = (b % tiles_x) * 64, y0 = (b / tiles_x) * 64;
  const gptr<const float> vx = gp((const float*)F.adj_rowv), vn = vx + (size_t)W * H;
  const bool have_conf = F.g_conf != nullptr;
  const gptr<const float> d = gp(F.g_d), conf = gp(F.g_conf);
  const gptr<float> gd = gp(F.grad_d), gc = gp(F.grad_conf);
  const int t = threadIdx.x;
  // whole tiles of images whose sizes are multiples of 4 move as float4 (16
  // consecutive bytes of a plane column or an image row); others one by one
  auto a16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool quads = (W & 3) == 0 && (H & 3) == 0 && x0 + 64 <= W && y0 + 64 <= H && a16(F.adj_rowv) &&
                     a16(F.g_d) && a16(F.g_conf) && a16(F.grad_d) && a16(F.grad_conf);
  if (quads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = t + 256 * k, c = i >> 4, qy = (i & 15) * 4;
      const size_t o = (size_t)(x0 + c) * H + y0 + qy;
      const float4 a = *(const gptr<const float4>)(vx + o), n = *(const gptr<const float4>)(vn + o);
      tx[c][qy] = a.x; tx[c][qy + 1] = a.y; tx[c][qy + 2] = a.z; tx[c][qy + 3] = a.w;
      tn[c][qy] = n.x; tn[c][qy + 1] = n.y; tn[c][qy + 2] = n.z; tn[c][qy + 3] = n.w;
    }
  } else {
    for (int c = t >> 6; c < 64; c += 4) {
      const int x = x0 + c, y = y0 + (t & 63);
      if (x < W && y < H) {
        tx[c][t & 63] = vx[(size_t)x * H + y];
        tn[c][t & 63] = vn[(size_t)x * H + y];
      }
    }
  }
  __syncthreads();
  if (quads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = t + 256 * k, r = i >> 4, qx = (i & 15) * 4;
      const size_t o = (size_t)(y0 + r) * W + x0 + qx;
      float v[4], n[4], g0[4], g1[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[e] = tx[qx + e][r];
        n[e] = tn[qx + e][r];
      }
      if (have_conf) {
        const float4 cf = *(const gptr<const float4>)(conf + o), dv = *(const gptr<const float4>)(d + o);
        const float cfa[4] = {cf.x, cf.y, cf.z, cf.w}, dva[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          g0[e] = cfa[e] * v[e];
          g1[e] = dva[e] * v[e] + n[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          g0[e] = v[e];
          g1[e] = 0.0f;
        }
      }
      if (gd) *(gptr<float4>)(gd + o) = make_float4(g0[0], g0[1], g0[2], g0[3]);
      if (gc) *(gptr<float4>)(gc + o) = make_float4(g1[0], g1[1], g1[2], g1[3]);
    }
  } else {
    for (int r = t >> 6; r < 64; r += 4) {
      const int x = x0 + (t & 63), y = y0 + r;
      if (x < W && y < H) {
        const size_t i = (size_t)y * W + x;
        const float v = tx[t & 63][r], n = tn[t & 63][r];
        if (gd) gd[i] = have_conf ? conf[i] * v : v;
        if (gc) gc[i] = have_conf ? d[i] * v + n : 0.0f;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Column pass. Lines are image columns, lanes columns x0 .. x0 + 63; the
// right-hand sides are the row pass's X and N planes. The out wave writes
// dense = X / N (NaN where N = 0, as the reference) row-major, the column
// results (X, N) to colxn when asked, and, during the forward sweep, resets the
// fusion z-buffer when asked (the fusion stage runs next).
// ---------------------------------------------------------------------------
struct UpColLds {
  float lw[256], lm[256];
  float M[kSlots][64 * kUS], B[kSlots][64 * kUS], R[kSlots][64 * kUS], N[kSlots][64 * kUS];
  float C[2][64 * kUS];
  float XO[2][64 * kUS], NO[2][64 * kUS];
};

template <int MODE>
struct ColBuf {                  // one producer wave's half of a chunk
  uint8_t px[kHalf + 2];         // key image at rows c0 - 1 + 8 HALF + j (modes 0, 1)
  float wv[kHalf + 1];           // column couplings (modes 2, 3)
  float rx[kHalf], rn[kHalf];    // row pass X, N of the half, from base (modes 0-2)
  float g[kHalf];                // dL/d(dense) (mode 3), row c0 + 8 HALF + j
  float zx[kHalf], zy[kHalf];    // the forward column result (X, N) there (mode 3)
};


template <int MODE>
__global__ __launch_bounds__(384) void k_up_col(const FrameDev* __restrict__ frames, BlockMap bm,
                                                const float* __restrict__ wlut, int clear_zbuf) {
  __shared__ UpColLds L;
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const double lam = F.lam;
  // waves: 0 and 1 producers (half a chunk each), 2 ch, 3 x, 4 n, 5 out; waves 0
  // and 4, 1 and 5 share SIMDs
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;  // wave: uniform
  if (MODE < 2 && threadIdx.x < 256) {
    const float w = wlut[threadIdx.x];
    L.lw[threadIdx.x] = w;
    L.lm[threadIdx.x] = (float)(-lam * (double)w);
  }
  const int x = min((blockIdx.x - bm.start[f]) * 64 + lane, W - 1);  // lanes past the image repeat column W - 1
  const Chunks ch(F.col_seg, H);
  const int nch = ch.n;
  const float edge = F.edge_c ? 1.0f : 0.0f;
  const gptr<float> ck = gp((float*)F.fwd);  // checkpoints [role][chunk][x]
  const int lo = lane * kUS;
  __syncthreads();

  auto producer = [&](auto half_tag) __attribute__((always_inline)) {
    constexpr int HALF = decltype(half_tag)::value;
    constexpr int h0 = kHalf * HALF;
    const int pitch = F.pitch;
    const gptr<const float> g_wcol = gp(F.g_wcol);
    const gptr<const float> colx = gp((const float*)F.rowx) + (size_t)x * H;
    const gptr<const float> coln = colx + (size_t)W * H;
    const gptr<const float> g_out = gp(F.g_out);
    const gptr<const float2> colz = gp((const float2*)F.colxn);
    auto load = [&](ColBuf<MODE>& b, int j) __attribute__((always_inline)) {
      int c0, cnt;
      bool first, last;
      ch.get(j, c0, cnt, first, last);
      if (MODE == 3) {
        // every load unconditional, at rows clamped to the line
#pragma unroll
        for (int k = 0; k <= kHalf; ++k) b.wv[k] = g_wcol[(size_t)min(c0 + h0 + k, H) * W + x];
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
          const size_t o = (size_t)min(c0 + h0 + k, H - 1) * W + x;
          b.g[k] = g_out[o];
          const float2 z = colz[o];
          b.zx[k] = z.x;
          b.zy[k] = z.y;
        }
        return;
      }
      if (MODE == 2) {
#pragma unroll
        for (int k = 0; k <= kHalf; ++k) b.wv[k] = g_wcol[(size_t)min(c0 + h0 + k, H) * W + x];
      } else {
        // rows of the key image: uniform row pointers, the lane's column as offset
#pragma unroll
        for (int k = 0; k < kHalf + 2; ++k)
          b.px[k] = (gp(F.key) + (size_t)clampi(c0 + h0 - 1 + k, 0, H - 1) * pitch)[(unsigned)x];
      }
      const int base = min(c0 + h0, H - kHalf);
      load8(colx, base, b.rx);
      load8(coln, base, b.rn);
    };
    auto produce = [&](ColBuf<MODE>& b, int j, int slot) __attribute__((always_inline)) {
      int c0, cnt;
      bool first, last;
      ch.get(j, c0, cnt, first, last);
      const int sh = c0 + h0 - min(c0 + h0, H - kHalf);
      if (MODE == 3) {
        // adjoint right-hand sides: out = X / N, so dL/dX = g / N and
        // dL/dN = -(g / N) (X / N); pixels with N = 0 (NaN output) give none
        half_couplings<HALF>(b.wv, lam, L.M[slot] + lo, L.B[slot] + lo);
        float* Rs = L.R[slot] + lo;
        float* Ns = L.N[slot] + lo;
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
          float gx = 0.0f, gn = 0.0f;
          if (b.zy[k] != 0.0f) {
            gx = b.g[k] / b.zy[k];
            gn = -gx * (b.zx[k] / b.zy[k]);
          }
          Rs[h0 + k] = gx;
          Ns[h0 + k] = gn;
        }
        return;
      }
      if (MODE == 2)
        half_couplings<HALF>(b.wv, lam, L.M[slot] + lo, L.B[slot] + lo);
      else
        half_couplings_lut<HALF>(b.px, c0, H, edge, lam, L.lw, L.lm, L.M[slot] + lo, L.B[slot] + lo);
      store8<HALF>(L.R[slot] + lo, sh, b.rx);
      store8<HALF>(L.N[slot] + lo, sh, b.rn);
    };
    ColBuf<MODE> A, Bf;
    for (int pass = 0; pass < 2; ++pass) {
      auto chunk_at = [&](int r) __attribute__((always_inline)) { return pass == 0 ? r : nch - 1 - r; };
      const int rounds = nch + 2 + 2 * pass;  // backward: the out wave trails by two rounds
      load(A, chunk_at(0));
      for (int r = 0; r < rounds; r += 2) {
        load(Bf, chunk_at(min(r + 1, nch - 1)));
        if (r < nch) produce(A, chunk_at(r), r % kSlots);
        __syncthreads();
        if (r + 1 < rounds) {
          load(A, chunk_at(min(r + 2, nch - 1)));
          if (r + 1 < nch) produce(Bf, chunk_at(r + 1), (r + 1) % kSlots);
          __syncthreads();
        }
      }
    }
  };

  if (wave == 0) {
    producer(std::integral_constant<int, 0>{});
  } else if (wave == 1) {
    producer(std::integral_constant<int, 1>{});
  } else if (wave == 2) {
    // ---------------- ch chain ----------------
    float st = 0.0f;
    for (int r = 0; r < nch + 2; ++r) {
      const int j = r - 1;
      if (j >= 0 && j < nch) {
        int c0, cnt;
        bool first, last;
        ch.get(j, c0, cnt, first, last);
        if (first) st = 0.0f;
        ck[(size_t)j * W + x] = st;
        st = ch_chunk(L.M[j % kSlots] + lo, L.B[j % kSlots] + lo, L.C[j & 1] + lo, st);
      }
      __syncthreads();
    }
    auto ck_at = [&](int q) __attribute__((always_inline)) { return ck[(size_t)(nch - 1 - clampi(q, 0, nch - 1)) * W + x]; };
    auto back_round = [&](int q, float s0) __attribute__((always_inline)) {
      if (q >= 0 && q < nch) {
        int c0, cnt;
        bool first, last;
        ch.get(nch - 1 - q, c0, cnt, first, last);
        ch_chunk(L.M[q % kSlots] + lo, L.B[q % kSlots] + lo, L.C[q & 1] + lo, s0);
      }
    };
    const int rounds = nch + 4;
    float cA = ck_at(-1), cB;
    for (int r = 0; r < rounds; r += 2) {
      cB = ck_at(r);
      back_round(r - 1, cA);
      __syncthreads();
      if (r + 1 < rounds) {
        cA = ck_at(r + 1);
        back_round(r, cB);
        __syncthreads();
      }
    }
  } else if (wave == 3 || wave == 4) {
    // ---------------- x (wave 3) and n (wave 4) chains ----------------
    const int role = wave == 3 ? 1 : 2;
    const float* Rsrc = role == 2 ? &L.N[0][0] : &L.R[0][0];
    float* Osrc = role == 2 ? &L.NO[0][0] : &L.XO[0][0];
    const size_t ckb = (size_t)role * nch * W;
    float st = 0.0f, v[kUK];
    for (int r = 0; r < nch + 2; ++r) {
      const int j = r - 2;
      if (j >= 0) {
        int c0, cnt;
        bool first, last;
        ch.get(j, c0, cnt, first, last);
        if (first) st = 0.0f;
        ck[ckb + (size_t)j * W + x] = st;
        const int s = j % kSlots;
        st = rhs_chunk<false>(L.M[s] + lo, L.B[s] + lo, Rsrc + s * 64 * kUS + lo, L.C[j & 1] + lo, st, v);
      }
      __syncthreads();
    }
    // backward: round r recomputes chunk q = r - 2 and sweeps back through
    // chunk q - 1 (registers from round r - 1) into the out wave's slot
    float nxt = 0.0f;
    auto ck_at = [&](int q) __attribute__((always_inline)) {
      return ck[ckb + (size_t)(nch - 1 - clampi(q, 0, nch - 1)) * W + x];
    };
    float va[kUK], ca[kUK], vb[kUK], cb[kUK];
#pragma unroll
    for (int k = 0; k < kUK; ++k) va[k] = ca[k] = vb[k] = cb[k] = 0.0f;
    auto back_round = [&](int q, float s0, float (&vn)[kUK], float (&cn)[kUK], float (&vp)[kUK],
                          const float (&cp)[kUK]) __attribute__((always_inline)) {
      if (q >= 0 && q <= nch) {
        const int qc = min(q, nch - 1);
        const int s = qc % kSlots;
        int c0, cnt;
        bool first, last;
        ch.get(nch - q, c0, cnt, first, last);  // chunk q - 1 (valid when q >= 1)
        const int stop = last ? cnt - 1 : -1;
        nxt = recompute_back<false>(L.M[s] + lo, L.B[s] + lo, Rsrc + s * 64 * kUS + lo, L.C[qc & 1] + lo, s0, vn, cn,
                                    stop, nxt, vp, cp);
        if (q >= 1) {
          float* O = Osrc + ((q - 1) & 1) * 64 * kUS + lo;
#pragma unroll
          for (int k = 0; k < kUK; ++k) O[k] = vp[k];
        }
      }
    };
    const int rounds = nch + 4;
    float sA = ck_at(-2), sB;
    for (int r = 0; r < rounds; r += 2) {
      sB = ck_at(r - 1);
      back_round(r - 2, sA, va, ca, vb, cb);
      __syncthreads();
      if (r + 1 < rounds) {
        sA = ck_at(r);
        back_round(r - 1, sB, vb, cb, va, ca);
        __syncthreads();
      }
    }
  } else if (MODE == 3) {
    // ---------------- adjoint out ----------------
    // (X, N) of the adjoint solve into adjxn ([column][row]); dL/dw of every
    // edge from the solutions v and the forward results z of its two steps,
    // -scale ((vx_i - vx_{i+1}) (zx_i - zx_{i+1}) + (vn_i - vn_{i+1}) (zy_i - zy_{i+1})),
    // or -scale (vx zx + vn zy) for an edge on a segment's end; the two
    // segments' terms of the split edge are summed. Chunks come from the line's
    // end; the inputs of a round are loaded a round ahead (two buffers).
    const gptr<float2> adj = gp(F.adjxn) + (size_t)x * H;
    const gptr<const float2> colz = gp((const float2*)F.colxn);
    const gptr<const float> wcol = gp(F.g_wcol);
    const gptr<float> gw = gp(F.grad_wcol);
    const float gscale = F.grad_w_unscaled ? 1.0f : (float)lam;
    for (int r = 0; r < nch + 2; ++r) __syncthreads();
    AdjOutIn A, B;
    auto load_o = [&](AdjOutIn& b, int q) __attribute__((always_inline)) {
      int c0, cnt;
      bool first, last;
      ch.get(nch - 1 - clampi(q, 0, nch - 1), c0, cnt, first, last);
#pragma unroll
      for (int k = 0; k < kUK; ++k) {
        const float2 z = colz[(size_t)min(c0 + k, H - 1) * W + x];
        b.zx[k] = z.x;
        b.zy[k] = z.y;
        b.w[k] = wcol[(size_t)min(c0 + k + 1, H) * W + x];
      }
    };
    AdjEdgeState st;
    // the weights of the edges that start the two segments (loaded here: a
    // load inside a round would make the round wait for the next one's inputs)
    const float w_seg0 = wcol[x], w_seg1 = wcol[(size_t)min(ch.seg, H) * W + x];
    auto out_round = [&](const AdjOutIn& b, int q) __attribute__((always_inline)) {
      if (q < 0 || q >= nch) return;
      const int j = nch - 1 - q;
      int c0, cnt;
      bool first, last;
      ch.get(j, c0, cnt, first, last);
      const bool seg1 = j >= ch.n0;
      const int s0 = seg1 ? ch.seg : 0, s1 = seg1 ? ch.len : ch.seg;
      const float* XO = L.XO[q & 1] + lo;
      const float* NO = L.NO[q & 1] + lo;
      // the solutions as the lane's 16 consecutive float2 (eight float4 for a
      // whole chunk, else one by one)
      if (cnt == kUK) {
        const gptr<f4_dword_aligned> q4 = (gptr<f4_dword_aligned>)(adj + c0);
#pragma unroll
        for (int k = 0; k < kUK / 2; ++k) q4[k] = make_float4(XO[2 * k], NO[2 * k], XO[2 * k + 1], NO[2 * k + 1]);
      } else {
#pragma unroll
        for (int k = 0; k < kUK; ++k)
          if (k < cnt) adj[c0 + k] = make_float2(XO[k], NO[k]);
      }
      float ge[kUK];  // edge c0 + k + 1
      if (cnt == kUK)
        adj_edges<true>(XO, NO, b, c0, cnt, s0, s1, H, gscale, seg1 ? w_seg1 : w_seg0, st, ge);
      else
        adj_edges<false>(XO, NO, b, c0, cnt, s0, s1, H, gscale, seg1 ? w_seg1 : w_seg0, st, ge);
      // edge c0 + k + 1 of column x: one coalesced store per step
      if (gw) {
        if (cnt == kUK) {
#pragma unroll
          for (int k = 0; k < kUK; ++k) gw[(size_t)(c0 + k + 1) * W + x] = ge[k];
        } else {
#pragma unroll
          for (int k = 0; k < kUK; ++k)
            if (k < cnt) gw[(size_t)(c0 + k + 1) * W + x] = ge[k];
        }
      }
    };
    // round r forms chunk r - 4, whose inputs were loaded two rounds ahead:
    // three buffers take turns (the loop is unrolled by three without
    // conditions; its last rounds follow it)
    const int rounds = nch + 4;
    AdjOutIn C;
    load_o(A, -4);
    load_o(B, -3);
    int r = 0;
    for (; r + 2 < rounds; r += 3) {
      load_o(C, r - 2);
      out_round(A, r - 4);
      __syncthreads();
      load_o(A, r - 1);
      out_round(B, r - 3);
      __syncthreads();
      load_o(B, r);
      out_round(C, r - 2);
      __syncthreads();
    }
    if (r < rounds) {
      out_round(A, r - 4);
      __syncthreads();
      ++r;
    }
    if (r < rounds) {
      out_round(B, r - 4);
      __syncthreads();
    }
    if (gw && st.has_first) gw[x] = st.e_first;
    if (F.lam_part) F.lam_part[(size_t)H + x] = st.part;
  } else {
    // ---------------- out ----------------
    const gptr<float> dense = gp(F.dense);
    const gptr<float> dense_out = gp(F.dense_out);
    const gptr<float2> colxn = gp(F.colxn);
    const gptr<unsigned long long> zbuf = gp(F.zbuf_clear);
    for (int r = 0; r < nch + 2; ++r) {
      if (clear_zbuf && r < nch) {
        int c0, cnt;
        bool first, last;
        ch.get(r, c0, cnt, first, last);
#pragma unroll
        for (int k = 0; k < kUK; ++k)
          if (k < cnt) zbuf[(size_t)(c0 + k) * W + x] = ~0ull;
      }
      __syncthreads();
    }
    for (int r = 0; r < nch + 4; ++r) {
      const int q = r - 4;
      if (q >= 0) {
        const int j = nch - 1 - q;
        int c0, cnt;
        bool first, last;
        ch.get(j, c0, cnt, first, last);
        const float* XO = L.XO[q & 1] + lo;
        const float* NO = L.NO[q & 1] + lo;
        float xv[kUK], nv[kUK], d[kUK];
#pragma unroll
        for (int k = 0; k < kUK; ++k) {
          xv[k] = XO[k];
          nv[k] = NO[k];
          d[k] = xv[k] / nv[k];
        }
        // 32-bit offsets (W H < 2^28); the optional outputs in branches of their own
        const unsigned o0 = (unsigned)c0 * (unsigned)W + (unsigned)x;
        if (cnt == kUK) {
#pragma unroll
          for (int k = 0; k < kUK; ++k) dense[o0 + (unsigned)k * (unsigned)W] = d[k];
        } else {
#pragma unroll
          for (int k = 0; k < kUK; ++k)
            if (k < cnt) dense[o0 + (unsigned)k * (unsigned)W] = d[k];
        }
        if (colxn) {
#pragma unroll
          for (int k = 0; k < kUK; ++k)
            if (k < cnt) colxn[o0 + (unsigned)k * (unsigned)W] = make_float2(xv[k], nv[k]);
        }
        if (dense_out) {
#pragma unroll
          for (int k = 0; k < kUK; ++k)
            if (k < cnt) dense_out[o0 + (unsigned)k * (unsigned)W] = d[k];
        }
      }
      __syncthreads();
    }
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
// z-buffer key of a warped seed: (float bits of its depth, NaN as one
// canonical NaN) above its source index, so that the minimum is the nearest
// seed and, among equal depths, the lower source
__device__ __forceinline__ unsigned long long warp_key(float nd, int idx) {
  uint32_t bits = __float_as_uint(nd);
  if (nd != nd) bits = 0x7FC00000u;
  return ((unsigned long long)bits << 32) | (unsigned long long)(uint32_t)idx;
}
__device__ __forceinline__ float grown_variance(float s2, float nd) { return (float)((double)s2 + (double)nd * 0.01); }

// The source seed of pixel idx after the warp: its own z-buffer winner, else
// (hole filling, src/depth_fusion.cu:121-144) the first non-empty neighbour
// in raster order whose position is in [2, W-2] x [2, H-2]; 0 for none. key
// receives the z-buffer key the source came from.
__device__ __forceinline__ int fused_source(const FrameDev& F, int idx, unsigned long long& key) {
  const int W = F.W, H = F.H;
  const int y = idx / W, x = idx - y * W;
  const unsigned long long empty = ~0ull;
  key = F.zbuf[idx];
  int src = key == empty ? 0 : (int)(uint32_t)(key & 0xFFFFFFFFull);
  if (src == 0) {
    for (int sy = y - 1; sy <= y + 1 && src == 0; ++sy)
      for (int sx = x - 1; sx <= x + 1; ++sx) {
        if (sx == x && sy == y) continue;
        if (sx < 2 || sx > W - 2 || sy < 2 || sy > H - 2) continue;
        const unsigned long long kk = F.zbuf[(size_t)sy * W + sx];
        if (kk != empty) {
          src = (int)(uint32_t)(kk & 0xFFFFFFFFull);
          key = kk;
          break;
        }
      }
  }
  return src;
}

// The warp of seed idx (src/depth_fusion.cu:48-119): false when the seed is
// bad or leaves the image; else s becomes the warped seed (depth nd, variance
// grown by 0.01 nd) and (pxi, pyi) its pixel.
__device__ __forceinline__ bool warp_seed(const FrameDev& F, int idx, float4& s, int& pxi, int& pyi) {
  const int W = F.W, H = F.H;
  const int y = idx / W, x = idx - y * W;
  if (is_badpoint(s.x, s.y)) return false;
  const V3 dir = ray_dir(x, y, F.fx, F.fy, F.cx, F.cy);
  V3 P;
  P.x = dir.x * s.z;
  P.y = dir.y * s.z;
  P.z = dir.z * s.z;
  const V3 q = se3_apply(F.T_fuse, P);
  const float nd = sqrtf((q.x * q.x + q.y * q.y) + q.z * q.z);
  if ((double)nd <= 0.5) return false;
  const float pxf = (F.fx * q.x) / q.z + F.cx;
  const float pyf = (F.fy * q.y) / q.z + F.cy;
  pxi = trunc_i32((double)pxf + 0.5);
  pyi = trunc_i32((double)pyf + 0.5);
  if (pxi >= W || pxi < 0 || pyi >= H || pyi < 0) return false;
  s.z = nd;
  s.w = grown_variance(s.w, nd);
  return true;
}

// The warp reads each prior seed once (a streaming load) and writes the
// warped seed back in place (the reference's layout) unless warp_skip: then
// the update rebuilds the winner's warped depth and variance from its
// z-buffer key and its prior seed, and the seeds stay as they were.
__global__ __launch_bounds__(256) void k_fuse_warp(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int idx = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  if (idx >= F.W * F.H) return;
  float4 s = load4_nt(gp((const float4*)F.seeds) + idx);
  int pxi, pyi;
  if (!warp_seed(F, idx, s, pxi, pyi)) return;
  if (!F.warp_skip) F.seeds[idx] = s;
  if (idx != 0) atomicMin(&F.zbuf[(size_t)pyi * F.W + pxi], warp_key(s.z, idx));
}

// Replays, for a batch run that skipped them, the warped seeds (in place) and
// the hole-filled warp table, from the run's prior seeds and its intact
// z-buffer: exactly what k_fuse_warp and k_fuse_update would have written.
__global__ __launch_bounds__(256) void k_fuse_replay(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int idx = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  if (idx >= F.W * F.H) return;
  unsigned long long key;
  F.table_out[idx] = fused_source(F, idx, key);
  float4 s = F.seeds[idx];
  int pxi, pyi;
  if (warp_seed(F, idx, s, pxi, pyi)) F.seeds[idx] = s;
}

// Hole filling from the z-buffer snapshot (an empty cell takes the first
// non-empty source in raster order whose position is in [2, W-2] x [2, H-2],
// src/depth_fusion.cu:121-144) and the per-pixel update
// (src/depth_fusion.cu:146-213). The dense depth, the source seed and both
// outputs are touched once per pixel here: streaming accesses.
__global__ __launch_bounds__(256) void k_fuse_update(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int idx = (blockIdx.x - bm.start[f]) * 256 + threadIdx.x;
  const int W = F.W, H = F.H;
  if (idx >= W * H) return;
  unsigned long long key;
  const int src = fused_source(F, idx, key);
  if (F.table_out) F.table_out[idx] = src;
  const float de = __builtin_nontemporal_load(&gp(F.dense)[idx]);
  float unc = fmaxf(0.5f, (float)((double)de * 0.2));
  unc = unc * unc;
  if (de <= 0.0f) unc = 1e9f;
  float a, b, mu, s2;
  if (src > 0) {
    float4 s = load4_nt(gp((const float4*)F.seeds) + src);
    if (F.warp_skip) {
      // the warp left the prior as it was: its depth is the key's, its variance grows by it
      s.z = __uint_as_float((uint32_t)(key >> 32));
      s.w = grown_variance(s.w, s.z);
    }
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
  if (F.depth_out) __builtin_nontemporal_store(is_goodpoint(a, b) ? mu_p : -1.0f, &gp(F.depth_out)[idx]);
  store4_nt(gp(F.seeds_new) + idx, make_float4(a_p, b_p, mu_p, s_p));
}

// ---------------------------------------------------------------------------
// rows_map / cols_map: the Thomas kernels' blocks (64 lines each) for the
// frames with only[i] != 0 (every frame when only is null)
static int rows_map(BlockMap& bm, const FrameDev* hframes, int n, const int* only = nullptr) {
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (!only || only[i]) ? (hframes[i].H + 63) / 64 : 0;
  return make_block_map(bm, counts, n);
}
static int cols_map(BlockMap& bm, const FrameDev* hframes, int n, const int* only = nullptr) {
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (!only || only[i]) ? (hframes[i].W + 63) / 64 : 0;
  return make_block_map(bm, counts, n);
}
static int pixels_map(BlockMap& bm, const FrameDev* hframes, int n) {
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].W * hframes[i].H + 255) / 256;
  return make_block_map(bm, counts, n);
}

void launch_upsample_row(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         hipStream_t s, int solver, bool clear_zbuf) {
  // the segmented-scan solver (modes 0 and 1) takes the frames whose rows fit
  // its lanes; the Thomas kernels the rest. clear_zbuf: the scan resets the
  // next run's fusion z-buffer when it takes every frame (row_scan_covers)
  int rest[kMaxFrames];
  const bool scan = solver == QM_UPSAMPLE_SCAN && mode <= 1;
  if (scan && launch_upsample_row_scan(dframes, hframes, n, mode, wlut, s, rest, clear_zbuf) == 0) return;
  BlockMap bm;
  const int nb = rows_map(bm, hframes, n, scan ? rest : nullptr);
  if (nb <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL(k_up_row<0>, dim3(nb), dim3(320), 0, s, dframes, bm, wlut);
  else if (mode == 1)
    hipLaunchKernelGGL(k_up_row<1>, dim3(nb), dim3(320), 0, s, dframes, bm, wlut);
  else if (mode == 2)
    hipLaunchKernelGGL(k_up_row<2>, dim3(nb), dim3(320), 0, s, dframes, bm, wlut);
  else
    hipLaunchKernelGGL(k_up_row<3>, dim3(nb), dim3(384), 0, s, dframes, bm, wlut);
}

void launch_adj_dconf(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) {
    const FrameDev& F = hframes[i];
    counts[i] = F.adj_rowv && (F.grad_d || F.grad_conf) ? ((F.W + 63) / 64) * ((F.H + 63) / 64) : 0;
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_adj_dconf, dim3(nb), dim3(256), 0, s, dframes, bm);
}

void launch_upsample_col(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         bool clear_zbuf, hipStream_t s, int solver) {
  int rest[kMaxFrames];
  const bool scan = solver == QM_UPSAMPLE_SCAN && mode <= 1;
  if (scan && launch_upsample_col_scan(dframes, hframes, n, wlut, clear_zbuf, s, rest) == 0) return;
  BlockMap bm;
  const int nb = cols_map(bm, hframes, n, scan ? rest : nullptr);
  if (nb <= 0) return;
  if (mode == 2)
    hipLaunchKernelGGL(k_up_col<2>, dim3(nb), dim3(384), 0, s, dframes, bm, wlut, (int)clear_zbuf);
  else if (mode == 3)
    hipLaunchKernelGGL(k_up_col<3>, dim3(nb), dim3(384), 0, s, dframes, bm, wlut, 0);
  else
    hipLaunchKernelGGL(k_up_col<0>, dim3(nb), dim3(384), 0, s, dframes, bm, wlut, (int)clear_zbuf);
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

void launch_fuse_replay(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  const int nb = pixels_map(bm, hframes, n);
  if (nb > 0) hipLaunchKernelGGL(k_fuse_replay, dim3(nb), dim3(256), 0, s, dframes, bm);
}

void launch_upsample(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                     bool clear_zbuf, hipStream_t s, int solver) {
  launch_upsample_row(dframes, hframes, n, mode, wlut, s, solver);
  launch_upsample_col(dframes, hframes, n, mode, wlut, clear_zbuf, s, solver);
}

void launch_fuse(const FrameDev* dframes, const FrameDev* hframes, int n, bool clear_zbuf, hipStream_t s) {
  if (clear_zbuf) launch_fill_zbuf(dframes, hframes, n, s);
  launch_fuse_warp(dframes, hframes, n, s);
  launch_fuse_update(dframes, hframes, n, s);
}

}  // namespace qm
