// k_scan.hip -- the upsample's line systems by segmented scans
// (qm_params.upsample_solver = QM_UPSAMPLE_SCAN). Same systems as the Thomas
// kernels of k_back.hip (reference: smooth_row / smooth_col,
// src/depth_upsample.cu:94-302): per line A = I + lambda L(w) with two
// right-hand sides, a_k = float(-lambda w(k)), c_k = float(-lambda w(k+1)),
// b_k = float(1 + lambda (w(k) + w(k+1))), the chain split at row_seg /
// col_seg, then dense = X / N after the column pass.
//
// P lanes of a wave solve one line (32 lanes for lines of up to 512 steps, two
// lines per wave; else 64); lane i runs steps [i m, (i + 1) m) with
// m = ceil(n / P), held in registers, steps past n padding the line. Each of the three recurrences of the
// Thomas solve is a map of its previous state:
//   ch   ch_k = c_k / (b_k - ch_{k-1} a_k), a Moebius map: the 2 x 2 matrix
//        [[0, c_k], [-a_k, b_k]] on (p, q), ch = p / q;
//   x, n dh_k = (r_k - dh_{k-1} a_k) y_k with y_k = 1 / den_k: affine; a lane
//        keeps each step's value from a zero start and its slope, and forms
//        dh_k = fma(slope, incoming, value) once the scan has given the
//        incoming state;
//   X, N X_k = dh_k - ch_k X_{k+1}, backward: affine, the same way.
// A lane composes its steps' map (in double), a scan over the 64 lanes
// (double; DPP moves inside rows of 16 lanes, lane broadcasts across rows)
// gives each lane the composition of all lanes before it (after it, for the
// backward map); the lane then runs its steps of the coefficient chain in
// float from the incoming state and combines the affine ones with it.
// A chain start (0 and seg) takes a_k = 0 in the recurrences, a chain end
// (seg - 1 and n - 1) ch_k = 0 in the backward one. Matrices are rescaled by a
// power of two (exact) every 8 steps of a lane and after every product of the
// scan. y_k is the correctly rounded reciprocal. The whole arithmetic is fixed,
// and the oracle's solve_line_scan (oracle/qm_oracle.cpp) restates it
// operation for operation; it is a different rounding of the same solve as
// the Thomas order, as accurate against FP64.
//
// Work is lane-parallel instead of a serial chain per line: a 752-pixel row
// takes 6 dependent scan rounds and 4 x 12 lane steps instead of 2 x 752
// chain steps, so the passes are no longer bound by the chains' latency.
// Results leave through an LDS transpose: the row pass writes the column-major
// X / N planes the column pass reads (a lane's segment of a column is then
// contiguous), the column pass writes dense row-major.
#include <type_traits>

#include "qm_device.h"
#include "qm_launch.h"

namespace qm {


__device__ __forceinline__ double absd(double v) { return v < 0.0 ? -v : v; }

// (a, b; c, d) * 2^-e, e the binary exponent of the largest |entry|: exact
__device__ __forceinline__ void mob_norm(double& a, double& b, double& c, double& d) {
  double mx = absd(a), t = absd(b);
  mx = t > mx ? t : mx;
  t = absd(c);
  mx = t > mx ? t : mx;
  t = absd(d);
  mx = t > mx ? t : mx;
  if (!(mx > 0.0) || !(mx <= 1.7976931348623157e308)) return;
  int e;
  (void)frexp(mx, &e);
  a = ldexp(a, -e);
  b = ldexp(b, -e);
  c = ldexp(c, -e);
  d = ldexp(d, -e);
}

__device__ __forceinline__ double shfl_up_d(double v, int o) { return __shfl_up(v, (unsigned)o, 64); }
__device__ __forceinline__ double shfl_down_d(double v, int o) { return __shfl_down(v, (unsigned)o, 64); }

// DPP moves of a double inside rows of 16 lanes (a lane whose source lies
// outside its row keeps its own value) and a lane's double broadcast
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  const unsigned l2 = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const unsigned h2 = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)h2 << 32) | l2));
}
__device__ __forceinline__ double lane_d(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// row_shr:o (source lane - o) and row_shl:o (source lane + o)
template <int O>
__device__ __forceinline__ double row_shr_d(double v) { return dpp_d<0x110 + O>(v); }
template <int O>
__device__ __forceinline__ double row_shl_d(double v) { return dpp_d<0x100 + O>(v); }

// The scans over the lanes of a line, one fixed tree (restated by the
// oracle's scan_tree): Hillis-Steele inside each row of 16 lanes (offsets 1,
// 2, 4, 8), then rows 1 and 3 combine with the last lane of rows 0 and 2,
// then (64-lane lines) rows 2 and 3 with lane 31. The suffix scan mirrors it:
// rows 0 and 2 combine with the first lane of rows 1 and 3, then rows 0 and 1
// with lane 32. A wave holds two 32-lane lines (lanes 0-31 and 32-63) or one
// 64-lane line. comb(v, h) combines this lane's value v with h, the value of lanes
// before it (prefix) or after it (suffix); move(v, O) is the DPP row shift by
// O (right for the prefix, left for the suffix).
template <int P, typename Vals, typename Move, typename Comb>
__device__ __forceinline__ void scan_tree(int lane, bool suffix, Vals& v, Move move, Comb comb) {
  const int j = lane & 15, row = lane >> 4;
  {
    Vals h = move(v, std::integral_constant<int, 1>{});
    if (suffix ? j + 1 < 16 : j >= 1) comb(v, h);
  }
  {
    Vals h = move(v, std::integral_constant<int, 2>{});
    if (suffix ? j + 2 < 16 : j >= 2) comb(v, h);
  }
  {
    Vals h = move(v, std::integral_constant<int, 4>{});
    if (suffix ? j + 4 < 16 : j >= 4) comb(v, h);
  }
  {
    Vals h = move(v, std::integral_constant<int, 8>{});
    if (suffix ? j + 8 < 16 : j >= 8) comb(v, h);
  }
  {
    Vals a = v.lane(suffix ? 16 : 15), b = v.lane(suffix ? 48 : 47);
    if (suffix ? (row & 1) == 0 : (row & 1) == 1) comb(v, row < 2 ? a : b);
  }
  if (P == 64) {
    Vals a = v.lane(suffix ? 32 : 31);
    if (suffix ? row < 2 : row >= 2) comb(v, a);
  }
}

struct Mob4 {
  double g11, g12, g21, g22;
  __device__ Mob4 lane(int l) const { return {lane_d(g11, l), lane_d(g12, l), lane_d(g21, l), lane_d(g22, l)}; }
};

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
struct Aff3 {
  float u, w, v;  // two intercepts sharing one slope
  __device__ Aff3 lane(int l) const { return {lane_f(u, l), lane_f(w, l), lane_f(v, l)}; }
};

// RN(1 / z): the hardware reciprocal refined by one FMA Newton step equals the
// IEEE quotient for every normal z in [2^-126, 2^126) (checked over all of
// them on gfx950). The pivots here lie in [1, 1 + 2 lambda]: frames with
// lambda above 2^100 run the Thomas kernels.
__device__ __forceinline__ float recip_exact(float z) {
  const float y = __builtin_amdgcn_rcpf(z);
  return fmaf(fmaf(-z, y, 1.0f), y, y);
}

// One line of n steps, chain split at seg: this lane's m steps s .. s + m - 1
// (m the same for every lane of the wave; steps past n pad the line: am = 0
// beyond the step n, bq = 1, zero right-hand sides) with raw couplings am[k] =
// float(-lambda w(s + k)) (k = 0 .. m), diagonal bq[k] and right-hand sides rx,
// rn; returns X, N of its steps.
template <int P, int M>
__device__ __forceinline__ void scan_line(int lane, int s, int m, int n, int seg, const float (&am)[M + 1],
                                          const float (&bq)[M], const float (&rx)[M], const float (&rn)[M],
                                          float (&X)[M], float (&Nn)[M]) {
  const int sg = seg < n ? seg : n;
  float ab[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int j = s + k;
    ab[k] = (j == 0 || j == sg || j >= n) ? 0.0f : am[k];
  }
  // ---- ch: lane map (rescaled every 8 steps and at the end), scan, steps ----
  Mob4 G{1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < m) {
      const double c = (double)am[k + 1], b = (double)bq[k], a = (double)ab[k];
      const double q11 = c * G.g21, q12 = c * G.g22;
      const double q21 = b * G.g21 - a * G.g11, q22 = b * G.g22 - a * G.g12;
      G = {q11, q12, q21, q22};
      if ((k & 7) == 7) mob_norm(G.g11, G.g12, G.g21, G.g22);
    }
  mob_norm(G.g11, G.g12, G.g21, G.g22);
  const int ll = lane & (P - 1);  // lane in the line
  scan_tree<P>(
      lane, false, G,
      [](const Mob4& g, auto o) {
        constexpr int O = decltype(o)::value;
        return Mob4{row_shr_d<O>(g.g11), row_shr_d<O>(g.g12), row_shr_d<O>(g.g21), row_shr_d<O>(g.g22)};
      },
      [](Mob4& g, const Mob4& h) {
        const double r11 = g.g11 * h.g11 + g.g12 * h.g21, r12 = g.g11 * h.g12 + g.g12 * h.g22;
        const double r21 = g.g21 * h.g11 + g.g22 * h.g21, r22 = g.g21 * h.g12 + g.g22 * h.g22;
        g = {r11, r12, r21, r22};  // not rescaled: six levels stay far inside the double range
      });
  float st;
  {
    const double p12 = shfl_up_d(G.g12, 1), p22 = shfl_up_d(G.g22, 1);
    st = ll == 0 ? 0.0f : (float)(p12 / p22);
  }
  float Y[M], CH[M];
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < m) {
      const float den = bq[k] - st * ab[k];
      Y[k] = recip_exact(den);
      st = am[k + 1] * Y[k];
      CH[k] = st;
    }
  // ---- x, n: each step's value from a zero start (u_k) and slope (V_k) of
  // the lane's affine map, scan of the lanes' maps, dh_k = fma(V_k, in, u_k) ----
  float Uu[M], Uw[M], Vs[M];
  Aff3 U{0.0f, 0.0f, 1.0f};
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < m) {
      U.u = (rx[k] - U.u * ab[k]) * Y[k];
      U.w = (rn[k] - U.w * ab[k]) * Y[k];
      U.v = -(ab[k] * Y[k]) * U.v;
      Uu[k] = U.u;
      Uw[k] = U.w;
      Vs[k] = U.v;
    }
  auto aff_comb = [](Aff3& g, const Aff3& h) {
    g.u = g.u + g.v * h.u;
    g.w = g.w + g.v * h.w;
    g.v = g.v * h.v;
  };
  scan_tree<P>(
      lane, false, U,
      [](const Aff3& g, auto o) {
        constexpr int O = decltype(o)::value;
        return Aff3{dpp_f<0x110 + O>(g.u), dpp_f<0x110 + O>(g.w), dpp_f<0x110 + O>(g.v)};
      },
      aff_comb);
  float dx, dn;
  {
    const float px = __shfl_up(U.u, 1u, 64), pn = __shfl_up(U.w, 1u, 64);
    dx = ll == 0 ? 0.0f : px;
    dn = ll == 0 ? 0.0f : pn;
  }
  float DH[M], NH[M];
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < m) {
      DH[k] = fmaf(Vs[k], dx, Uu[k]);
      NH[k] = fmaf(Vs[k], dn, Uw[k]);
    }
  // ---- X, N: the same backward from the lane's end, suffix scan,
  // X_k = fma(Q_k, in, p_k) ----
  float cc[M];
#pragma unroll
  for (int k = 0; k < M; ++k) cc[k] = (s + k == sg - 1 || s + k == n - 1) ? 0.0f : CH[k];
  Aff3 Pm{0.0f, 0.0f, 1.0f};
#pragma unroll
  for (int k = M - 1; k >= 0; --k)
    if (k < m) {
      Pm.u = DH[k] - cc[k] * Pm.u;
      Pm.w = NH[k] - cc[k] * Pm.w;
      Pm.v = -cc[k] * Pm.v;
      Uu[k] = Pm.u;
      Uw[k] = Pm.w;
      Vs[k] = Pm.v;
    }
  scan_tree<P>(
      lane, true, Pm,
      [](const Aff3& g, auto o) {
        constexpr int O = decltype(o)::value;
        return Aff3{dpp_f<0x100 + O>(g.u), dpp_f<0x100 + O>(g.w), dpp_f<0x100 + O>(g.v)};
      },
      aff_comb);
  float xe, ne;
  {
    const float nx = __shfl_down(Pm.u, 1u, 64), nn = __shfl_down(Pm.w, 1u, 64);
    xe = ll == P - 1 ? 0.0f : nx;
    ne = ll == P - 1 ? 0.0f : nn;
  }
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < m) {
      X[k] = fmaf(Vs[k], xe, Uu[k]);
      Nn[k] = fmaf(Vs[k], ne, Uw[k]);
    }
}

// A lane's grey levels px[k] = I(s - 1 + k), k = 0 .. M + 1, clamped to the
// line, loaded together; the empty asm takes them as operands, so every load
// is issued up front rather than sunk into a branch of the lanes that use it.
template <int M, typename PixFn>
__device__ __forceinline__ void lane_pixels(int s, int n, PixFn pix, int (&px)[M + 2]) {
#pragma unroll
  for (int k = 0; k < M + 2; ++k) px[k] = pix(min(max(s - 1 + k, 0), n - 1));
#pragma unroll
  for (int k = 0; k < M + 2; ++k) asm volatile("" : "+v"(px[k]));
}

// Couplings of a lane's m steps from its grey levels: w(0) = 0, w(n) = edge,
// else the table entry of |I(j) - I(j - 1)|; am = float(-lambda w), and b;
// past the step n (padding) am = 0 and b = 1.
template <int M>
__device__ __forceinline__ void lane_couplings(int s, int n, float edge, double lam, const float* lw,
                                               const float* lm, const int (&px)[M + 2], float (&am)[M + 1],
                                               float (&bq)[M]) {
  const float a0 = (float)(-lam * 0.0), ae = (float)(-lam * (double)edge);
  float wt[M + 1], at[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) {
    const int d = px[k + 1] > px[k] ? px[k + 1] - px[k] : px[k] - px[k + 1];
    wt[k] = lw[d];
    at[k] = lm[d];
  }
  // every lane reads its table entries (selects below, not branches around the reads)
#pragma unroll
  for (int k = 0; k <= M; ++k) asm volatile("" : "+v"(wt[k]), "+v"(at[k]));
  float wv[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) {
    const int j = s + k;
    wv[k] = j == 0 ? 0.0f : j == n ? edge : j > n ? 0.0f : wt[k];
    am[k] = j == 0 ? a0 : j == n ? ae : j > n ? 0.0f : at[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) bq[k] = s + k < n ? (float)(1.0 + lam * (double)(wv[k] + wv[k + 1])) : 1.0f;
}

// ---------------------------------------------------------------------------
// Row pass: a workgroup of WAVES waves solves LINES = WAVES * 64 / P rows, each
// row on P lanes of a wave. MODE 0 composes the sparse map (BP winner depth at
// node pixels (4x, 4y), else the pre-fill, else 0) and writes it to
// sparse_out when asked; MODE 1 reads an explicit sparse map. X and N leave
// through the LDS tile [x][row] to the planes rowx[x][y] and rowx[W H + x][y].
// ---------------------------------------------------------------------------
template <int P, int M, int WAVES, int MODE>
__global__ __launch_bounds__(64 * WAVES) void k_up_row_scan(const FrameDev* __restrict__ frames, BlockMap bm,
                                                            const float* __restrict__ wlut, int clear_zbuf) {
  constexpr int LINES = WAVES * 64 / P;
  __shared__ float lw[256], lm[256];
  __shared__ float tile[P * M * (LINES + 1)];
  const int f = __builtin_amdgcn_readfirstlane(find_frame(bm, blockIdx.x));  // uniform: scalar loads of F
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H, Wc = F.Wc, Hc = F.Hc;
  const double lam = F.lam;
  if (clear_zbuf) {
    // the next run's fusion z-buffer over this workgroup's rows, reset to
    // empty as whole lines: 16-byte streaming stores (an odd first or last
    // entry on its own)
    typedef unsigned U4v __attribute__((ext_vector_type(4)));
    const gptr<unsigned long long> z = gp(F.zbuf_clear);
    const int ya = (blockIdx.x - bm.start[f]) * LINES;
    size_t a = (size_t)ya * W;
    size_t e = (size_t)min(ya + LINES, H) * W;
    if ((a & 1) && a < e) {
      if (threadIdx.x == 0) z[a] = ~0ull;
      ++a;
    }
    if (((e - a) & 1) && threadIdx.x == 0) z[e - 1] = ~0ull;
    const size_t pairs = (e - a) >> 1;
    const gptr<U4v> z2 = (gptr<U4v>)(z + a);
    const U4v ones = {~0u, ~0u, ~0u, ~0u};
    for (size_t k = threadIdx.x; k < pairs; k += 64 * WAVES) __builtin_nontemporal_store(ones, z2 + k);
  }
  for (int t = threadIdx.x; t < 256; t += 64 * WAVES) {
    const float w = wlut[t];
    lw[t] = w;
    lm[t] = (float)(-lam * (double)w);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r = (threadIdx.x >> 6) * (64 / P) + lane / P;  // this lane's row in the workgroup
  const int y0 = (blockIdx.x - bm.start[f]) * LINES;
  const int y = min(y0 + r, H - 1);  // rows past the image repeat row H - 1 and are not stored
  const int m = (W + P - 1) / P;     // steps per lane; lane steps at x >= W pad the row
  const int s = (lane & (P - 1)) * m;
  const gptr<const uint8_t> img = gp(F.key + (size_t)y * F.pitch);
  int px[M + 2];
  lane_pixels<M>(s, W, [&](int j) { return (int)img[j]; }, px);
  float am[M + 1], bq[M];
  lane_couplings<M>(s, W, F.edge_r ? 1.0f : 0.0f, lam, lw, lm, px, am, bq);
  float rx[M], rn[M];
  const size_t row = (size_t)y * W;
  if (MODE == 1) {
    const gptr<const float> sp = gp(F.sparse_in) + row;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const float v = sp[min(s + k, W - 1)];
      rx[k] = s + k < W ? v : 0.0f;
    }
  } else {
    const bool have_prefill = F.prefill != nullptr;
    const gptr<const float> pf = gp(have_prefill ? F.prefill : (const float*)F.dense) + row;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const float v = pf[min(s + k, W - 1)];
      rx[k] = (have_prefill && s + k < W) ? v : 0.0f;
    }
    if ((y & 3) == 0 && (y >> 2) < Hc) {
      const size_t cb = (size_t)(y >> 2) * Wc;
      const gptr<const int32_t> label = gp((const int32_t*)F.label) + cb;
      const gptr<const float> dcoarse = gp((const float*)F.dcoarse) + cb;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int j = s + k, cx = j >> 2;
        if (k < m && (j & 3) == 0 && cx < Wc && label[cx] >= 0) rx[k] = dcoarse[cx];
      }
    }
    if (F.sparse_out && y0 + r < H) {
      const gptr<float> so = gp(F.sparse_out) + row;
#pragma unroll
      for (int k = 0; k < M; ++k)
        if (k < m && s + k < W) so[s + k] = rx[k];
    }
  }
#pragma unroll
  for (int k = 0; k < M; ++k) rn[k] = rx[k] > 0.0f ? 1.0f : 0.0f;
  float X[M], Nn[M];
  scan_line<P, M>(lane, s, m, W, F.row_seg, am, bq, rx, rn, X, Nn);
  const int rows = min(LINES, H - y0);
  const gptr<float> rowx = gp(F.rowx);
#pragma unroll
  for (int plane = 0; plane < 2; ++plane) {
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (k < m && s + k < W) tile[(s + k) * (LINES + 1) + r] = plane == 0 ? X[k] : Nn[k];
    __syncthreads();
    const gptr<float> dst = rowx + (plane == 0 ? (size_t)0 : (size_t)W * H) + y0;
    for (int i = threadIdx.x; i < W * LINES; i += 64 * WAVES) {
      const int x = i / LINES, rr = i - x * LINES;
      if (rr < rows) dst[(size_t)x * H + rr] = tile[x * (LINES + 1) + rr];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Column pass: a workgroup of WAVES waves solves LINES = WAVES * 64 / P
// columns, each on P lanes of a wave; the right-hand sides are the row pass's
// planes, read along the column (a lane's steps are contiguous). dense = X / N
// (NaN where N = 0, as the reference) leaves through the LDS tile [y][column]
// row-major, with dense_out and the fusion z-buffer reset (clear_zbuf) beside
// it; the column results (X, N) go to colxn when asked.
// ---------------------------------------------------------------------------
template <int P, int M, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_up_col_scan(const FrameDev* __restrict__ frames, BlockMap bm,
                                                            const float* __restrict__ wlut, int clear_zbuf) {
  constexpr int LINES = WAVES * 64 / P;
  __shared__ float lw[256], lm[256];
  __shared__ float tile[P * M * (LINES + 1)];
  // neighbouring workgroups read the same image lines: keep them on one XCD's L2
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int f = __builtin_amdgcn_readfirstlane(find_frame(bm, lb));  // uniform: scalar loads of F
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const double lam = F.lam;
  for (int t = threadIdx.x; t < 256; t += 64 * WAVES) {
    const float w = wlut[t];
    lw[t] = w;
    lm[t] = (float)(-lam * (double)w);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int c = (threadIdx.x >> 6) * (64 / P) + lane / P;  // this lane's column in the workgroup
  const int x0 = (lb - bm.start[f]) * LINES;
  const int x = min(x0 + c, W - 1);  // columns past the image repeat column W - 1 and are not stored
  const int m = (H + P - 1) / P;     // steps per lane; lane steps at y >= H pad the column
  const int s = (lane & (P - 1)) * m;
  const gptr<const uint8_t> key = gp(F.key) + x;
  const int pitch = F.pitch;
  int pv[M + 2];
  lane_pixels<M>(s, H, [&](int j) { return (int)key[(size_t)j * pitch]; }, pv);
  float am[M + 1], bq[M];
  lane_couplings<M>(s, H, F.edge_c ? 1.0f : 0.0f, lam, lw, lm, pv, am, bq);
  float rx[M], rn[M];
  const gptr<const float> px = gp((const float*)F.rowx) + (size_t)x * H;
  const gptr<const float> pn = px + (size_t)W * H;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int j = min(s + k, H - 1);
    const float a = px[j], b = pn[j];
    rx[k] = s + k < H ? a : 0.0f;
    rn[k] = s + k < H ? b : 0.0f;
  }
  float X[M], Nn[M];
  scan_line<P, M>(lane, s, m, H, F.col_seg, am, bq, rx, rn, X, Nn);
  const bool col_ok = x0 + c < W;
  if (F.colxn && col_ok) {
    const gptr<float2> cz = gp(F.colxn);
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (k < m && s + k < H) cz[(size_t)(s + k) * W + x] = make_float2(X[k], Nn[k]);
  }
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < m && s + k < H) tile[(s + k) * (LINES + 1) + c] = X[k] / Nn[k];
  __syncthreads();
  const int cols = min(LINES, W - x0);
  const gptr<float> dense = gp(F.dense) + x0;
  const gptr<float> dense_out = gp(F.dense_out);
  const gptr<unsigned long long> zbuf = gp(F.zbuf_clear) + x0;
  for (int i = threadIdx.x; i < H * LINES; i += 64 * WAVES) {
    const int yy = i / LINES, cc = i - yy * LINES;
    if (cc < cols) {
      const float v = tile[yy * (LINES + 1) + cc];
      const size_t o = (size_t)yy * W + cc;
      dense[o] = v;
      if (dense_out) dense_out[o + x0] = v;
      if (clear_zbuf) zbuf[o] = ~0ull;
    }
  }
}

// ---------------------------------------------------------------------------
// Launches. Lines of up to 512 steps run on 32 lanes, lines of up to 1536 on
// 64 (the oracle's solve_line_by follows the same rule); a frame's rows
// (columns) go to the kernel of the smallest shape that holds them: 12 or 16
// steps per lane on 32 lanes, 12 or 24 on 64. Longer lines, and frames with
// lambda above 2^100, are left to the Thomas kernels (rest[i] = 1, the return
// value counts them).
// ---------------------------------------------------------------------------
// The waves per workgroup (which do not change the arithmetic) measured on
// MI355X: the row pass is fastest with many rows per workgroup (its transpose
// writes 64 B runs of a column plane), the column pass with four waves.
struct ScanShape {
  int P, M, waves;
};
constexpr ScanShape kRowShapes[4] = {{32, 12, 16}, {32, 16, 8}, {64, 12, 16}, {64, 24, 8}};
constexpr ScanShape kColShapes[4] = {{32, 12, 4}, {32, 16, 4}, {64, 12, 4}, {64, 24, 4}};

static int scan_shape_of(int n) {
  const int P = n <= 512 ? 32 : 64;
  for (int i = 0; i < 4; ++i)
    if (kRowShapes[i].P == P && (n + P - 1) / P <= kRowShapes[i].M) return i;
  return -1;
}

template <int I, int MODE>
static void row_launch(const FrameDev* dframes, const BlockMap& bm, int nb, const float* wlut, int clear,
                       hipStream_t s) {
  constexpr ScanShape sh = kRowShapes[I];
  hipLaunchKernelGGL((k_up_row_scan<sh.P, sh.M, sh.waves, MODE>), dim3(nb), dim3(64 * sh.waves), 0, s, dframes, bm,
                     wlut, clear);
}
template <int I>
static void col_launch(const FrameDev* dframes, const BlockMap& bm, int nb, const float* wlut, int clear,
                       hipStream_t s) {
  constexpr ScanShape sh = kColShapes[I];
  hipLaunchKernelGGL((k_up_col_scan<sh.P, sh.M, sh.waves>), dim3(nb), dim3(64 * sh.waves), 0, s, dframes, bm, wlut,
                     clear);
}

// The row scan takes a frame whose rows fit its lanes and whose lambda is in
// its range; row_scan_covers: it takes every frame of the launch.
static bool row_scan_takes(const FrameDev& h) { return scan_shape_of(h.W) >= 0 && h.lam <= 0x1p100; }
bool row_scan_covers(const FrameDev* hframes, int n) {
  for (int i = 0; i < n; ++i)
    if (!row_scan_takes(hframes[i])) return false;
  return true;
}

int launch_upsample_row_scan(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                             hipStream_t s, int* rest, bool clear_zbuf) {
  int left = 0;
  for (int i = 0; i < n; ++i) {
    rest[i] = row_scan_takes(hframes[i]) ? 0 : 1;
    left += rest[i];
  }
  // the z-buffer reset moves here only when every frame is the scan's
  const int cz = clear_zbuf && left == 0 ? 1 : 0;
  for (int sh = 0; sh < 4; ++sh) {
    const int lines = kRowShapes[sh].waves * 64 / kRowShapes[sh].P;
    int counts[kMaxFrames];
    for (int i = 0; i < n; ++i)
      counts[i] = !rest[i] && scan_shape_of(hframes[i].W) == sh ? (hframes[i].H + lines - 1) / lines : 0;
    BlockMap bm;
    const int nb = make_block_map(bm, counts, n);
    if (nb <= 0) continue;
    const bool m1 = mode == 1;
    if (sh == 0) m1 ? row_launch<0, 1>(dframes, bm, nb, wlut, cz, s) : row_launch<0, 0>(dframes, bm, nb, wlut, cz, s);
    if (sh == 1) m1 ? row_launch<1, 1>(dframes, bm, nb, wlut, cz, s) : row_launch<1, 0>(dframes, bm, nb, wlut, cz, s);
    if (sh == 2) m1 ? row_launch<2, 1>(dframes, bm, nb, wlut, cz, s) : row_launch<2, 0>(dframes, bm, nb, wlut, cz, s);
    if (sh == 3) m1 ? row_launch<3, 1>(dframes, bm, nb, wlut, cz, s) : row_launch<3, 0>(dframes, bm, nb, wlut, cz, s);
  }
  return left;
}

int launch_upsample_col_scan(const FrameDev* dframes, const FrameDev* hframes, int n, const float* wlut,
                             bool clear_zbuf, hipStream_t s, int* rest) {
  int left = 0;
  for (int i = 0; i < n; ++i) {
    rest[i] = scan_shape_of(hframes[i].H) < 0 || !(hframes[i].lam <= 0x1p100) ? 1 : 0;
    left += rest[i];
  }
  for (int sh = 0; sh < 4; ++sh) {
    const int lines = kColShapes[sh].waves * 64 / kColShapes[sh].P;
    int counts[kMaxFrames];
    for (int i = 0; i < n; ++i)
      counts[i] = !rest[i] && scan_shape_of(hframes[i].H) == sh ? (hframes[i].W + lines - 1) / lines : 0;
    BlockMap bm;
    const int nb = make_block_map(bm, counts, n);
    if (nb <= 0) continue;
    const int cz = clear_zbuf ? 1 : 0;
    if (sh == 0) col_launch<0>(dframes, bm, nb, wlut, cz, s);
    if (sh == 1) col_launch<1>(dframes, bm, nb, wlut, cz, s);
    if (sh == 2) col_launch<2>(dframes, bm, nb, wlut, cz, s);
    if (sh == 3) col_launch<3>(dframes, bm, nb, wlut, cz, s);
  }
  return left;
}

}  // namespace qm
