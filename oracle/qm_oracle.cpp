// qm_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Scalar C++ restatement of QuadtreeMapping's per-keyframe depth pipeline, used
// only as the correctness checker (tests/, __graft_entry__.smoke()) and as the
// CPU baseline leg of bench.py. Nothing in the shipped library links or calls it.
//
// Every stage is a template over Real:
//   Real = float  -> the FP32 oracle. It follows the reference's operation order
//                    with one pinned arithmetic (no FMA contraction, IEEE
//                    division/sqrt, double literals where the reference has them).
//                    OpenMP-parallel; it is also the CPU baseline.
//   Real = double -> the FP64 oracle, used for error bounds only.
//
// Reference files followed (paths relative to the reference repository):
//   quadtree ............ src/quadtree.cu:17-77
//   cost volume ......... src/depth_extract.cu:135-244, include/quadmap/pixel_cost.cuh,
//                         include/quadmap/camera_model/pinhole_camera.cuh:50-62,
//                         include/quadmap/se3.cuh:121-201
//   HBP + extraction .... src/bpextract.cu:37-395
//   upsample ............ src/depth_upsample.cu:20-302
//   fusion .............. src/depth_fusion.cu:13-213, src/seed_matrix.cu:553-597
//   keyframe flow ....... src/seed_matrix.cu:448-551
//
// Pinned semantics where the reference is racy or undefined (identical in the
// HIP kernels):
//   * quadtree: the recursive block definition (the reference breaks out of a
//     loop that contains __syncthreads; its intended result is this one).
//   * fusion z-buffer: nearest transformed depth wins, ties -> lower source
//     index; source index 0 is the "empty" sentinel and is never stored
//     (src/depth_fusion.cu:94-118 retries a CAS at most 5 times).
//   * hole filling: every source reads the table as it was after the z-buffer
//     (snapshot); an empty cell takes the first eligible source in raster order
//     (src/depth_fusion.cu:121-144 lets the first CAS win and can cascade).
//   * upsample: the row chain is split at 16*floor(W/16) and the column chain at
//     8*floor(H/8) (src/depth_upsample.cu:111,216 leave the remainder unwritten);
//     the remainder is solved as its own chain.
//   * float->int conversions truncate toward zero, NaN -> 0, saturating.
//   * exp(x) for float is (float)exp((double)x); 1/sqrt is 1.0f/sqrtf.

#include <cstdint>
#include <limits>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Cam {
  float fx, fy, cx, cy;
};

inline int trunc_i32(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (-2147483647 - 1);
  return (int)v;
}

template <typename Real> inline Real pinned_exp(Real x);
template <> inline float pinned_exp<float>(float x) { return (float)std::exp((double)x); }
template <> inline double pinned_exp<double>(double x) { return std::exp(x); }

template <typename Real> inline Real pinned_rsqrt(Real x) { return Real(1) / std::sqrt(x); }

template <typename Real> struct V3 {
  Real x, y, z;
};

// cam2world + normalize: include/quadmap/camera_model/pinhole_camera.cuh:50-55 and
// the normalize() of the reference's vector helpers (v * rsqrt(dot(v, v))).
template <typename Real> inline V3<Real> ray_dir(int x, int y, const Cam& c) {
  V3<Real> v;
  v.x = ((Real)x - (Real)c.cx) / (Real)c.fx;
  v.y = ((Real)y - (Real)c.cy) / (Real)c.fy;
  v.z = Real(1);
  Real d = (v.x * v.x + v.y * v.y) + v.z * v.z;
  Real inv = pinned_rsqrt<Real>(d);
  v.x = v.x * inv;
  v.y = v.y * inv;
  v.z = v.z * inv;
  return v;
}

// SE3<float> * float3: include/quadmap/se3.cuh:121-127,145-151,197-201 (rotate, then translate).
template <typename Real> inline V3<Real> se3_apply(const float* T, const V3<Real>& p) {
  V3<Real> q;
  q.x = (((Real)T[0] * p.x + (Real)T[1] * p.y) + (Real)T[2] * p.z) + (Real)T[3];
  q.y = (((Real)T[4] * p.x + (Real)T[5] * p.y) + (Real)T[6] * p.z) + (Real)T[7];
  q.z = (((Real)T[8] * p.x + (Real)T[9] * p.y) + (Real)T[10] * p.z) + (Real)T[11];
  return q;
}

// ---------------------------------------------------------------------------
// Quadtree level assignment (src/quadtree.cu:17-77).
// level(p) = max(2, max i <= 4 such that the aligned 2^i block holding p lies in
// the image and it and every aligned sub-block (down to 2x2) keeps all of its
// pixels within 10 grey levels of its own mean).
// ---------------------------------------------------------------------------
void quadtree_impl(const uint8_t* img, int W, int H, int pitch, uint8_t* levels) {
  std::vector<std::vector<uint8_t>> pass(5);
  for (int i = 1; i <= 4; ++i) {
    const int s = 1 << i;
    const int nbx = (W + s - 1) / s, nby = (H + s - 1) / s;
    pass[i].assign((size_t)nbx * nby, 0);
    const int pnbx = (W + s / 2 - 1) / (s / 2);
#pragma omp parallel for schedule(static)
    for (int by = 0; by < nby; ++by) {
      for (int bx = 0; bx < nbx; ++bx) {
        const int x0 = bx * s, y0 = by * s;
        if (x0 + s > W || y0 + s > H) continue;  // block not fully inside the image
        bool ok = true;
        if (i > 1) {
          for (int cy = 0; cy < 2 && ok; ++cy)
            for (int cx = 0; cx < 2 && ok; ++cx)
              ok = pass[i - 1][(size_t)(2 * by + cy) * pnbx + (2 * bx + cx)] != 0;
        }
        if (!ok) continue;
        long sum = 0;
        for (int y = y0; y < y0 + s; ++y)
          for (int x = x0; x < x0 + s; ++x) sum += img[(size_t)y * pitch + x];
        const long n = (long)s * s;
        // |I - sum/n| > 10  <=>  |n*I - sum| > 10*n  (exact; the reference's
        // float arithmetic is exact here because all values are integers < 2^24)
        for (int y = y0; y < y0 + s && ok; ++y)
          for (int x = x0; x < x0 + s && ok; ++x) {
            long dv = n * (long)img[(size_t)y * pitch + x] - sum;
            if (dv < 0) dv = -dv;
            if (dv > 10 * n) ok = false;
          }
        pass[i][(size_t)by * nbx + bx] = ok ? 1 : 0;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      int lv = 2;
      for (int i = 4; i >= 3; --i) {
        const int s = 1 << i;
        const int nbx = (W + s - 1) / s;
        if (pass[i][(size_t)(y / s) * nbx + (x / s)]) {
          lv = i;
          break;
        }
      }
      levels[(size_t)y * W + x] = (uint8_t)lv;
    }
  }
}

// ---------------------------------------------------------------------------
// Multi-view patch cost volume (src/depth_extract.cu:135-229). normalize_the_cost
// (:231-244) divides by add_num, which the 64 sample threads of a node's
// image_to_cost block (launched 64 x 5 at src/seed_matrix.cu:503-512) each
// incremented once inside "if (frame_id == 0)" (:225-228): count 64, an exact
// division by 64 of every node's costs (including the 100 of a sample no
// reference sees); cells that are not nodes keep 0.
// Slots f = 0..4 read reference (int)(0.9 * f) = {0,0,1,2,3}
// (src/depth_extract.cu:153-166 with the age table that is never written).
// ---------------------------------------------------------------------------
// Two projection arithmetics (the library's qm_params.cost_arith):
//   arith 0, reference order: P = dir * depth, q = R P + t with every product
//     rounded, px = (fx q.x) / q.z + cx (src/depth_extract.cu:178-191), texel
//     (int)((px + u2u i) + v2u j + 0.5) (:195-197);
//   arith 1, fused: per cell a = (fx (R0.dir), fy (R1.dir), R2.dir), each dot
//     product fma(R2, d.z, fma(R1, d.y, R0 d.x)); per sample U = fma(depth, a,
//     (fx t0, fy t1, t2)), r = 1 / U.z, p5 = fma(U.xy, r, c + 0.5) (the pixel
//     + 0.5), texel floor((p5 + u2u i) + v2u j), in view when 2 <= p5 < W - 2.
//     The same algebra with fewer roundings; the HIP kernel computes the same
//     bits (its reciprocal is exact where it replaces the division).
template <typename Real> inline Real fma_r(Real a, Real b, Real c) { return std::fma(a, b, c); }
template <typename Real> inline int floor_sat(Real v) {
  v = std::fmin(std::fmax(v, Real(-1.0e9)), Real(1.0e9));
  return (int)std::floor(v);
}

template <typename Real>
void cost_impl(const uint8_t* key, const uint8_t* const* refs, int W, int H, int pitch,
               const float* T /*4x12*/, const Cam& cam, int N, const Real* depth_table,
               const uint8_t* levels, Real* cost, int32_t* count, int arith) {
  const int Wc = W / 4, Hc = H / 4;
#pragma omp parallel for schedule(dynamic, 4)
  for (int cy = 0; cy < Hc; ++cy) {
    for (int cxi = 0; cxi < Wc; ++cxi) {
      Real* out = cost + ((size_t)cy * Wc + cxi) * N;
      for (int d = 0; d < N; ++d) out[d] = Real(0);
      if (count) count[(size_t)cy * Wc + cxi] = 0;
      const int x = 4 * cxi, y = 4 * cy;
      const int qs = 1 << levels[(size_t)y * W + x];
      if (x >= W - 1 || y >= H - 1 || x <= 0 || y <= 0 || x % qs != 0 || y % qs != 0) continue;
      Real patch[3][3];
      for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) patch[i + 1][j + 1] = (Real)key[(size_t)(y + j) * pitch + (x + i)];
      const V3<Real> dir = ray_dir<Real>(x, y, cam);
      Real ax[4], ay[4], az[4];
      for (int r = 0; r < 4; ++r) {
        const float* Tr = T + 12 * r;
        ax[r] = (Real)cam.fx * fma_r<Real>(Tr[2], dir.z, fma_r<Real>(Tr[1], dir.y, (Real)Tr[0] * dir.x));
        ay[r] = (Real)cam.fy * fma_r<Real>(Tr[6], dir.z, fma_r<Real>(Tr[5], dir.y, (Real)Tr[4] * dir.x));
        az[r] = fma_r<Real>(Tr[10], dir.z, fma_r<Real>(Tr[9], dir.y, (Real)Tr[8] * dir.x));
      }
      for (int d = 0; d < N; ++d) {
        const Real depth = depth_table[d];
        V3<Real> P;
        P.x = dir.x * depth;
        P.y = dir.y * depth;
        P.z = dir.z * depth;
        Real cr[4], nr[4];
        for (int r = 0; r < 4; ++r) {
          const float* Tr = T + 12 * r;
          if (arith == 1) {
            const Real U = fma_r<Real>(depth, ax[r], (Real)cam.fx * (Real)Tr[3]);
            const Real V = fma_r<Real>(depth, ay[r], (Real)cam.fy * (Real)Tr[7]);
            const Real Z = fma_r<Real>(depth, az[r], (Real)Tr[11]);
            const Real rz = Real(1) / Z;
            const Real px5 = fma_r<Real>(U, rz, (Real)cam.cx + Real(0.5));
            const Real py5 = fma_r<Real>(V, rz, (Real)cam.cy + Real(0.5));
            if (px5 >= Real(2) && px5 < (Real)(W - 2) && py5 >= Real(2) && py5 < (Real)(H - 2)) {
              const Real u2u = (Real)Tr[0], u2v = (Real)Tr[4], v2u = (Real)Tr[1], v2v = (Real)Tr[5];
              const uint8_t* ref = refs[r];
              Real sad = Real(0);
              for (int j = -1; j <= 1; ++j) {
                for (int i = -1; i <= 1; ++i) {
                  int chx = floor_sat<Real>((px5 + u2u * (Real)i) + v2u * (Real)j);
                  int chy = floor_sat<Real>((py5 + u2v * (Real)i) + v2v * (Real)j);
                  chx = chx < 0 ? 0 : (chx > W - 1 ? W - 1 : chx);
                  chy = chy < 0 ? 0 : (chy > H - 1 ? H - 1 : chy);
                  sad += std::fabs(patch[i + 1][j + 1] - (Real)ref[(size_t)chy * pitch + chx]);
                }
              }
              cr[r] = sad;
              nr[r] = Real(1);
            } else {
              cr[r] = Real(0);
              nr[r] = Real(0);
            }
            continue;
          }
          const V3<Real> q = se3_apply<Real>(Tr, P);
          const Real px = ((Real)cam.fx * q.x) / q.z + (Real)cam.cx;
          const Real py = ((Real)cam.fy * q.y) / q.z + (Real)cam.cy;
          const int pxi = trunc_i32((double)px + 0.5);
          const int pyi = trunc_i32((double)py + 0.5);
          if (pxi >= 2 && pxi < W - 2 && pyi >= 2 && pyi < H - 2) {
            const Real u2u = (Real)Tr[0], u2v = (Real)Tr[4], v2u = (Real)Tr[1], v2v = (Real)Tr[5];
            const uint8_t* ref = refs[r];
            Real sad = Real(0);
            for (int j = -1; j <= 1; ++j) {
              for (int i = -1; i <= 1; ++i) {
                int chx = trunc_i32((double)((px + u2u * (Real)i) + v2u * (Real)j) + 0.5);
                int chy = trunc_i32((double)((py + u2v * (Real)i) + v2v * (Real)j) + 0.5);
                chx = chx < 0 ? 0 : (chx > W - 1 ? W - 1 : chx);
                chy = chy < 0 ? 0 : (chy > H - 1 ? H - 1 : chy);
                sad += std::fabs(patch[i + 1][j + 1] - (Real)ref[(size_t)chy * pitch + chx]);
              }
            }
            cr[r] = sad;
            nr[r] = Real(1);
          } else {
            cr[r] = Real(0);
            nr[r] = Real(0);
          }
        }
        // shared-memory tree over the 5 slots: ((c0 + c4) + c2) + (c1 + c3)
        const Real tot = ((cr[0] + cr[3]) + cr[1]) + (cr[0] + cr[2]);
        const Real n = ((nr[0] + nr[3]) + nr[1]) + (nr[0] + nr[2]);
        out[d] = (n > Real(0) ? (tot / n) / Real(255) : Real(100)) / Real(64);
      }
      if (count) count[(size_t)cy * Wc + cxi] = 64;
    }
  }
}

// ---------------------------------------------------------------------------
// Hierarchical BP and winner extraction (src/bpextract.cu:37-395).
// ---------------------------------------------------------------------------
struct BpParams {
  int N;
  float P1, P2;
  int iters[4];  // fine -> coarse
  double idepth_min, idepth_step;
};

template <typename Real> struct Level {
  int w, h;
  std::vector<Real> cost;    // h*w*N
  std::vector<Real> msg[4];  // up, dm, lm, rm : h*w*N each
};

template <typename Real>
void bp_sweep(Level<Real>& L, int l, bool A, const uint8_t* levels, int W, int H, const BpParams& p) {
  const int N = p.N, w = L.w, h = L.h;
  const int hw = (w + 1) / 2;
  Real* up = L.msg[0].data();
  Real* dm = L.msg[1].data();
  Real* lm = L.msg[2].data();
  Real* rm = L.msg[3].data();
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y) {
    std::vector<Real> hb(4 * N), m(N), s(N);
    for (int bx = 0; bx < hw; ++bx) {
      const int lev = l + 2;
      if (lev <= 4) {
        // quirk kept on purpose: the lookup uses the block index bx, not the
        // checkerboard x (src/bpextract.cu:209-228)
        int X = bx << lev, Y = y << lev;
        X = X > W - 1 ? W - 1 : X;
        Y = Y > H - 1 ? H - 1 : Y;
        if ((int)levels[(size_t)Y * W + X] > lev) continue;
      }
      const int x = A ? 2 * bx + (y % 2) : 2 * bx + ((y + 1) % 2);
      if (x >= w || y >= h) continue;
      const Real* data = L.cost.data() + ((size_t)y * w + x) * N;
      const Real* from_up = y > 0 ? dm + ((size_t)(y - 1) * w + x) * N : nullptr;
      const Real* from_down = y < h - 1 ? up + ((size_t)(y + 1) * w + x) * N : nullptr;
      const Real* from_left = x > 0 ? rm + ((size_t)y * w + x - 1) * N : nullptr;
      const Real* from_right = x < w - 1 ? lm + ((size_t)y * w + x + 1) * N : nullptr;
      Real outm[4][256];
      for (int dir = 0; dir < 4; ++dir) {
        Real* hd = hb.data() + dir * N;
        for (int d = 0; d < N; ++d) {
          Real v = data[d];
          if (dir != 0 && from_up) v += from_up[d];
          if (dir != 1 && from_down) v += from_down[d];
          if (dir != 2 && from_left) v += from_left[d];
          if (dir != 3 && from_right) v += from_right[d];
          hd[d] = v;
        }
        Real hmin = hd[0];
        for (int d = 1; d < N; ++d) hmin = hd[d] < hmin ? hd[d] : hmin;
        for (int d = 0; d < N; ++d) {
          Real mc = hd[d];
          if (d > 0) mc = std::fmin(mc, hd[d - 1] + (Real)p.P1);
          if (d < N - 1) mc = std::fmin(mc, hd[d + 1] + (Real)p.P1);
          mc = std::fmin(mc, hmin + (Real)p.P2);
          m[d] = mc;
          s[d] = mc;
        }
        // pairwise tree sum in the order of src/bpextract.cu:293-300
        for (int off = N / 2; off > 0; off /= 2)
          for (int d = 0; d < off; ++d) s[d] += s[d + off];
        const Real mean = s[0] / (Real)N;
        for (int d = 0; d < N; ++d) outm[dir][d] = m[d] - mean;
      }
      for (int dir = 0; dir < 4; ++dir) {
        Real* dst = L.msg[dir].data() + ((size_t)y * w + x) * N;
        for (int d = 0; d < N; ++d) dst[d] = outm[dir][d];
      }
    }
  }
}

template <typename Real>
void bp_extract_impl(const Real* cost0, int W, int H, const uint8_t* levels, const BpParams& p,
                     Real* depth_coarse, int32_t* label, Real* msg_out /*4*Hc*Wc*N or null*/) {
  const int N = p.N;
  const int Wc = W / 4, Hc = H / 4;
  Level<Real> L[4];
  L[0].w = Wc;
  L[0].h = Hc;
  for (int i = 1; i < 4; ++i) {
    L[i].w = (L[i - 1].w + 1) / 2;
    L[i].h = (L[i - 1].h + 1) / 2;
  }
  L[0].cost.assign(cost0, cost0 + (size_t)Wc * Hc * N);
  // cost pyramid: src/bpextract.cu:169-198, children summed in the order
  // (2x,2y), (2x,2y+1), (2x+1,2y), (2x+1,2y+1), starting from 0.0f
  for (int i = 1; i < 4; ++i) {
    const int w = L[i].w, h = L[i].h, w0 = L[i - 1].w, h0 = L[i - 1].h;
    L[i].cost.assign((size_t)w * h * N, Real(0));
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x)
        for (int d = 0; d < N; ++d) {
          Real sum = Real(0);
          for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj)
              if (2 * x + ii < w0 && 2 * y + jj < h0)
                sum += L[i - 1].cost[((size_t)(2 * y + jj) * w0 + (2 * x + ii)) * N + d];
          L[i].cost[((size_t)y * w + x) * N + d] = sum;
        }
  }
  for (int dir = 0; dir < 4; ++dir) L[3].msg[dir].assign((size_t)L[3].w * L[3].h * N, Real(0));
  for (int l = 3; l >= 0; --l) {
    if (l < 3) {
      // message upsample: src/bpextract.cu:374-395
      const int w = L[l].w, h = L[l].h, w1 = L[l + 1].w;
      for (int dir = 0; dir < 4; ++dir) {
        L[l].msg[dir].assign((size_t)w * h * N, Real(0));
        for (int y = 0; y < h; ++y)
          for (int x = 0; x < w; ++x)
            std::memcpy(&L[l].msg[dir][((size_t)y * w + x) * N],
                        &L[l + 1].msg[dir][((size_t)(y / 2) * w1 + (x / 2)) * N], sizeof(Real) * N);
        L[l + 1].msg[dir].clear();
        L[l + 1].msg[dir].shrink_to_fit();
      }
    }
    bool A = true;
    for (int it = 0; it < p.iters[l]; ++it) {
      bp_sweep<Real>(L[l], l, A, levels, W, H, p);
      A = !A;
    }
  }
  // winner extraction: src/bpextract.cu:314-372
  const Real* up = L[0].msg[0].data();
  const Real* dm = L[0].msg[1].data();
  const Real* lm = L[0].msg[2].data();
  const Real* rm = L[0].msg[3].data();
#pragma omp parallel for schedule(static)
  for (int y = 0; y < Hc; ++y) {
    std::vector<Real> b(N), mc(N);
    std::vector<int> mi(N);
    for (int x = 0; x < Wc; ++x) {
      const size_t ci = (size_t)y * Wc + x;
      const int qs = 1 << levels[(size_t)(4 * y) * W + 4 * x];
      if ((4 * x) % qs != 0 || (4 * y) % qs != 0) {
        depth_coarse[ci] = Real(0);
        label[ci] = -1;
        continue;
      }
      const Real* data = L[0].cost.data() + ci * N;
      for (int d = 0; d < N; ++d) {
        Real v = data[d];
        if (x != 0) v += rm[(ci - 1) * N + d];
        if (x != Wc - 1) v += lm[(ci + 1) * N + d];
        if (y != 0) v += dm[(ci - Wc) * N + d];
        if (y != Hc - 1) v += up[(ci + Wc) * N + d];
        b[d] = v;
        mc[d] = v;
        mi[d] = d;
      }
      for (int off = N / 2; off > 0; off /= 2)
        for (int d = 0; d < off; ++d)
          if (mc[d + off] < mc[d]) {
            mc[d] = mc[d + off];
            mi[d] = mi[d + off];
          }
      const int id = mi[0];
      Real disp = (Real)id;
      if (id > 0 && id < N - 1) {
        const Real pre = b[id - 1], post = b[id + 1];
        const Real a = (pre - Real(2) * mc[0]) + post;
        const Real bb = (-pre) + post;
        const Real ba = bb / a;
        if (std::isfinite(ba)) disp = (Real)id - ba / Real(2);
      }
      depth_coarse[ci] = (Real)(1.0 / (p.idepth_step * (double)disp + p.idepth_min));
      label[ci] = id;
    }
  }
  if (msg_out) {
    const size_t n = (size_t)Wc * Hc * N;
    for (int dir = 0; dir < 4; ++dir) std::memcpy(msg_out + dir * n, L[0].msg[dir].data(), sizeof(Real) * n);
  }
}

// ---------------------------------------------------------------------------
// Global tridiagonal upsampling (src/depth_upsample.cu:20-302), generalised:
//   rhs_x = conf ? conf*d : d,  rhs_n = conf ? conf : 1{d > 0}
//   row pass with w_row (H x (W+1)), then column pass with w_col ((H+1) x W),
//   out = x / n.
// A = I + lambda*L(w) per line; the chain is split at row_seg / col_seg.
// ---------------------------------------------------------------------------
template <typename Real>
inline void thomas_coeffs(Real wl, Real wr, double lam, Real& a, Real& b, Real& c) {
  a = (Real)(-lam * (double)wl);
  b = (Real)(1.0 + lam * (double)(wl + wr));
  c = (Real)(-lam * (double)wr);
}

template <typename Real>
void solve_line(int n, const Real* w /*n+1 couplings, stride ws*/, int ws, const Real* rx, const Real* rn,
                int rstride, Real* ox, Real* on, int ostride, int seg, double lam, Real* tmp /*3n*/) {
  int segs[3] = {0, seg < n ? seg : n, n};
  for (int sgi = 0; sgi < 2; ++sgi) {
    const int s0 = segs[sgi], s1 = segs[sgi + 1];
    if (s1 <= s0) continue;
    Real fc = 0, fd = 0, fn = 0;
    for (int i = s0; i < s1; ++i) {
      Real a, b, c;
      thomas_coeffs<Real>(w[(size_t)i * ws], w[(size_t)(i + 1) * ws], lam, a, b, c);
      const Real den = b - fc * a;
      const Real ch = c / den;
      const Real dh = (rx[(size_t)i * rstride] - fd * a) / den;
      const Real nh = (rn[(size_t)i * rstride] - fn * a) / den;
      tmp[3 * i] = ch;
      tmp[3 * i + 1] = dh;
      tmp[3 * i + 2] = nh;
      fc = ch;
      fd = dh;
      fn = nh;
    }
    fd = 0;
    fn = 0;
    for (int i = s1 - 1; i >= s0; --i) {
      const Real xv = tmp[3 * i + 1] - tmp[3 * i] * fd;
      const Real nv = tmp[3 * i + 2] - tmp[3 * i] * fn;
      ox[(size_t)i * ostride] = xv;
      on[(size_t)i * ostride] = nv;
      fd = xv;
      fn = nv;
    }
  }
}

// ---------------------------------------------------------------------------
// The same line systems by segmented scans (qm_params.upsample_solver =
// QM_UPSAMPLE_SCAN): P lanes (32 for lines of up to 512 steps, else 64), lane i
// running steps [i m, (i + 1) m), m = ceil(n / P); steps past the
// line's end pad it (a = 0, b = 1, c = 0, zero
// right-hand sides: their states stay finite and never reach a real step).
// The three recurrences of the Thomas solve are maps of their previous state,
// so each lane composes its steps' map, an exclusive scan over the lanes gives
// every lane its incoming state, and the lane then runs its steps from it:
//   ch   ch_k = c_k / (b_k - ch_{k-1} a_k): a Moebius map, the 2 x 2 matrix
//        [[0, c_k], [-a_k, b_k]] on (p, q) with ch = p / q;
//   x, n dh_k = (r_k - dh_{k-1} a_k) y_k, y_k = 1 / den_k: affine; a lane
//        keeps each step's value from a zero start u_k and slope V_k, and
//        dh_k = fma(V_k, incoming, u_k);
//   X, N X_k = dh_k - ch_k X_{k+1} backward: affine, the same with a suffix
//        scan.
// A chain start k (0 and seg) uses a_k = 0 in the recurrences (its state is 0
// before it), a chain end (seg - 1 and n - 1) uses ch_k = 0 in the backward
// one. The Moebius lane maps and their scan are in double (their products
// cancel badly in float when lambda is large), rescaled by a power of two
// (exact) after every 8th step of a lane and at its end; the affine maps are in
// float. The scans follow one fixed tree
// (scan_tree). The arithmetic is fixed here and in the HIP kernels
// (k_scan.hip), which match it bit for bit; it is a different rounding of the
// same solve as solve_line. Lines with more than 64 kScanMaxSeg steps use
// solve_line.
// ---------------------------------------------------------------------------
constexpr int kScanLanes = 64;
constexpr int kScanMaxSeg = 24;

template <typename Real>
struct Mob {
  Real g11, g12, g21, g22;
};

// G * 2^-e with e the binary exponent of max |entry| (frexp): exact
template <typename Real>
inline void mob_norm(Mob<Real>& G) {
  Real mx = std::fabs(G.g11), t = std::fabs(G.g12);
  mx = t > mx ? t : mx;
  t = std::fabs(G.g21);
  mx = t > mx ? t : mx;
  t = std::fabs(G.g22);
  mx = t > mx ? t : mx;
  if (!(mx > Real(0)) || !(mx <= std::numeric_limits<Real>::max())) return;
  int e;
  (void)std::frexp(mx, &e);
  G.g11 = std::ldexp(G.g11, -e);
  G.g12 = std::ldexp(G.g12, -e);
  G.g21 = std::ldexp(G.g21, -e);
  G.g22 = std::ldexp(G.g22, -e);
}

// A B (B applied first); the scan's products are not rescaled (six levels of
// products of rescaled lane maps stay far inside the double range)
template <typename Real>
inline Mob<Real> mob_mul(const Mob<Real>& A, const Mob<Real>& B) {
  Mob<Real> R;
  R.g11 = A.g11 * B.g11 + A.g12 * B.g21;
  R.g12 = A.g11 * B.g12 + A.g12 * B.g22;
  R.g21 = A.g21 * B.g11 + A.g22 * B.g21;
  R.g22 = A.g21 * B.g12 + A.g22 * B.g22;
  return R;
}

// The scans over the 64 lanes, one fixed tree (the HIP kernels' scan_tree):
// Hillis-Steele inside each row of 16 lanes (offsets 1, 2, 4, 8), then rows 1
// and 3 combine with the last lane of rows 0 and 2, then rows 2 and 3 with
// lane 31; the suffix scan mirrors it (rows 0 and 2 with the first lane of
// rows 1 and 3, then rows 0 and 1 with lane 32). comb(a, b) combines lane
// value a with b, the value of lanes before it (prefix) or after it (suffix).
template <typename T, typename Comb>
void scan_tree(std::vector<T>& v, bool suffix, Comb comb) {
  const int P = (int)v.size();  // 32 (rows 0 and 1 of the tree) or 64
  for (int o = 1; o < 16; o <<= 1) {
    const std::vector<T> t = v;
    for (int i = 0; i < P; ++i) {
      const int j = i & 15;
      if (suffix ? j + o < 16 : j >= o) v[i] = comb(t[i], t[suffix ? i + o : i - o]);
    }
  }
  {
    const std::vector<T> t = v;
    for (int i = 0; i < P; ++i) {
      const int row = i >> 4;
      if (suffix ? (row & 1) == 0 : (row & 1) == 1)
        v[i] = comb(t[i], t[suffix ? (row < 2 ? 16 : 48) : (row < 2 ? 15 : 47)]);
    }
  }
  if (P == 64) {
    const std::vector<T> t = v;
    for (int i = 0; i < P; ++i) {
      const int row = i >> 4;
      if (suffix ? row < 2 : row >= 2) v[i] = comb(t[i], t[suffix ? 32 : 31]);
    }
  }
}

template <typename Real>
void solve_line_scan(int P, int n, const Real* w, int ws, const Real* rx, const Real* rn, int rstride, Real* ox,
                     Real* on, int ostride, int seg, double lam) {
  using D = double;  // the Moebius lane maps and their scan
  const int m = (n + P - 1) / P, L = P * m;
  const int sg = seg < n ? seg : n;
  // every lane runs m steps; steps k >= n pad the line (a = 0, b = 1, c = 0,
  // zero right-hand sides) and their results are dropped
  std::vector<Real> A(L, Real(0)), B(L, Real(1)), Cc(L, Real(0)), R(L, Real(0)), RN(L, Real(0));
  std::vector<Real> Y(L), CH(L), DH(L), NH(L);
  for (int k = 0; k < n; ++k) {
    thomas_coeffs<Real>(w[(size_t)k * ws], w[(size_t)(k + 1) * ws], lam, A[k], B[k], Cc[k]);
    if (k == 0 || k == sg) A[k] = Real(0);
    R[k] = rx[(size_t)k * rstride];
    RN[k] = rn[(size_t)k * rstride];
  }
  // ---- ch: lane maps (rescaled after every 8th step and at the end), scan, steps ----
  std::vector<Mob<D>> G(P);
  for (int i = 0; i < P; ++i) {
    Mob<D> g{D(1), D(0), D(0), D(1)};
    for (int t = 0; t < m; ++t) {
      const int k = i * m + t;
      Mob<D> q;
      q.g11 = (D)Cc[k] * g.g21;
      q.g12 = (D)Cc[k] * g.g22;
      q.g21 = (D)B[k] * g.g21 - (D)A[k] * g.g11;
      q.g22 = (D)B[k] * g.g22 - (D)A[k] * g.g12;
      if ((t & 7) == 7) mob_norm(q);
      g = q;
    }
    mob_norm(g);
    G[i] = g;
  }
  scan_tree(G, false, [](const Mob<D>& a, const Mob<D>& b) { return mob_mul(a, b); });
  for (int i = 0; i < P; ++i) {
    Real st = i == 0 ? Real(0) : (Real)(G[i - 1].g12 / G[i - 1].g22);
    for (int k = i * m; k < (i + 1) * m; ++k) {
      const Real den = B[k] - st * A[k];
      Y[k] = Real(1) / den;
      st = Cc[k] * Y[k];
      CH[k] = st;
    }
  }
  // ---- x, n: per step the from-zero values and slopes of the lane's affine
  // map (u_k, V_k), scan of the lanes' maps, dh_k = fma(V_k, incoming, u_k) ----
  struct Aff {
    Real u, w, v;  // two intercepts sharing one slope
  };
  auto aff = [](const Aff& a, const Aff& b) { return Aff{a.u + a.v * b.u, a.w + a.v * b.w, a.v * b.v}; };
  std::vector<Real> Uu(L), Uw(L), Vs(L);
  std::vector<Aff> U(P);
  for (int i = 0; i < P; ++i) {
    Real ux = 0, un = 0, v = 1;
    for (int k = i * m; k < (i + 1) * m; ++k) {
      ux = (R[k] - ux * A[k]) * Y[k];
      un = (RN[k] - un * A[k]) * Y[k];
      v = -(A[k] * Y[k]) * v;
      Uu[k] = ux;
      Uw[k] = un;
      Vs[k] = v;
    }
    U[i] = Aff{ux, un, v};
  }
  scan_tree(U, false, aff);
  for (int i = 0; i < P; ++i) {
    const Real dx = i == 0 ? Real(0) : U[i - 1].u, dn = i == 0 ? Real(0) : U[i - 1].w;
    for (int k = i * m; k < (i + 1) * m; ++k) {
      DH[k] = std::fma(Vs[k], dx, Uu[k]);
      NH[k] = std::fma(Vs[k], dn, Uw[k]);
    }
  }
  // ---- X, N: the same backward: from-zero values and slopes (p_k, Q_k) from
  // the lane's end, suffix scan, X_k = fma(Q_k, incoming, p_k) ----
  auto cut = [&](int k) { return (k == sg - 1 || k == n - 1) ? Real(0) : CH[k]; };
  std::vector<Aff> Q(P);
  for (int i = 0; i < P; ++i) {
    Real px = 0, pn = 0, q = 1;
    for (int k = (i + 1) * m - 1; k >= i * m; --k) {
      const Real c = cut(k);
      px = DH[k] - c * px;
      pn = NH[k] - c * pn;
      q = -c * q;
      Uu[k] = px;
      Uw[k] = pn;
      Vs[k] = q;
    }
    Q[i] = Aff{px, pn, q};
  }
  scan_tree(Q, true, aff);
  for (int i = 0; i < P; ++i) {
    const Real xe = i == P - 1 ? Real(0) : Q[i + 1].u, ne = i == P - 1 ? Real(0) : Q[i + 1].w;
    for (int k = i * m; k < (i + 1) * m && k < n; ++k) {
      ox[(size_t)k * ostride] = std::fma(Vs[k], xe, Uu[k]);
      on[(size_t)k * ostride] = std::fma(Vs[k], ne, Uw[k]);
    }
  }
}

// one line by the chosen solver: the scan over 32 lanes for lines of up to 512
// steps, over 64 lanes up to 64 kScanMaxSeg, Thomas beyond
template <typename Real>
void solve_line_by(int solver, int n, const Real* w, int ws, const Real* rx, const Real* rn, int rstride, Real* ox,
                   Real* on, int ostride, int seg, double lam, Real* tmp) {
  // (lambda beyond 2^100 keeps Thomas: the scan's pivots stay below 2^126)
  if (solver == 1 && n <= 512 && lam <= 0x1p100)
    solve_line_scan<Real>(32, n, w, ws, rx, rn, rstride, ox, on, ostride, seg, lam);
  else if (solver == 1 && n <= kScanLanes * kScanMaxSeg && lam <= 0x1p100)
    solve_line_scan<Real>(64, n, w, ws, rx, rn, rstride, ox, on, ostride, seg, lam);
  else
    solve_line<Real>(n, w, ws, rx, rn, rstride, ox, on, ostride, seg, lam, tmp);
}

template <typename Real>
void upsample_impl(int W, int H, const Real* d, const Real* conf, const Real* w_row, const Real* w_col,
                   double lam, int row_seg, int col_seg, Real* out, Real* row_x, Real* row_n, int solver = 0) {
  std::vector<Real> rxv, rnv, rhx, rhn;
  Real* RX = row_x;
  Real* RN = row_n;
  if (!RX) {
    rxv.resize((size_t)W * H);
    RX = rxv.data();
  }
  if (!RN) {
    rnv.resize((size_t)W * H);
    RN = rnv.data();
  }
  rhx.resize((size_t)W * H);
  rhn.resize((size_t)W * H);
  for (size_t i = 0; i < (size_t)W * H; ++i) {
    if (conf) {
      rhx[i] = conf[i] * d[i];
      rhn[i] = conf[i];
    } else {
      rhx[i] = d[i];
      rhn[i] = d[i] > Real(0) ? Real(1) : Real(0);
    }
  }
#pragma omp parallel
  {
    std::vector<Real> tmp(3 * (size_t)(W > H ? W : H));
#pragma omp for schedule(static)
    for (int y = 0; y < H; ++y)
      solve_line_by<Real>(solver, W, w_row + (size_t)y * (W + 1), 1, rhx.data() + (size_t)y * W,
                          rhn.data() + (size_t)y * W, 1, RX + (size_t)y * W, RN + (size_t)y * W, 1, row_seg, lam,
                          tmp.data());
  }
  std::vector<Real> cx((size_t)W * H), cn((size_t)W * H);
#pragma omp parallel
  {
    std::vector<Real> tmp(3 * (size_t)(W > H ? W : H));
#pragma omp for schedule(static)
    for (int x = 0; x < W; ++x)
      solve_line_by<Real>(solver, H, w_col + x, W, RX + x, RN + x, W, cx.data() + x, cn.data() + x, W, col_seg,
                          lam, tmp.data());
  }
  for (size_t i = 0; i < (size_t)W * H; ++i) out[i] = cx[i] / cn[i];
}

// Smoothness weights from the keyframe (src/depth_upsample.cu:58-93):
// w = expf(-(dI)^2/400) on interior edges; the x=0 / y=0 edges are 0; the x=W
// (y=H) edge is 1 when W (H) is not a multiple of 16, because the 16x16 launch
// grid then covers it and the clamped texture gives dI = 0; otherwise 0.
template <typename Real> Real weight_of_delta(int delta);
template <> float weight_of_delta<float>(int delta) {
  const float arg = (float)(-(double)(delta * delta) / 400.0);
  return pinned_exp<float>(arg);
}
template <> double weight_of_delta<double>(int delta) { return std::exp(-(double)(delta * delta) / 400.0); }

template <typename Real>
void weights_impl(const uint8_t* img, int W, int H, int pitch, Real* w_row, Real* w_col) {
  Real lut[256];
  for (int i = 0; i < 256; ++i) lut[i] = weight_of_delta<Real>(i);
  const Real edge_r = (W % 16) ? Real(1) : Real(0);
  const Real edge_c = (H % 16) ? Real(1) : Real(0);
  for (int y = 0; y < H; ++y) {
    w_row[(size_t)y * (W + 1)] = Real(0);
    for (int x = 1; x < W; ++x)
      w_row[(size_t)y * (W + 1) + x] =
          lut[std::abs((int)img[(size_t)y * pitch + x] - (int)img[(size_t)y * pitch + x - 1])];
    w_row[(size_t)y * (W + 1) + W] = edge_r;
  }
  for (int x = 0; x < W; ++x) w_col[x] = Real(0);
  for (int y = 1; y < H; ++y)
    for (int x = 0; x < W; ++x)
      w_col[(size_t)y * W + x] = lut[std::abs((int)img[(size_t)y * pitch + x] - (int)img[(size_t)(y - 1) * pitch + x])];
  for (int x = 0; x < W; ++x) w_col[(size_t)H * W + x] = edge_c;
}

// ---------------------------------------------------------------------------
// Gaussian x Beta fusion (src/depth_fusion.cu:13-28,48-213).
// seeds: H x W x 4 (a, b, mu, sigma^2), updated in place by the warp step.
// ---------------------------------------------------------------------------
template <typename Real> inline bool is_badpoint(Real a, Real b) {
  return ((double)a < 0.001) || ((double)(a / (a + b)) < 0.45);
}
template <typename Real> inline bool is_goodpoint(Real a, Real b) { return (double)(a / (a + b)) > 0.60; }

template <typename Real> inline Real normpdf(Real x, Real mu, Real s2) {
  const Real e = pinned_exp<Real>((-(x - mu)) * (x - mu) / (Real(2) * s2));
  const Real r = pinned_rsqrt<Real>((Real)(2.0 * M_PI * (double)s2));
  return e * r;
}

template <typename Real>
void fuse_update_impl(int W, int H, const int32_t* filled, const Real* seeds, const Real* depth_in, Real* depth_out,
                      Real* seeds_out);

template <typename Real>
void fuse_impl(int W, int H, const Cam& cam, const float* T /*12*/, Real* seeds, const Real* depth_in, Real* depth_out,
               Real* seeds_out, int32_t* table_out) {
  const size_t n = (size_t)W * H;
  // 1) warp the previous seeds into the current frame. The z-buffer keeps, per
  //    target pixel, the smallest key (float bits of the warped depth, source
  //    index); NaN depths use the canonical quiet-NaN bits, so they lose to any
  //    number. Source index 0 is the empty sentinel and is never stored.
  std::vector<uint64_t> key(n, ~0ull);
  std::vector<uint64_t> cand_k(n, ~0ull);
  std::vector<int32_t> cand_t(n, -1);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const size_t idx = (size_t)y * W + x;
      Real* s = seeds + 4 * idx;
      if (is_badpoint<Real>(s[0], s[1])) continue;
      const V3<Real> dir = ray_dir<Real>(x, y, cam);
      V3<Real> P;
      P.x = dir.x * s[2];
      P.y = dir.y * s[2];
      P.z = dir.z * s[2];
      const V3<Real> q = se3_apply<Real>(T, P);
      const Real nd = std::sqrt((q.x * q.x + q.y * q.y) + q.z * q.z);
      if ((double)nd <= 0.5) continue;  // false for NaN: the reference carries on
      const Real px = ((Real)cam.fx * q.x) / q.z + (Real)cam.cx;
      const Real py = ((Real)cam.fy * q.y) / q.z + (Real)cam.cy;
      const int pxi = trunc_i32((double)px + 0.5);
      const int pyi = trunc_i32((double)py + 0.5);
      if (pxi >= W || pxi < 0 || pyi >= H || pyi < 0) continue;  // no write-back either
      s[2] = nd;
      s[3] = (Real)((double)s[3] + (double)nd * 0.01);
      float ndf = (float)nd;
      uint32_t bits;
      std::memcpy(&bits, &ndf, 4);
      if (ndf != ndf) bits = 0x7FC00000u;
      cand_k[idx] = ((uint64_t)bits << 32) | (uint64_t)(uint32_t)idx;
      cand_t[idx] = pyi * W + pxi;
    }
  }
  for (size_t idx = 1; idx < n; ++idx) {
    const int32_t t = cand_t[idx];
    if (t < 0) continue;
    if (cand_k[idx] < key[t]) key[t] = cand_k[idx];
  }
  std::vector<int32_t> table(n, 0);
  for (size_t t = 0; t < n; ++t) table[t] = key[t] == ~0ull ? 0 : (int32_t)(uint32_t)(key[t] & 0xFFFFFFFFu);
  // 2) hole filling from the snapshot
  std::vector<int32_t> filled(table);
  for (int ty = 0; ty < H; ++ty)
    for (int tx = 0; tx < W; ++tx) {
      const size_t t = (size_t)ty * W + tx;
      if (table[t] != 0) continue;
      for (int sy = ty - 1; sy <= ty + 1 && filled[t] == 0; ++sy)
        for (int sx = tx - 1; sx <= tx + 1; ++sx) {
          if (sx == tx && sy == ty) continue;
          if (sx < 2 || sx > W - 2 || sy < 2 || sy > H - 2) continue;
          const int32_t v = table[(size_t)sy * W + sx];
          if (v != 0) {
            filled[t] = v;
            break;
          }
        }
    }
  // 3) per-pixel Gaussian x Beta update
  fuse_update_impl<Real>(W, H, filled.data(), seeds, depth_in, depth_out, seeds_out);
  if (table_out) std::memcpy(table_out, filled.data(), sizeof(int32_t) * n);
}

// The per-pixel update of src/depth_fusion.cu:146-213 given the hole-filled
// table and the warped seeds (the half the differentiable path varies).
template <typename Real>
void fuse_update_impl(int W, int H, const int32_t* filled, const Real* seeds, const Real* depth_in, Real* depth_out,
                      Real* seeds_out) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const size_t idx = (size_t)y * W + x;
      const Real de = depth_in[idx];
      Real unc = (Real)std::fmax((Real)0.5, (Real)((double)de * 0.2));
      unc = unc * unc;
      if (de <= Real(0)) unc = (Real)1e9;
      const int32_t pre = filled[idx];
      Real a, b, mu, s2;
      if (pre > 0) {
        const Real* s = seeds + 4 * (size_t)pre;
        a = s[0];
        b = s[1];
        mu = s[2];
        s2 = s[3];
      } else {
        a = 10;
        b = 10;
        mu = de;
        s2 = 2500;
      }
      if ((de - mu) * (de - mu) > unc + s2) {
        a = 10;
        b = 10;
        mu = de;
        s2 = 2500;
      }
      const Real new_sq = unc * s2 / (unc + s2);
      const Real new_mu = (de * s2 + mu * unc) / (unc + s2);
      Real c1 = (a / (a + b)) * normpdf<Real>(de, mu, unc + s2);
      Real c2 = (b / (a + b)) * Real(1) / Real(50);
      const Real norm = c1 + c2;
      c1 = c1 / norm;
      c2 = c2 / norm;
      const Real f = c1 * ((a + Real(1)) / (a + b + Real(1))) + c2 * (a / (a + b + Real(1)));
      const Real e = c1 * (((a + Real(1)) * (a + Real(2))) / ((a + b + Real(1)) * (a + b + Real(2)))) +
                     c2 * (a * (a + Real(1)) / ((a + b + Real(1)) * (a + b + Real(2))));
      const Real mu_p = c1 * new_mu + c2 * mu;
      const Real s_p = c1 * (new_sq + new_mu * new_mu) + c2 * (s2 + mu * mu) - mu_p * mu_p;
      const Real a_p = (e - f) / (f - e / f);
      const Real b_p = a_p * (Real(1) - f) / f;
      depth_out[idx] = is_goodpoint<Real>(a, b) ? mu_p : Real(-1);
      Real* so = seeds_out + 4 * idx;
      so[0] = a_p;
      so[1] = b_p;
      so[2] = mu_p;
      so[3] = s_p;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (ctypes) -- test infrastructure only.
// Layouts: images H x W row-major (pitch in bytes for u8 images); cost volume
// Hc x Wc x N; seeds H x W x 4; transforms 3x4 row-major float[12].
// ---------------------------------------------------------------------------
extern "C" {
// qm_prior.cpp
void qmo_quadtree_prior(const float* prior_depth, int W, int H, uint8_t* levels);
void qmo_prior_to_cost(const float* prior, int W, int H, const uint8_t* levels, int N, double imin, double imax,
                       float* cost, int32_t* count);
}

extern "C" {

int qmo_version() { return 1; }

void qmo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int qmo_get_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void qmo_quadtree(const uint8_t* img, int W, int H, int pitch, uint8_t* levels) {
  quadtree_impl(img, W, H, pitch, levels);
}

// inverse-depth samples (src/depth_extract.cu:179): (float)(1.0/(step*d + min))
void qmo_depth_table_f32(int N, double imin, double imax, float* out) {
  const double step = (imax - imin) / (double)(N - 1);
  for (int d = 0; d < N; ++d) out[d] = (float)(1.0 / (step * d + imin));
}
void qmo_depth_table_f64(int N, double imin, double imax, double* out) {
  const double step = (imax - imin) / (double)(N - 1);
  for (int d = 0; d < N; ++d) out[d] = 1.0 / (step * d + imin);
}

void qmo_cost_f32(const uint8_t* key, const uint8_t* const* refs, int W, int H, int pitch, const float* T, float fx,
                  float fy, float cx, float cy, int N, const float* table, const uint8_t* levels, float* cost,
                  int32_t* count, int arith) {
  Cam c{fx, fy, cx, cy};
  cost_impl<float>(key, refs, W, H, pitch, T, c, N, table, levels, cost, count, arith);
}
void qmo_cost_f64(const uint8_t* key, const uint8_t* const* refs, int W, int H, int pitch, const float* T, float fx,
                  float fy, float cx, float cy, int N, const double* table, const uint8_t* levels, double* cost,
                  int32_t* count, int arith) {
  Cam c{fx, fy, cx, cy};
  cost_impl<double>(key, refs, W, H, pitch, T, c, N, table, levels, cost, count, arith);
}

static BpParams make_bp(int N, float P1, float P2, const int* iters, double imin, double imax) {
  BpParams p;
  p.N = N;
  p.P1 = P1;
  p.P2 = P2;
  for (int i = 0; i < 4; ++i) p.iters[i] = iters[i];
  p.idepth_min = imin;
  p.idepth_step = (imax - imin) / (double)(N - 1);
  return p;
}

void qmo_bp_extract_f32(const float* cost, int W, int H, const uint8_t* levels, int N, float P1, float P2,
                        const int* iters, double imin, double imax, float* depth_coarse, int32_t* label,
                        float* msg_out) {
  bp_extract_impl<float>(cost, W, H, levels, make_bp(N, P1, P2, iters, imin, imax), depth_coarse, label, msg_out);
}
void qmo_bp_extract_f64(const double* cost, int W, int H, const uint8_t* levels, int N, float P1, float P2,
                        const int* iters, double imin, double imax, double* depth_coarse, int32_t* label,
                        double* msg_out) {
  bp_extract_impl<double>(cost, W, H, levels, make_bp(N, P1, P2, iters, imin, imax), depth_coarse, label, msg_out);
}

void qmo_weights_f32(const uint8_t* img, int W, int H, int pitch, float* w_row, float* w_col) {
  weights_impl<float>(img, W, H, pitch, w_row, w_col);
}
void qmo_weights_f64(const uint8_t* img, int W, int H, int pitch, double* w_row, double* w_col) {
  weights_impl<double>(img, W, H, pitch, w_row, w_col);
}
float qmo_weight_of_delta_f32(int delta) { return weight_of_delta<float>(delta); }

void qmo_upsample_f32(int W, int H, const float* d, const float* conf, const float* w_row, const float* w_col,
                      double lam, int row_seg, int col_seg, float* out, float* row_x, float* row_n, int solver) {
  upsample_impl<float>(W, H, d, conf, w_row, w_col, lam, row_seg, col_seg, out, row_x, row_n, solver);
}
void qmo_upsample_f64(int W, int H, const double* d, const double* conf, const double* w_row, const double* w_col,
                      double lam, int row_seg, int col_seg, double* out, double* row_x, double* row_n, int solver) {
  upsample_impl<double>(W, H, d, conf, w_row, w_col, lam, row_seg, col_seg, out, row_x, row_n, solver);
}

void qmo_fuse_f32(int W, int H, float fx, float fy, float cx, float cy, const float* T, float* seeds,
                  const float* depth_in, float* depth_out, float* seeds_out, int32_t* table_out) {
  Cam c{fx, fy, cx, cy};
  fuse_impl<float>(W, H, c, T, seeds, depth_in, depth_out, seeds_out, table_out);
}
void qmo_fuse_f64(int W, int H, float fx, float fy, float cx, float cy, const float* T, double* seeds,
                  const double* depth_in, double* depth_out, double* seeds_out, int32_t* table_out) {
  Cam c{fx, fy, cx, cy};
  fuse_impl<double>(W, H, c, T, seeds, depth_in, depth_out, seeds_out, table_out);
}

void qmo_fuse_update_f32(int W, int H, const int32_t* table, const float* seeds_warped, const float* depth_in,
                         float* depth_out, float* seeds_out) {
  fuse_update_impl<float>(W, H, table, seeds_warped, depth_in, depth_out, seeds_out);
}
void qmo_fuse_update_f64(int W, int H, const int32_t* table, const double* seeds_warped, const double* depth_in,
                         double* depth_out, double* seeds_out) {
  fuse_update_impl<double>(W, H, table, seeds_warped, depth_in, depth_out, seeds_out);
}

// Sparse map handed to the upsample (src/bpextract.cu:370 writes the node
// depths over the semi-dense pre-fill of debug_image).
void qmo_compose_sparse_f32(int W, int H, const float* depth_coarse, const int32_t* label, const float* prefill,
                            float* sparse) {
  const int Wc = W / 4, Hc = H / 4;
  for (size_t i = 0; i < (size_t)W * H; ++i) sparse[i] = prefill ? prefill[i] : 0.0f;
  for (int y = 0; y < Hc; ++y)
    for (int x = 0; x < Wc; ++x)
      if (label[(size_t)y * Wc + x] >= 0) sparse[(size_t)(4 * y) * W + 4 * x] = depth_coarse[(size_t)y * Wc + x];
}
void qmo_compose_sparse_f64(int W, int H, const double* depth_coarse, const int32_t* label, const double* prefill,
                            double* sparse) {
  const int Wc = W / 4, Hc = H / 4;
  for (size_t i = 0; i < (size_t)W * H; ++i) sparse[i] = prefill ? prefill[i] : 0.0;
  for (int y = 0; y < Hc; ++y)
    for (int x = 0; x < Wc; ++x)
      if (label[(size_t)y * Wc + x] >= 0) sparse[(size_t)(4 * y) * W + 4 * x] = depth_coarse[(size_t)y * Wc + x];
}

// One keyframe through every stage (src/seed_matrix.cu:448-551 then :553-597),
// FP32; prior / prior_depth (nullable) switch on the prior-driven alternates
// of qm_prior.cpp after the quadtree and the cost volume. stage_ms receives quadtree, cost, bp, upsample, fusion wall times.
// seeds is the persistent fusion state (updated: it receives the new seeds).
void qmo_keyframe_f32(const uint8_t* key, const uint8_t* const* refs, int W, int H, const float* T_refs, float fx,
                      float fy, float cx, float cy, int N, double imin, double imax, const float* prefill,
                      const float* prior, const float* prior_depth, float* seeds, const float* T_last_to_cur,
                      float* depth_out, float* sparse_out, uint8_t* levels_out, int32_t* label_out, double* stage_ms,
                      int arith, int solver) {
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  const int Wc = W / 4, Hc = H / 4;
  const size_t npx = (size_t)W * H, nc = (size_t)Wc * Hc;
  std::vector<uint8_t> lv_local;
  uint8_t* levels = levels_out;
  if (!levels) {
    lv_local.resize(npx);
    levels = lv_local.data();
  }
  std::vector<float> table(N), cost(nc * N), dcoarse(nc), wr((size_t)(W + 1) * H), wc((size_t)W * (H + 1));
  std::vector<int32_t> lab_local;
  int32_t* label = label_out;
  if (!label) {
    lab_local.resize(nc);
    label = lab_local.data();
  }
  std::vector<float> sp_local, dense(npx), seeds_new(npx * 4);
  float* sparse = sparse_out;
  if (!sparse) {
    sp_local.resize(npx);
    sparse = sp_local.data();
  }
  qmo_depth_table_f32(N, imin, imax, table.data());
  Cam c{fx, fy, cx, cy};
  const int iters[4] = {4, 10, 10, 10};
  auto t0 = clk::now();
  quadtree_impl(key, W, H, W, levels);
  if (prior_depth) qmo_quadtree_prior(prior_depth, W, H, levels);
  auto t1 = clk::now();
  cost_impl<float>(key, refs, W, H, W, T_refs, c, N, table.data(), levels, cost.data(), nullptr, arith);
  if (prior) qmo_prior_to_cost(prior, W, H, levels, N, imin, imax, cost.data(), nullptr);
  auto t2 = clk::now();
  bp_extract_impl<float>(cost.data(), W, H, levels, make_bp(N, 0.003f, 0.01f, iters, imin, imax), dcoarse.data(),
                         label, nullptr);
  qmo_compose_sparse_f32(W, H, dcoarse.data(), label, prefill, sparse);
  auto t3 = clk::now();
  weights_impl<float>(key, W, H, W, wr.data(), wc.data());
  upsample_impl<float>(W, H, sparse, nullptr, wr.data(), wc.data(), 10.0, 16 * (W / 16), 8 * (H / 8), dense.data(),
                       nullptr, nullptr, solver);
  auto t4 = clk::now();
  fuse_impl<float>(W, H, c, T_last_to_cur, seeds, dense.data(), depth_out, seeds_new.data(), nullptr);
  std::memcpy(seeds, seeds_new.data(), sizeof(float) * npx * 4);
  auto t5 = clk::now();
  if (stage_ms) {
    stage_ms[0] = ms(t0, t1);
    stage_ms[1] = ms(t1, t2);
    stage_ms[2] = ms(t2, t3);
    stage_ms[3] = ms(t3, t4);
    stage_ms[4] = ms(t4, t5);
  }
}

}  // extern "C"
