// k_front.hip -- quadtree level assignment, multi-view cost volume and the
// cost pyramid for the hierarchical BP, for a batch of keyframes.
#include "qm_device.h"
#include "qm_launch.h"

namespace qm {

// ---------------------------------------------------------------------------
// Quadtree (reference: src/quadtree.cu:17-77).
// One lane owns one 4x4 pixel block; a wave covers four 16x16 tiles side by
// side. Inside a tile the 16 lanes are in Morton order, so the 8x8 parent of a
// block is the lane quad (xor 1, 2) and the 16x16 tile is the lane row (xor 4, 8).
// All sums are of integers, so the per-level test |n*I - sum| <= 10*n is exact.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_quadtree(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H;
  const int lane = threadIdx.x & 63;
  const int wid = (blockIdx.x - bm.start[f]) * 4 + (threadIdx.x >> 6);
  const int tiles_x = (W + 15) >> 4, tiles_y = (H + 15) >> 4;
  const int quads_x = (tiles_x + 3) >> 2;
  const int ty = wid / quads_x, tq = wid - ty * quads_x;
  if (ty >= tiles_y) return;  // whole wave
  const int tx = tq * 4 + (lane >> 4);
  const int m = lane & 15;
  const int bx = (m & 1) | ((m >> 1) & 2);
  const int by = ((m >> 1) & 1) | ((m >> 2) & 2);
  const int X0 = tx * 16 + bx * 4, Y0 = ty * 16 + by * 4;
  const bool inside = (X0 + 4 <= W) && (Y0 + 4 <= H);

  int px[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) px[i] = 0;
  if (inside) {
    const uint8_t* src = F.key + (size_t)Y0 * F.pitch + X0;
    if ((((uintptr_t)src) & 3) == 0 && (F.pitch & 3) == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t w4 = *reinterpret_cast<const uint32_t*>(src + (size_t)r * F.pitch);
#pragma unroll
        for (int c = 0; c < 4; ++c) px[r * 4 + c] = (w4 >> (8 * c)) & 0xFF;
      }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) px[r * 4 + c] = src[(size_t)r * F.pitch + c];
    }
  }
  // level 1 (2x2) and level 2 (4x4) inside the lane
  int p2 = inside ? 1 : 0;
#pragma unroll
  for (int sy = 0; sy < 2; ++sy)
#pragma unroll
    for (int sx = 0; sx < 2; ++sx) {
      const int a = px[(2 * sy) * 4 + 2 * sx], b = px[(2 * sy) * 4 + 2 * sx + 1];
      const int c = px[(2 * sy + 1) * 4 + 2 * sx], d = px[(2 * sy + 1) * 4 + 2 * sx + 1];
      const int s4 = a + b + c + d;
      if (abs(4 * a - s4) > 40 || abs(4 * b - s4) > 40 || abs(4 * c - s4) > 40 || abs(4 * d - s4) > 40) p2 = 0;
    }
  int s16 = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) s16 += px[i];
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (abs(16 * px[i] - s16) > 160) p2 = 0;
  // level 3 (8x8): lane quad
  int s64 = s16 + __shfl_xor(s16, 1, 64);
  s64 += __shfl_xor(s64, 2, 64);
  int all2 = p2 & __shfl_xor(p2, 1, 64);
  all2 &= __shfl_xor(all2, 2, 64);
  int dev3 = 1;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (abs(64 * px[i] - s64) > 640) dev3 = 0;
  dev3 &= __shfl_xor(dev3, 1, 64);
  dev3 &= __shfl_xor(dev3, 2, 64);
  const int p3 = all2 & dev3;
  // level 4 (16x16): lane row of 16
  int s256 = s64 + __shfl_xor(s64, 4, 64);
  s256 += __shfl_xor(s256, 8, 64);
  int all3 = p3 & __shfl_xor(p3, 4, 64);
  all3 &= __shfl_xor(all3, 8, 64);
  int dev4 = 1;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (abs(256 * px[i] - s256) > 2560) dev4 = 0;
  // the per-lane deviation test spans all 16 lanes of the tile
  dev4 &= __shfl_xor(dev4, 1, 64);
  dev4 &= __shfl_xor(dev4, 2, 64);
  dev4 &= __shfl_xor(dev4, 4, 64);
  dev4 &= __shfl_xor(dev4, 8, 64);
  const int p4 = all3 & dev4;
  const int lv = p4 ? 4 : (p3 ? 3 : 2);

  uint8_t* dst = F.levels + (size_t)Y0 * W + X0;
  if (inside) {
    const uint32_t v4 = 0x01010101u * (uint32_t)lv;
    if ((((uintptr_t)dst) & 3) == 0 && (W & 3) == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) *reinterpret_cast<uint32_t*>(dst + (size_t)r * W) = v4;
    } else {
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) dst[(size_t)r * W + c] = (uint8_t)lv;
    }
  } else if (X0 < W && Y0 < H) {
    for (int r = 0; r < 4 && Y0 + r < H; ++r)
      for (int c = 0; c < 4 && X0 + c < W; ++c) dst[(size_t)r * W + c] = (uint8_t)lv;
  }
}

// ---------------------------------------------------------------------------
// Cost volume (reference: src/depth_extract.cu:135-244).
// One lane per coarse cell (4x4 pixels), 64 consecutive cells per wave, looping
// over N/4 inverse-depth samples: for one sample the taps of neighbouring cells
// fall on neighbouring texels of the reference images, so a wave's gather
// touches a few cache lines instead of one per lane. The four waves of a block
// take four sample ranges of the same cells; their results meet in LDS and
// leave as contiguous rows. Non-node cells are written as zeros, so the volume
// needs no separate clear. Slots {0,0,1,2,3} read reference 0 twice: it is evaluated
// once and counted twice in the reference's slot tree ((c0+c4)+c2)+(c1+c3).
//
// Patch taps. Tap (i, j) of reference r samples texel
// round((px + u2u*i) + v2u*j, (py + u2v*i) + v2v*j), with round(v) =
// (int)((double)v + 0.5) and the texel index clamped to the image. The SAD of
// integer grey levels is an integer below 2^24, so it is accumulated exactly
// in an integer register (v_sad_u8 on zero-extended bytes) and converted once.
// When the centre texel lies in [2, W-3] x [2, H-3] and the rotation
// terms satisfy |u2u| + |v2u| <= 1.75 and |u2v| + |v2v| <= 1.75 (true for any
// rotation), every tap lies in [-0.25, W-0.75) (the float error of the two
// additions is below 2e-3 for W, H <= 16384), so the clamp can never act and
// the rounding is a single v_cvt_rpi_i32_f32 (floor(v + 0.5), exact on the
// hardware for every float; the argument is >= -0.5 here, where floor and
// truncation agree). Other transforms take the general path with the clamp.
//
// Two projection arithmetics, each shared bit for bit with the oracle:
//  * reference order (QM_COST_REFERENCE): q = R (dir * depth) + t summed as
//    ((R0 p.x + R1 p.y) + R2 p.z) + t, then two IEEE divisions
//    px = (fx q.x) / q.z + cx, py = (fy q.y) / q.z + cy, the order of
//    src/depth_extract.cu:178-191 with every product rounded.
//  * fused (QM_COST_FUSED): the per-cell products are hoisted out of the
//    sample loop, a = (fx (R0.dir), fy (R1.dir), R2.dir) with each dot product
//    as fma(R2, d.z, fma(R1, d.y, R0 d.x)), b = (fx t0, fy t1, t2); per sample
//    U = fma(depth, a, b) for the three rows, r = RN(1 / U.z) and
//    (px + 0.5, py + 0.5) = (fma(U.x, r, cx + 0.5), fma(U.y, r, cy + 0.5));
//    the texel of a tap is floor of the tap sum formed from px + 0.5 in the
//    reference's order ((p + u2u i) + v2u j). The same algebra in 4 roundings
//    instead of 12, so its errors against an FP64 evaluation are no larger.
//    RN(1 / z) is v_rcp_f32 refined by one FMA Newton step, which equals the
//    IEEE quotient for every normal |z| in [2^-126, 2^126) (checked
//    exhaustively over all such floats on gfx950); outside that range, and
//    for zero, infinities and NaN, the IEEE division is used.
// ---------------------------------------------------------------------------
// The tap bound above, evaluated on the host for every reference of a launch
// (same float arithmetic as the device); false for NaN.
static bool cost_fast_taps(const float* Tr) {
  return (fabsf(Tr[0]) + fabsf(Tr[1]) <= 1.75f) && (fabsf(Tr[4]) + fabsf(Tr[5]) <= 1.75f);
}

__device__ __forceinline__ int cvt_rpi(float v) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}
// floor(v) as an int (saturating, NaN -> 0)
__device__ __forceinline__ int cvt_flr(float v) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}
// Texel index of a tap sum: round-half-up of the reference's pixel in the
// reference order, floor of (pixel + 0.5) in the fused arithmetic.
template <bool FUSED>
__device__ __forceinline__ int tap_index(float v) {
  return FUSED ? cvt_flr(v) : cvt_rpi(v);
}

// RN(1 / z): one Newton step on the hardware reciprocal, with the IEEE
// division where that step is not exact (|z| outside [2^-126, 2^126), zero,
// infinities, NaN). The branch is wave-uniform in practice: depths in the
// reference frame are far from both limits.
__device__ __forceinline__ float recip_rn(float z) {
  const float y = __builtin_amdgcn_rcpf(z);
  const float e = fmaf(-z, y, 1.0f);
  float r = fmaf(e, y, y);
  const float az = fabsf(z);
  if (__builtin_expect(!(az >= 0x1p-126f && az < 0x1p126f), 0)) r = 1.0f / z;
  return r;
}

// RN(1 / v) for v in [1, 5] (a normal number: the Newton step is exact there).
__device__ __forceinline__ float recip_small(float v) {
  const float y = __builtin_amdgcn_rcpf(v);
  return fmaf(fmaf(-v, y, 1.0f), y, y);
}
// RN(a / b) from y = RN(1 / b): q = a y, then one correction with the exact
// residual a - b q (used where the whole input domain has been checked).
__device__ __forceinline__ float div_by_recip(float a, float b, float y) {
  const float q = a * y;
  return fmaf(fmaf(-b, q, a), y, q);
}
constexpr float kRecip255 = 1.0f / 255.0f;  // RN(1 / 255), folded by the compiler

// The fused arithmetic's floor of a tap sum on the general path, saturating
// as the oracle's (NaN -> below 0, clamped to texel 0 there).
__device__ __forceinline__ int floor_sat(float v) {
  v = fminf(fmaxf(v, -1.0e9f), 1.0e9f);
  return (int)floorf(v);
}

// Image bytes through a buffer resource: a tap's address is a 32-bit byte
// offset (no 64-bit address arithmetic per load), and an offset past the
// image reads 0 instead of touching memory outside it.
typedef __amdgpu_buffer_rsrc_t ImgRsrc;
__device__ __forceinline__ ImgRsrc img_rsrc(const uint8_t* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ unsigned img_dword(ImgRsrc r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ unsigned img_byte(ImgRsrc r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}

// Fast path: the three taps of a patch row (fixed j) usually land on three
// consecutive texels of one image row. Such a row is read with one unaligned
// dword load (its fourth byte masked off) instead of three byte gathers, which
// cuts the texture-address work that bounds this kernel; lanes whose row is not
// of that shape, or whose dword would reach past the image, gather the three
// bytes instead. Either way the SAD is the same integer.
typedef unsigned u32_unaligned __attribute__((aligned(1)));

// Offsets y * pitch + x of the nine taps, [j][i] for tap (i - 1, j - 1), on
// the fast path (no clamp); (px, py) is the pixel (reference order) or the
// pixel + 0.5 (fused).
template <bool FUSED>
__device__ __forceinline__ void tap_offsets(float px, float py, const float* Tr, int pitch, unsigned (&o)[3][3]) {
  const float u2u = Tr[0], u2v = Tr[4], v2u = Tr[1], v2v = Tr[5];
  float bx[3], by[3];
#pragma unroll
  for (int i = -1; i <= 1; ++i) {
    // u2u * i is exact (+-u2u or a signed zero, which the rounding ignores)
    bx[i + 1] = i == 0 ? px : px + (i < 0 ? -u2u : u2u);
    by[i + 1] = i == 0 ? py : py + (i < 0 ? -u2v : u2v);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float sx = j == 1 ? bx[i] : bx[i] + (j < 1 ? -v2u : v2u);
      const float sy = j == 1 ? by[i] : by[i] + (j < 1 ? -v2v : v2v);
      o[j][i] = __umul24((unsigned)tap_index<FUSED>(sy), (unsigned)pitch) + (unsigned)tap_index<FUSED>(sx);
    }
}

// General path: clamped taps, float SAD in the reference's tap order.
template <bool FUSED>
__device__ __forceinline__ float patch_sad_clamped(gptr<const uint8_t> ref, int W, int H, int pitch, float px,
                                                   float py, const float* Tr, const unsigned (&patch)[3][3]) {
  const float u2u = Tr[0], u2v = Tr[4], v2u = Tr[1], v2v = Tr[5];
  float sad = 0.0f;
#pragma unroll
  for (int j = -1; j <= 1; ++j) {
#pragma unroll
    for (int i = -1; i <= 1; ++i) {
      const float sx = (px + u2u * (float)i) + v2u * (float)j;
      const float sy = (py + u2v * (float)i) + v2v * (float)j;
      const int chx = min(max(FUSED ? floor_sat(sx) : round_half_up(sx), 0), W - 1);
      const int chy = min(max(FUSED ? floor_sat(sy) : round_half_up(sy), 0), H - 1);
      sad += fabsf((float)patch[i + 1][j + 1] - (float)ref[(size_t)chy * pitch + chx]);
    }
  }
  return sad;
}

// Levels 1-3 of the cost pyramid (reference: src/bpextract.cu:59-78,169-198)
// for the 8 x 8 level-0 tile (tx, ty) held in LDS (cell c = 8 row + column at
// t0[c * stride], samples in order): its 4 x 4 level-1, 2 x 2 level-2 cells and
// its level-3 cell. Children are summed from 0.0f in the order (2x,2y),
// (2x,2y+1), (2x+1,2y), (2x+1,2y+1), skipping those outside their level, as the
// stand-alone pyramid kernel does; the results go to the same positions (row
// layout or sample order) as level 0.
template <int N, int THREADS>
__device__ __forceinline__ void cost_pyramid_tile(const FrameDev& F, const float* t0, int stride, int tx, int ty,
                                                  int natural) {
  __shared__ float t1[16 * N], t2[4 * N];
  const int w0 = F.lw[0], h0 = F.lh[0], w1 = F.lw[1], h1 = F.lh[1], w2 = F.lw[2], h2 = F.lh[2];
  auto pos = [&](int d) __attribute__((always_inline)) { return natural ? d : row_pos<N>(d); };
  // level 1
  for (int i = threadIdx.x; i < 16 * N; i += THREADS) {
    const int c = i / N, d = i - c * N;
    const int lx = c & 3, ly = c >> 2;  // in the tile's 4 x 4 level-1 cells
    float sum = 0.0f;
#pragma unroll
    for (int i0 = 0; i0 < 2; ++i0)
#pragma unroll
      for (int j0 = 0; j0 < 2; ++j0) {
        const int cx = 2 * lx + i0, cy = 2 * ly + j0;
        if (8 * tx + cx < w0 && 8 * ty + cy < h0) sum += t0[(8 * cy + cx) * stride + d];
      }
    t1[c * N + d] = sum;
    const int x1 = 4 * tx + lx, y1 = 4 * ty + ly;
    if (x1 < w1 && y1 < h1) gp(F.cost[1])[((size_t)y1 * w1 + x1) * N + pos(d)] = sum;
  }
  __syncthreads();
  // level 2
  for (int i = threadIdx.x; i < 4 * N; i += THREADS) {
    const int c = i / N, d = i - c * N;
    const int lx = c & 1, ly = c >> 1;
    float sum = 0.0f;
#pragma unroll
    for (int i1 = 0; i1 < 2; ++i1)
#pragma unroll
      for (int j1 = 0; j1 < 2; ++j1) {
        const int cx = 2 * lx + i1, cy = 2 * ly + j1;
        if (4 * tx + cx < w1 && 4 * ty + cy < h1) sum += t1[(4 * cy + cx) * N + d];
      }
    t2[c * N + d] = sum;
    const int x2 = 2 * tx + lx, y2 = 2 * ty + ly;
    if (x2 < w2 && y2 < h2) gp(F.cost[2])[((size_t)y2 * w2 + x2) * N + pos(d)] = sum;
  }
  __syncthreads();
  // level 3: one cell, the tile's own
  for (int d = threadIdx.x; d < N; d += THREADS) {
    float sum = 0.0f;
#pragma unroll
    for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
      for (int j2 = 0; j2 < 2; ++j2)
        if (2 * tx + i2 < w2 && 2 * ty + j2 < h2) sum += t2[(2 * j2 + i2) * N + d];
    gp(F.cost[3])[((size_t)ty * F.lw[3] + tx) * N + pos(d)] = sum;
  }
}

template <int N, bool FAST, int WAVES, bool FUSED>
__global__ __launch_bounds__(64 * WAVES) void k_cost(const FrameDev* __restrict__ frames, BlockMap bm,
                                                     const float* __restrict__ depth_table, int nblocks, int natural,
                                                     int pyramid) {
  // block = an 8 x 8 tile of cells x N samples (lane = cell); wave w takes
  // samples [w*N/WAVES, (w+1)*N/WAVES) of every cell
  constexpr int kPerWave = N / WAVES;
  constexpr int kRow = N + 1;  // LDS row stride in floats (conflict-free column writes)
  __shared__ float tile[64 * kRow];
  const int lb_global = xcd_logical_block(blockIdx.x, nblocks);
  const int f = find_frame(bm, lb_global);
  const FrameDev& F = frames[f];
  const int W = F.W, H = F.H, Wc = F.Wc, Hc = F.Hc;
  const int tiles_x = (Wc + 7) >> 3;
  const int tl = lb_global - bm.start[f];
  const int ty = tl / tiles_x, tx = tl - ty * tiles_x;
  const int lane = threadIdx.x & 63;
  // the wave index is uniform: the sample loop then reads the depth table
  // through scalar loads
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cxx = 8 * tx + (lane & 7), cyy = 8 * ty + (lane >> 3);
  const int cell = cyy * Wc + cxx;
  const int dbase = kPerWave * w;
  float* trow = tile + lane * kRow + dbase;
  if (cxx < Wc && cyy < Hc) {
    const int x = 4 * cxx, y = 4 * cyy;
    const int qs = 1 << gp((const uint8_t*)F.levels)[(size_t)y * W + x];
    const bool node = !(x >= W - 1 || y >= H - 1 || x <= 0 || y <= 0 || (x % qs) != 0 || (y % qs) != 0);
    if (F.count && w == 0) gp(F.count)[cell] = node ? 64 : 0;
    if (!node) {
#pragma unroll
      for (int t = 0; t < kPerWave; ++t) trow[t] = 0.0f;
    } else {
      // descriptor fields into registers once (uniform: scalar registers)
      const int pitch = F.pitch;
      const float fx = F.fx, fy = F.fy, cx = F.cx, cy = F.cy;
      float T[4][12];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int i = 0; i < 12; ++i) T[r][i] = F.T[r][i];
      const unsigned img_bytes = (unsigned)H * (unsigned)pitch;
      gptr<const uint8_t> ref[4];
      ImgRsrc rs[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        ref[r] = gp(F.ref[r]);
        rs[r] = img_rsrc(F.ref[r], img_bytes);
      }
      const gptr<const uint8_t> key = gp(F.key);
      unsigned patch[3][3], prow[3];
#pragma unroll
      for (int j = -1; j <= 1; ++j)
#pragma unroll
        for (int i = -1; i <= 1; ++i) patch[i + 1][j + 1] = key[(y + j) * pitch + (x + i)];
      // patch row j as bytes i = -1, 0, 1 (little-endian, like a texel dword)
#pragma unroll
      for (int j = 0; j < 3; ++j) prow[j] = patch[0][j] | (patch[1][j] << 8) | (patch[2][j] << 16);
      const V3 dir = ray_dir(x, y, fx, fy, cx, cy);
      // fused arithmetic: the per-cell rows a = (fx R0.dir, fy R1.dir, R2.dir)
      float ax[4], ay[4], az[4], bx[4], by[4];
      if (FUSED) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          bx[r] = fx * T[r][3];
          by[r] = fy * T[r][7];
          ax[r] = fx * fmaf(T[r][2], dir.z, fmaf(T[r][1], dir.y, T[r][0] * dir.x));
          ay[r] = fy * fmaf(T[r][6], dir.z, fmaf(T[r][5], dir.y, T[r][4] * dir.x));
          az[r] = fmaf(T[r][10], dir.z, fmaf(T[r][9], dir.y, T[r][8] * dir.x));
        }
      }
      // in-view bounds: 2 <= (int)((double)p + 0.5) < W - 2 is 1.5 <= p < W - 2.5
      // on the reference-order pixel and 2 <= p5 < W - 2 on the fused p + 0.5
      // (exact for W, H <= 16384; false for NaN)
      const float lo = FUSED ? 2.0f : 1.5f;
      const float wlim = FUSED ? (float)W - 2.0f : (float)W - 2.5f;
      const float hlim = FUSED ? (float)H - 2.0f : (float)H - 2.5f;
      const float safe = FUSED ? 2.5f : 2.0f;  // an in-bounds centre for lanes out of view
      const float cx5 = cx + 0.5f, cy5 = cy + 0.5f;
      const unsigned last4 = img_bytes - 4u;  // last offset a dword may start at
#pragma unroll 2
      for (int t = 0; t < kPerWave; ++t) {
        const float depth = depth_table[dbase + t];
        // all four projections and view tests first, then all taps: the texel
        // loads and their SADs then sit in one basic block. nr[r] is 1 for a
        // reference that sees the sample, else 0.
        float pxs[4], pys[4], nr[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float pxf, pyf;
          if (FUSED) {
            const float U = fmaf(depth, ax[r], bx[r]);
            const float V = fmaf(depth, ay[r], by[r]);
            const float Z = fmaf(depth, az[r], T[r][11]);
            const float rz = recip_rn(Z);
            pxf = fmaf(U, rz, cx5);
            pyf = fmaf(V, rz, cy5);
          } else {
            V3 P;
            P.x = dir.x * depth;
            P.y = dir.y * depth;
            P.z = dir.z * depth;
            const V3 q = se3_apply(T[r], P);
            pxf = (fx * q.x) / q.z + cx;
            pyf = (fy * q.y) / q.z + cy;
          }
          const bool in = pxf >= lo && pxf < wlim && pyf >= lo && pyf < hlim;
          nr[r] = in ? 1.0f : 0.0f;
          // out of view, the fast path samples around a safe centre, always in
          // bounds, and drops the result, so the taps of all references issue
          // together
          pxs[r] = (FAST && !in) ? safe : pxf;
          pys[r] = (FAST && !in) ? safe : pyf;
        }
        float cr[4];
        if (FAST) {
          // Tap offsets of every reference, then one dword per patch row where
          // the row's three taps are consecutive bytes (offsets o, o + 1, o + 2,
          // with o + 4 inside the image), and byte gathers, in a divergent
          // branch, for the rows of other lanes; t keeps the offsets, replaced
          // by the texels in such rows, which are then packed into the row's
          // dword. All loads are issued before any SAD.
          unsigned t[4][3][3], v[4][3];
          bool rw[4][3];
          bool gathered = false;  // this lane gathered some row
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            tap_offsets<FUSED>(pxs[r], pys[r], T[r], pitch, t[r]);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              // (t1, t2) == (t0 + 1, t0 + 2) as one 64-bit compare
              const unsigned long long got = ((unsigned long long)t[r][j][2] << 32) | t[r][j][1];
              const unsigned long long want = ((unsigned long long)(t[r][j][0] + 2u) << 32) | (t[r][j][0] + 1u);
              rw[r][j] = (got == want) & (t[r][j][0] <= last4);
              // unconditional: a dword past the image end (a row starting in
              // its last three bytes, which also gathers) reads as zero, no
              // fault, through the buffer's range check
              v[r][j] = img_dword(rs[r], t[r][j][0]);
            }
#pragma unroll
            for (int j = 0; j < 3; ++j)
              if (!rw[r][j]) {
#pragma unroll
                for (int i = 0; i < 3; ++i) t[r][j][i] = img_byte(rs[r], t[r][j][i]);
                gathered = true;
              }
          }
          // the gathered rows' three texels packed like a dword row: one
          // wave-uniform branch, skipped (scalar) when no lane gathered any
          // row, as with rotation-free poses, where every row is contiguous
          if (__builtin_amdgcn_ballot_w64(gathered)) {
            asm volatile("");  // a real branch: not speculated into selects
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
              for (int j = 0; j < 3; ++j) {
                // bytes (t0, t1, 0, 0), then (t0, t1, t2, 0): two byte permutes
                const unsigned lo = __builtin_amdgcn_perm(t[r][j][1], t[r][j][0], 0x0C0C0400u);
                const unsigned g = __builtin_amdgcn_perm(t[r][j][2], lo, 0x0C040100u);
                v[r][j] = rw[r][j] ? v[r][j] : g;
              }
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            unsigned sad = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) sad = __builtin_amdgcn_sad_u8(v[r][j] & 0xFFFFFFu, prow[j], sad);
            cr[r] = nr[r] != 0.0f ? (float)sad : 0.0f;
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            cr[r] = nr[r] != 0.0f ? patch_sad_clamped<FUSED>(ref[r], W, H, pitch, pxs[r], pys[r], T[r], patch) : 0.0f;
        }
        const float tot = ((cr[0] + cr[3]) + cr[1]) + (cr[0] + cr[2]);
        const float n = ((nr[0] + nr[3]) + nr[1]) + (nr[0] + nr[2]);
        // (tot / n) / 255 with both quotients correctly rounded, as FMA-corrected
        // products with the reciprocal (exact for every integer tot <= 5*9*255
        // and n in 1..5, and for every such quotient over 255: checked over the
        // whole domain); normalize_the_cost divides by add_num, which each of
        // the 64 sample threads of a node's block incremented once: an exact
        // / 64. Branch-free: the quotient is computed for every lane (n >= 1
        // where used).
        const float nn = fmaxf(n, 1.0f);
        const float q = div_by_recip(div_by_recip(tot, nn, recip_small(nn)), 255.0f, kRecip255);
        trow[t] = (n > 0.0f ? q : 100.0f) * 0.015625f;
      }
    }
  }
  __syncthreads();
  // each tile row's 8 cost rows are contiguous in the volume: written as
  // float4, in the row layout of the pipeline or, for callers, in sample
  // order, with streaming stores (BP reads them in later launches)
  constexpr int kQ = N / 4;  // float4 per row
#pragma unroll
  for (int k = 0; k < (16 * N) / (64 * WAVES); ++k) {
    const int i = threadIdx.x + 64 * WAVES * k;  // float4 index in the tile
    const int c = i / kQ, q = i - (i / kQ) * kQ;
    const int ox = 8 * tx + (c & 7), oy = 8 * ty + (c >> 3);
    if (ox < Wc && oy < Hc) {
      const float* src = tile + c * kRow;
      float4 v;
      if (natural) {
        v = make_float4(src[4 * q], src[4 * q + 1], src[4 * q + 2], src[4 * q + 3]);
      } else {
        v = make_float4(src[row_sample<N>(4 * q)], src[row_sample<N>(4 * q + 1)], src[row_sample<N>(4 * q + 2)],
                        src[row_sample<N>(4 * q + 3)]);
      }
      store4_nt((gptr<float4>)(gp(F.cost[0]) + ((size_t)oy * Wc + ox) * N) + q, v);
    }
  }
  if (pyramid) cost_pyramid_tile<N, 64 * WAVES>(F, tile, kRow, tx, ty, natural);
}

// ---------------------------------------------------------------------------
// Cost pyramid (reference: src/bpextract.cu:59-78,169-198). One wave builds
// one level-3 cell and its whole subtree (4 level-2, 16 level-1 cells), reading
// level 0 once. Children are summed from 0.0f in the order (2x,2y), (2x,2y+1),
// (2x+1,2y), (2x+1,2y+1), skipping those outside the level.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_pyramid(const FrameDev* __restrict__ frames, BlockMap bm) {
  const int f = find_frame(bm, blockIdx.x);
  const FrameDev& F = frames[f];
  const int lane = threadIdx.x & 63;
  const int c3 = (blockIdx.x - bm.start[f]) * 4 + (threadIdx.x >> 6);
  if (c3 >= F.lw[3] * F.lh[3]) return;
  const int y3 = c3 / F.lw[3], x3 = c3 - y3 * F.lw[3];
  const gptr<const float> C0 = gp((const float*)F.cost[0]);
  const gptr<float> C1 = gp(F.cost[1]);
  const gptr<float> C2 = gp(F.cost[2]);
  const gptr<float> C3 = gp(F.cost[3]);
  const int w0 = F.lw[0], h0 = F.lh[0], w1 = F.lw[1], h1 = F.lh[1], w2 = F.lw[2], h2 = F.lh[2];
#pragma unroll
  for (int k = 0; k < N / 64; ++k) {
    const int d = lane + 64 * k;
    float s3 = 0.0f;
    for (int i2 = 0; i2 < 2; ++i2)
      for (int j2 = 0; j2 < 2; ++j2) {
        const int x2 = 2 * x3 + i2, y2 = 2 * y3 + j2;
        if (x2 >= w2 || y2 >= h2) continue;
        float s2 = 0.0f;
        for (int i1 = 0; i1 < 2; ++i1)
          for (int j1 = 0; j1 < 2; ++j1) {
            const int x1 = 2 * x2 + i1, y1 = 2 * y2 + j1;
            if (x1 >= w1 || y1 >= h1) continue;
            float s1 = 0.0f;
            for (int i0 = 0; i0 < 2; ++i0)
              for (int j0 = 0; j0 < 2; ++j0) {
                const int xx = 2 * x1 + i0, yy = 2 * y1 + j0;
                if (xx < w0 && yy < h0) s1 += C0[((size_t)yy * w0 + xx) * N + d];
              }
            C1[((size_t)y1 * w1 + x1) * N + d] = s1;
            s2 += s1;
          }
        C2[((size_t)y2 * w2 + x2) * N + d] = s2;
        s3 += s2;
      }
    C3[((size_t)c3) * N + d] = s3;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_quadtree(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) {
    const int tx = (hframes[i].W + 15) / 16, ty = (hframes[i].H + 15) / 16;
    counts[i] = (((tx + 3) / 4) * ty + 3) / 4;
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb > 0) hipLaunchKernelGGL(k_quadtree, dim3(nb), dim3(256), 0, s, dframes, bm);
}

template <int N, bool FAST, bool FUSED>
static void launch_cost_shape(int waves, int nb, hipStream_t s, const FrameDev* dframes, const BlockMap& bm,
                              const float* table, int natural, int pyramid) {
  if (waves == 2)
    hipLaunchKernelGGL((k_cost<N, FAST, 2, FUSED>), dim3(nb), dim3(128), 0, s, dframes, bm, table, nb, natural,
                       pyramid);
  else if (waves == 8)
    hipLaunchKernelGGL((k_cost<N, FAST, 8, FUSED>), dim3(nb), dim3(512), 0, s, dframes, bm, table, nb, natural,
                       pyramid);
  else
    hipLaunchKernelGGL((k_cost<N, FAST, 4, FUSED>), dim3(nb), dim3(256), 0, s, dframes, bm, table, nb, natural,
                       pyramid);
}

template <int N>
static void launch_cost_n(bool fast, bool fused, int waves, int nb, hipStream_t s, const FrameDev* dframes,
                          const BlockMap& bm, const float* table, int natural, int pyramid) {
  if (fused) {
    if (fast) launch_cost_shape<N, true, true>(waves, nb, s, dframes, bm, table, natural, pyramid);
    else launch_cost_shape<N, false, true>(waves, nb, s, dframes, bm, table, natural, pyramid);
  } else {
    if (fast) launch_cost_shape<N, true, false>(waves, nb, s, dframes, bm, table, natural, pyramid);
    else launch_cost_shape<N, false, false>(waves, nb, s, dframes, bm, table, natural, pyramid);
  }
}

void launch_cost(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const float* table, bool natural,
                 int arith, const Tuning& t, hipStream_t s, bool pyramid) {
  BlockMap bm;
  int counts[kMaxFrames];
  bool fast = true;
  for (int i = 0; i < n; ++i) {
    counts[i] = ((hframes[i].Wc + 7) / 8) * ((hframes[i].Hc + 7) / 8);  // 8 x 8 tiles of cells
    for (int r = 0; r < 4; ++r) fast = fast && cost_fast_taps(hframes[i].T[r]);
  }
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  const int pyr = pyramid ? 1 : 0;
  const bool fused = arith == QM_COST_FUSED;
  if (N == 64) launch_cost_n<64>(fast, fused, t.cost_waves, nb, s, dframes, bm, table, (int)natural, pyr);
  else launch_cost_n<128>(fast, fused, t.cost_waves, nb, s, dframes, bm, table, (int)natural, pyr);
}

void launch_pyramid(const FrameDev* dframes, const FrameDev* hframes, int n, int N, hipStream_t s) {
  BlockMap bm;
  int counts[kMaxFrames];
  for (int i = 0; i < n; ++i) counts[i] = (hframes[i].lw[3] * hframes[i].lh[3] + 3) / 4;
  const int nb = make_block_map(bm, counts, n);
  if (nb <= 0) return;
  if (N == 64)
    hipLaunchKernelGGL(k_pyramid<64>, dim3(nb), dim3(256), 0, s, dframes, bm);
  else
    hipLaunchKernelGGL(k_pyramid<128>, dim3(nb), dim3(256), 0, s, dframes, bm);
}

}  // namespace qm
