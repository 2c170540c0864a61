// k_semidense.hip -- semi-dense (LSD-style) keyframe seeds that pre-fill the
// sparse map every frame, and the image gradient that feeds them (reference:
// src/seed_init.cu:26-42, src/lsd_semidense.cu:64-753, drivers
// src/seed_matrix.cu:211-304,399-446).
//
// One thread per pixel throughout. Seeds are two float4 per pixel:
// (valid, blacklist, validity counter, 0) and (idepth, variance, smoothed
// idepth, smoothed variance), the reference's DepthSeed fields.
// Pinned semantics shared with the oracle:
// * FP32 without contraction, IEEE division and sqrt; the reference's double
//   literals keep double arithmetic ("+ 0.5", "2.0 *", "+= 4.0", the result
//   variance factor);
// * texture fetches at texel centres are exact; bilinear fetches (the
//   epipolar search) weight the four texels by (256-a)(256-b), a(256-b), ...
//   with a, b the fractional coordinates rounded to 1/256 (CUDA's 8-bit
//   filtering weights), computed exactly in integers on the integer images;
// * the reference's races are resolved deterministically: regularisation and
//   hole filling read a snapshot (kernels write a second buffer); the
//   propagation table keeps the largest inverse depth (ties: lower source
//   index); the sparse-map projection keeps the nearest depth (ties: lower
//   source index).
#include "qm_device.h"
#include "qm_internal.h"

namespace qm {

namespace sd {

constexpr float kMinGrad = 5.0f;
constexpr double kMinInvDepth = 0.01;
constexpr double kMaxInvDepth = 2.0;
constexpr double kVarianceMax = 4.0;
constexpr float kMaxDiffConst = 40.0f * 40.0f;
constexpr float kMaxDiffGradMult = 0.5f * 0.5f;
constexpr float kRegDistVar = 0.075f * 0.075f;
constexpr float kValSumMinCreate = 24.0f, kValSumMinKeep = 24.0f, kValSumMinUnblacklist = 100.0f;
constexpr float kMinBlacklist = -1.0f;
constexpr int kMaxSearch = 100;
constexpr float kMaxErrorStereo = 1300.0f;
constexpr float kMinDistanceErrorStereo = 1.5f;
constexpr float kValidityInc = 5.0f, kValidityDec = 5.0f, kValidityInitial = 5.0f;
constexpr float kValidityMax = 5.0f, kValidityMaxVariable = 250.0f;
constexpr float kSuccVarIncFac = 1.01f, kFailVarIncFac = 1.1f;

struct Params {
  int W, H, pitch;
  float fx, fy, cx, cy;
  float k2i[12];  // keyframe -> income
  float i2k[12];  // income -> keyframe
  const uint8_t* key;
  const float2* key_grad;
  const uint8_t* inc;
  const float2* inc_grad;
};

__device__ __forceinline__ bool interior(int x, int y, int W, int H) {
  return !(x >= W - 3 || y >= H - 3 || x <= 2 || y <= 2);
}

__device__ __forceinline__ float texel(const uint8_t* img, int W, int H, int pitch, int x, int y) {
  x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
  y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
  return (float)img[(size_t)y * pitch + x];
}

__device__ float tex_bilinear(const uint8_t* img, int W, int H, int pitch, float u, float v) {
  float xb = u - 0.5f, yb = v - 0.5f;
  xb = fminf(fmaxf(xb, -2.0f), (float)W + 1.0f);
  yb = fminf(fmaxf(yb, -2.0f), (float)H + 1.0f);
  const float fx0 = floorf(xb), fy0 = floorf(yb);
  const int a = (int)rintf((xb - fx0) * 256.0f), b = (int)rintf((yb - fy0) * 256.0f);
  const int i0 = (int)fx0, j0 = (int)fy0;
  const int ia = i0 < 0 ? 0 : (i0 > W - 1 ? W - 1 : i0);
  const int ib = i0 + 1 < 0 ? 0 : (i0 + 1 > W - 1 ? W - 1 : i0 + 1);
  const int ja = j0 < 0 ? 0 : (j0 > H - 1 ? H - 1 : j0);
  const int jb = j0 + 1 < 0 ? 0 : (j0 + 1 > H - 1 ? H - 1 : j0 + 1);
  const int t00 = img[(size_t)ja * pitch + ia], t10 = img[(size_t)ja * pitch + ib];
  const int t01 = img[(size_t)jb * pitch + ia], t11 = img[(size_t)jb * pitch + ib];
  const int s = (256 - a) * (256 - b) * t00 + a * (256 - b) * t10 + (256 - a) * b * t01 + a * b * t11;
  return (float)s * (1.0f / 65536.0f);
}

__device__ __forceinline__ float3 rotate(const float* T, float3 p) {
  return make_float3((T[0] * p.x + T[1] * p.y) + T[2] * p.z, (T[4] * p.x + T[5] * p.y) + T[6] * p.z,
                     (T[8] * p.x + T[9] * p.y) + T[10] * p.z);
}
__device__ __forceinline__ float3 transform(const float* T, float3 p) {
  const float3 r = rotate(T, p);
  return make_float3(r.x + T[3], r.y + T[7], r.z + T[11]);
}

// src/lsd_semidense.cu:529-753
__device__ float search_point(int x, int y, const Params& P, float epx, float epy, float gradient_max,
                              float my_gradient, float& result_idep, float& result_var) {
  const int W = P.W, H = P.H;
  const float min_idep = (float)kMinInvDepth, max_idep = (float)kMaxInvDepth;
  float this_patch[5];
#pragma unroll
  for (int i = -2; i <= 2; ++i)
    this_patch[i + 2] = tex_bilinear(P.key, W, H, P.pitch, (float)((double)((float)x + (float)i * epx) + 0.5),
                                     (float)((double)((float)y + (float)i * epy) + 0.5));
  const float3 KiP = make_float3(((float)x - P.cx) / P.fx, ((float)y - P.cy) / P.fy, 1.0f);
  const float3 RKiP = rotate(P.k2i, KiP);
  const float3 T = make_float3(P.k2i[3], P.k2i[7], P.k2i[11]);
  const float3 KRKiP = make_float3(P.fx * RKiP.x + P.cx * RKiP.z, P.fy * RKiP.y + P.cy * RKiP.z, RKiP.z);
  const float3 KT = make_float3(P.fx * T.x + P.cx * T.z, P.fy * T.y + P.cy * T.z, T.z);
  const float3 np = make_float3(KRKiP.x + KT.x * max_idep, KRKiP.y + KT.y * max_idep, KRKiP.z + KT.z * max_idep);
  const float3 fp = make_float3(KRKiP.x + KT.x * min_idep, KRKiP.y + KT.y * min_idep, KRKiP.z + KT.z * min_idep);
  float near_u = np.x / np.z, near_v = np.y / np.z;
  float far_u = fp.x / fp.z, far_v = fp.y / fp.z;
  const float du = near_u - far_u, dv = near_v - far_v;
  float search_length = sqrtf(du * du + dv * dv);
  const float sdx = du / search_length, sdy = dv / search_length;
  if ((double)search_length < 3.0) {
    near_u += sdx * 2.0f;
    near_v += sdy * 2.0f;
    far_u -= sdx * 2.0f;
    far_v -= sdy * 2.0f;
    search_length = (float)((double)search_length + 4.0);
  }
  if (far_u <= 1.0f || far_v <= 1.0f || far_u >= (float)(W - 2) || far_v >= (float)(H - 2)) return -1.0f;
  float that_patch[5];
#pragma unroll
  for (int i = -2; i <= 1; ++i)
    that_patch[i + 2] = tex_bilinear(P.inc, W, H, P.pitch, (float)((double)(far_u + (float)i * sdx) + 0.5),
                                     (float)((double)(far_v + (float)i * sdy) + 0.5));
  float best_step = -1.0f, best_score = 1e9f, pre_best_score = 1e9f, post_best_score = 1e9f;
  bool last_is_best = false;
  float best_pre = __int_as_float(0x7fc00000), best_post = __int_as_float(0x7fc00000);
  const float qnan = __int_as_float(0x7fc00000);
  float e1A = qnan, e1B = qnan, e2A = qnan, e2B = qnan, e3A = qnan, e3B = qnan, e4A = qnan, e4B = qnan, e5A = qnan,
        e5B = qnan;
  float second_best_score = 1e9f, second_best_step = -1.0f;
  float pre_score = 1e9f, this_score = 1e9f;
  for (int step = 0; (float)step <= search_length && step <= kMaxSearch; step += 1) {
    pre_score = this_score;
    this_score = 0.0f;
    const float spu = far_u + (float)step * sdx, spv = far_v + (float)step * sdy;
    if (spu <= 1.0f || spu >= (float)(W - 1) || spv <= 1.0f || spv >= (float)(H - 1)) break;
    that_patch[4] = tex_bilinear(P.inc, W, H, P.pitch, (float)(((double)spu + 2.0 * (double)sdx) + 0.5),
                                 (float)(((double)spv + 2.0 * (double)sdy) + 0.5));
    if (step % 2 == 0) {
      e1A = that_patch[4] - this_patch[4]; this_score += e1A * e1A;
      e2A = that_patch[3] - this_patch[3]; this_score += e2A * e2A;
      e3A = that_patch[2] - this_patch[2]; this_score += e3A * e3A;
      e4A = that_patch[1] - this_patch[1]; this_score += e4A * e4A;
      e5A = that_patch[0] - this_patch[0]; this_score += e5A * e5A;
    } else {
      e1B = that_patch[4] - this_patch[4]; this_score += e1B * e1B;
      e2B = that_patch[3] - this_patch[3]; this_score += e2B * e2B;
      e3B = that_patch[2] - this_patch[2]; this_score += e3B * e3B;
      e4B = that_patch[1] - this_patch[1]; this_score += e4B * e4B;
      e5B = that_patch[0] - this_patch[0]; this_score += e5B * e5B;
    }
    if (last_is_best) {
      post_best_score = this_score;
      best_post = (((e1A * e1B + e2A * e2B) + e3A * e3B) + e4A * e4B) + e5A * e5B;
      last_is_best = false;
    }
    if (this_score < best_score) {
      second_best_score = best_score;
      second_best_step = best_step;
      best_score = this_score;
      best_step = (float)step;
      pre_best_score = pre_score;
      best_pre = (((e1A * e1B + e2A * e2B) + e3A * e3B) + e4A * e4B) + e5A * e5B;
      best_post = -1.0f;
      post_best_score = 1e9f;
      last_is_best = true;
    } else if (this_score < second_best_score) {
      second_best_score = this_score;
      second_best_step = (float)step;
    }
    that_patch[0] = that_patch[1];
    that_patch[1] = that_patch[2];
    that_patch[2] = that_patch[3];
    that_patch[3] = that_patch[4];
  }
  if (best_score > 4.0f * kMaxErrorStereo) return -3.0f;
  if (fabsf(second_best_step - best_step) > 1.0f && kMinDistanceErrorStereo * best_score > second_best_score)
    return -2.0f;
  float best_match_sub = best_step;
  bool did_subpixel = false;
  {
    const float gradPre_pre = -(pre_best_score - best_pre);
    const float gradPre_this = +(best_score - best_pre);
    const float gradPost_this = -(best_score - best_post);
    const float gradPost_post = +(post_best_score - best_post);
    bool interp_post = false, interp_pre = false;
    if ((gradPre_pre < 0) ^ (gradPre_this < 0)) {
      if (!((gradPost_post < 0) ^ (gradPost_this < 0))) interp_pre = true;
    } else if ((gradPost_post < 0) ^ (gradPost_this < 0)) {
      interp_post = true;
    }
    if (interp_pre) {
      const float d = gradPre_this / (gradPre_this - gradPre_pre);
      best_match_sub -= d;
      best_score = (best_score - (2.0f * d) * gradPre_this) - ((gradPre_pre - gradPre_this) * d) * d;
      did_subpixel = true;
    } else if (interp_post) {
      const float d = gradPost_this / (gradPost_this - gradPost_post);
      best_match_sub -= d;
      best_score = (best_score + (2.0f * d) * gradPost_this) + ((gradPost_post - gradPost_this) * d) * d;
      did_subpixel = true;
    }
  }
  float grad_along = 0.0f;
#pragma unroll
  for (int i = 1; i < 5; ++i) grad_along += (this_patch[i] - this_patch[i - 1]) * (this_patch[i] - this_patch[i - 1]);
  if (best_score > kMaxErrorStereo + sqrtf(grad_along) * 20.0f) return -3.0f;
  float best_idepth, alpha;
  if (sdx * sdx > sdy * sdy) {
    const float best_u = far_u + best_match_sub * sdx;
    best_idepth = (best_u * KRKiP.z - KRKiP.x) / (KT.x - best_u * KT.z);
    alpha = (KRKiP.z * KT.x - KT.z * KRKiP.x) / ((KT.x - best_u * KT.z) * (KT.x - best_u * KT.z));
  } else {
    const float best_v = far_v + best_match_sub * sdy;
    best_idepth = (best_v * KRKiP.z - KRKiP.y) / (KT.y - best_v * KT.z);
    alpha = (KRKiP.z * KT.y - KT.z * KRKiP.y) / ((KT.y - best_v * KT.z) * (KT.y - best_v * KT.z));
  }
  best_idepth = (double)best_idepth < kMinInvDepth ? (float)kMinInvDepth : best_idepth;
  best_idepth = (double)best_idepth > kMaxInvDepth ? (float)kMaxInvDepth : best_idepth;
  const float photo = 4.0f * 4.0f / my_gradient / my_gradient;
  const float geo = 0.25f * gradient_max * gradient_max / (my_gradient * my_gradient);
  result_var = (float)((double)(alpha * alpha) * (((did_subpixel ? 0.05 : 0.5) + (double)geo) + (double)photo));
  result_idep = best_idepth;
  return best_score;
}

}  // namespace sd

using namespace sd;

__global__ __launch_bounds__(256) void k_sd_gradient(const uint8_t* __restrict__ img, int W, int H, int pitch,
                                                     float2* __restrict__ grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  float2 g = make_float2(0.0f, 0.0f);
  if (x > 0 && y > 0 && x < W - 1 && y < H - 1) {
    const float r = img[(size_t)y * pitch + x + 1], l = img[(size_t)y * pitch + x - 1];
    const float d = img[(size_t)(y + 1) * pitch + x], u = img[(size_t)(y - 1) * pitch + x];
    g = make_float2((float)((double)(r - l) / 2.0), (float)((double)(d - u) / 2.0));
  }
  grad[i] = g;
}

// initial keyframe: the default seed on every interior pixel with |gradient| >= 5
__global__ __launch_bounds__(256) void k_sd_init(const float2* __restrict__ grad, int W, int H,
                                                 float4* __restrict__ seeds) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  float4 a = make_float4(0, 0, 0, 0), b = make_float4(0, 0, 0, 0);
  const float2 g = grad[i];
  if (interior(x, y, W, H) && !(g.x * g.x + g.y * g.y < kMinGrad * kMinGrad)) {
    a = make_float4(1.0f, 0.0f, 20.0f, 0.0f);
    b = make_float4(1.0f, (float)kVarianceMax, 1.0f, (float)kVarianceMax);
  }
  seeds[2 * i] = a;
  seeds[2 * i + 1] = b;
}

// src/lsd_semidense.cu:389-527: one keyframe pixel against the income frame, in place
__global__ __launch_bounds__(64) void k_sd_update(Params P, float4* __restrict__ seeds) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int W = P.W, H = P.H;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  if (!interior(x, y, W, H)) return;
  const float2 g = P.key_grad[i];
  const float gradient_max = sqrtf(g.x * g.x + g.y * g.y);
  if (gradient_max < kMinGrad) return;
  float4 t = seeds[2 * i], d = seeds[2 * i + 1];
  if (t.y < kMinBlacklist) return;
  const bool first_observe = !(t.x > 0.0f);
  const float tx = P.i2k[3], ty = P.i2k[7], tz = P.i2k[11];
  float epx = (-P.fx) * tx + tz * ((float)x - P.cx);
  float epy = (-P.fy) * ty + tz * ((float)y - P.cy);
  const float ep_len = sqrtf(epx * epx + epy * epy);
  if ((double)ep_len < 1.0) return;
  epx = epx / ep_len;
  epy = epy / ep_len;
  const float grad_along = fabsf(epx * g.x + epy * g.y);
  if ((double)grad_along < 2.0) return;
  if ((double)(grad_along / gradient_max) < 0.3) return;
  float result_idep = 0.0f, result_var = 0.0f;
  const float err = search_point(x, y, P, epx, epy, gradient_max, grad_along, result_idep, result_var);
  const float deff = result_idep - d.z;
  if (first_observe) {
    if (err == -3.0f || err == -2.0f) t.y -= 1.0f;
    if (err > 0.0f && (double)result_var < kVarianceMax) {
      t = make_float4(1.0f, 0.0f, kValidityInitial, 0.0f);
      d = make_float4(result_idep, result_var, 1.0f, (float)kVarianceMax);
    }
  } else {
    if (err == -2.0f) {
      t.z -= kValidityDec;
      t.z = t.z < 0.0f ? 0.0f : t.z;
      d.y *= kFailVarIncFac;
      if ((double)d.y > kVarianceMax) {
        t.x = -1.0f;
        t.y -= 1.0f;
      }
    } else if (err == -1.0f || err == -3.0f || err == -4.0f) {
      return;
    } else if (deff * deff > result_var + d.w) {
      d.y *= kFailVarIncFac;
      if ((double)d.y > kVarianceMax) t.x = -1.0f;
    } else {
      const float inc_var = d.y * kSuccVarIncFac;
      const float new_var = (inc_var * result_var) / (inc_var + result_var);
      const float new_idepth = (d.x * result_var + result_idep * inc_var) / (inc_var + result_var);
      if (new_var < d.y) d.y = new_var;
      d.x = new_idepth;
      t.z += kValidityInc;
      const float max_valid = kValidityMax + gradient_max * kValidityMaxVariable / 255.0f;
      t.z = t.z > max_valid ? max_valid : t.z;
    }
  }
  seeds[2 * i] = t;
  seeds[2 * i + 1] = d;
}

// src/lsd_semidense.cu:345-387 from a snapshot; the gradient test reads the income frame
__global__ __launch_bounds__(256) void k_sd_fill(const float4* __restrict__ in, float4* __restrict__ out, int W,
                                                 int H, const float2* __restrict__ inc_grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  float4 t = in[2 * i], d = in[2 * i + 1];
  if (interior(x, y, W, H) && !(t.x > 0.0f)) {
    const float2 g = inc_grad[i];
    if (!(g.x * g.x + g.y * g.y < kMinGrad * kMinGrad)) {
      float sum_idepth = 0.0f, sum_ivar = 0.0f, sum_valid = 0.0f;
      for (int j = -2; j <= 2; ++j)
        for (int k = -2; k <= 2; ++k) {
          const size_t o = (size_t)(y + j) * W + (x + k);
          const float4 ot = in[2 * o];
          if (ot.x > 0.0f) {
            const float4 od = in[2 * o + 1];
            sum_valid += ot.z;
            const float w = 1.0f / od.y;
            sum_idepth += od.x * w;
            sum_ivar += w;
          }
        }
      const bool create = (sum_valid > kValSumMinUnblacklist) || (t.y >= kMinBlacklist && sum_valid > kValSumMinCreate);
      if (create) {
        t = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
        d = make_float4(sum_idepth / sum_ivar, 1.0f / sum_ivar, 1.0f, (float)kVarianceMax);
      }
    }
  }
  out[2 * i] = t;
  out[2 * i + 1] = d;
}

// src/lsd_semidense.cu:296-344 from a snapshot
__global__ __launch_bounds__(256) void k_sd_regularize(const float4* __restrict__ in, float4* __restrict__ out, int W,
                                                       int H, int remove_occlusions) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  float4 t = in[2 * i], d = in[2 * i + 1];
  if (interior(x, y, W, H) && t.x > 0.0f) {
    float sum_idepth = 0.0f, sum_ivar = 0.0f, sum_valid = 0.0f;
    int occlusion = 0, not_occlusion = 0;
    for (int j = -2; j <= 2; ++j)
      for (int k = -2; k <= 2; ++k) {
        const size_t o = (size_t)(y + j) * W + (x + k);
        const float4 ot = in[2 * o];
        if (!(ot.x > 0.0f)) continue;
        const float4 od = in[2 * o + 1];
        const float diff = d.x - od.x;
        if (diff * diff > d.y + od.y) {
          if (d.x < od.x) occlusion++;
          continue;
        }
        sum_valid += ot.z;
        not_occlusion++;
        const float spatial = kRegDistVar * (float)(k * k + j * j);
        const float ivar = 1.0f / (od.y + spatial);
        sum_idepth += ivar * od.x;
        sum_ivar += ivar;
      }
    if (sum_valid < kValSumMinKeep) t.x = -1.0f;
    if (remove_occlusions && occlusion > not_occlusion) t.x = -1.0f;
    d.z = sum_idepth / sum_ivar;
    d.w = 1.0f / sum_ivar;
  }
  out[2 * i] = t;
  out[2 * i + 1] = d;
}

// projection of a keyframe seed into the income frame (shared by the sparse
// map projection and the propagation)
__device__ __forceinline__ bool project_seed(const Params& P, const float* T, int x, int y, float s_idepth,
                                             float3& np, float& new_idepth, int& nx, int& ny) {
  const int W = P.W, H = P.H;
  float3 kp = make_float3(((float)x - P.cx) / P.fx, ((float)y - P.cy) / P.fy, 1.0f);
  kp.x = kp.x / s_idepth;
  kp.y = kp.y / s_idepth;
  kp.z = kp.z / s_idepth;
  np = transform(T, kp);
  new_idepth = 1.0f / np.z;
  nx = trunc_i32((double)((np.x * new_idepth) * P.fx + P.cx) + 0.5);
  ny = trunc_i32((double)((np.y * new_idepth) * P.fy + P.cy) + 0.5);
  if (nx <= 2 || nx >= W - 3 || ny <= 2 || ny >= H - 3) return false;
  const float2 g = P.inc_grad[(size_t)ny * W + nx];
  const float g2 = g.x * g.x + g.y * g.y;
  if (g2 < kMinGrad * kMinGrad) return false;
  const float diff = texel(P.key, W, H, P.pitch, x, y) - texel(P.inc, W, H, P.pitch, nx, ny);
  if (diff * diff > (kMaxDiffConst + kMaxDiffGradMult * g2)) return false;
  return true;
}

__global__ __launch_bounds__(256) void k_sd_fill_u64(unsigned long long* __restrict__ p, size_t n,
                                                     unsigned long long v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// src/lsd_semidense.cu:181-239: nearest depth per target pixel
__global__ __launch_bounds__(256) void k_sd_project(Params P, const float4* __restrict__ seeds,
                                                    unsigned long long* __restrict__ best) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int W = P.W, H = P.H;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  if (!interior(x, y, W, H)) return;
  const float4 t = seeds[2 * i];
  if (!(t.x > 0.0f) || t.z < 20.0f) return;
  const float4 d = seeds[2 * i + 1];
  float3 np;
  float nid;
  int nx, ny;
  if (!project_seed(P, P.k2i, x, y, d.z, np, nid, nx, ny)) return;
  const float nd = sqrtf((np.x * np.x + np.y * np.y) + np.z * np.z);
  if (nd != nd || (double)d.w > 0.02) return;
  const unsigned long long k = ((unsigned long long)__float_as_uint(nd) << 32) | (unsigned long long)(uint32_t)i;
  atomicMin(&best[(size_t)ny * W + nx], k);
}

__global__ __launch_bounds__(256) void k_sd_resolve(const unsigned long long* __restrict__ best, float* __restrict__ out,
                                                    int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = best[i];
  out[i] = k == ~0ull ? 0.0f : __uint_as_float((uint32_t)(k >> 32));
}

// src/lsd_semidense.cu:241-294: the (inverse depth, variance) prior of the
// income frame, read by the prior-driven alternates. The reference returns when
// the projected depth and variance are both finite (:290-291), so it never
// stores a finite prior and its prior map stays zero; here the test is the one
// that code intends: finite values are kept. Nearest depth wins a pixel, as
// for the depth projection.
__global__ __launch_bounds__(256) void k_sd_project_prior(Params P, const float4* __restrict__ seeds,
                                                          unsigned long long* __restrict__ best,
                                                          float4* __restrict__ info) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int W = P.W, H = P.H;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  if (!interior(x, y, W, H)) return;
  const float4 t = seeds[2 * i];
  if (!(t.x > 0.0f) || t.z < 10.0f) return;
  const float4 d = seeds[2 * i + 1];
  float3 np;
  float nid;
  int nx, ny;
  if (!project_seed(P, P.k2i, x, y, d.z, np, nid, nx, ny)) return;
  const float nd = sqrtf((np.x * np.x + np.y * np.y) + np.z * np.z);
  const float vs = nd / np.z;
  float var = (vs * vs) * d.w;
  if (!(isfinite(nd) && isfinite(var))) return;
  var = (double)var < 0.01 ? (float)0.01 : var;
  info[i] = make_float4(1.0f / nd, var, 0.0f, 0.0f);
  const unsigned long long k = ((unsigned long long)__float_as_uint(nd) << 32) | (unsigned long long)(uint32_t)i;
  atomicMin(&best[(size_t)ny * W + nx], k);
}

__global__ __launch_bounds__(256) void k_sd_resolve_prior(const unsigned long long* __restrict__ best,
                                                          const float4* __restrict__ info, float2* __restrict__ out,
                                                          int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = best[i];
  float2 v = make_float2(0.0f, 0.0f);
  if (k != ~0ull) {
    const float4 in = info[(uint32_t)(k & 0xFFFFFFFFull)];
    v = make_float2(in.x, in.y);
  }
  out[i] = v;
}

__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// src/lsd_semidense.cu:95-160: propagation of the old keyframe into the income frame
__global__ __launch_bounds__(256) void k_sd_propagate(Params P, const float4* __restrict__ seeds,
                                                      unsigned long long* __restrict__ table,
                                                      float4* __restrict__ info) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int W = P.W, H = P.H;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  if (!interior(x, y, W, H)) return;
  const float4 t = seeds[2 * i];
  if (!(t.x > 0.0f)) return;
  const float4 d = seeds[2 * i + 1];
  float3 np;
  float nid;
  int nx, ny;
  if (!project_seed(P, P.k2i, x, y, d.z, np, nid, nx, ny)) return;
  float r4 = nid / d.z;
  r4 *= r4;
  r4 *= r4;
  const float nvar = r4 * d.w;
  info[i] = make_float4(nid, nvar, t.z, 0.0f);
  const unsigned long long k =
      ((unsigned long long)fkey(nid) << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)i);
  atomicMax(&table[(size_t)ny * W + nx], k);
}

__global__ __launch_bounds__(256) void k_sd_init_table(const unsigned long long* __restrict__ table,
                                                       const float4* __restrict__ info, int W, int H,
                                                       float4* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int y = i / W, x = i - y * W;
  float4 t = make_float4(0, 0, 0, 0), d = make_float4(0, 0, 0, 0);
  const unsigned long long k = table[i];
  if (interior(x, y, W, H) && k != 0) {
    const uint32_t src = 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFull);
    if (src != 0) {
      const float4 in = info[src];
      t = make_float4(1.0f, 0.0f, in.z, 0.0f);
      d = make_float4(in.x, in.y, 1.0f, (float)kVarianceMax);
    }
  }
  out[2 * i] = t;
  out[2 * i + 1] = d;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int blocks(size_t n, int per) { return (int)((n + per - 1) / per); }

static void se3_inv_host(const float* d, float* r) {
  r[0] = d[0]; r[1] = d[4]; r[2] = d[8];
  r[4] = d[1]; r[5] = d[5]; r[6] = d[9];
  r[8] = d[2]; r[9] = d[6]; r[10] = d[10];
  r[3] = ((-d[0]) * d[3] - d[4] * d[7]) - d[8] * d[11];
  r[7] = ((-d[1]) * d[3] - d[5] * d[7]) - d[9] * d[11];
  r[11] = ((-d[2]) * d[3] - d[6] * d[7]) - d[10] * d[11];
}

void sd_gradient(const uint8_t* img, int W, int H, int pitch, float2* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_sd_gradient, dim3(blocks((size_t)W * H, 256)), dim3(256), 0, s, img, W, H, pitch, grad);
}

void sd_init(const float2* grad, int W, int H, float4* seeds, hipStream_t s) {
  hipLaunchKernelGGL(k_sd_init, dim3(blocks((size_t)W * H, 256)), dim3(256), 0, s, grad, W, H, seeds);
}

// One frame of the semi-dense module (src/seed_matrix.cu:399-446, then
// :247-304 when switch_keyframe): seeds[0] holds the state on entry and on
// exit; seeds[1] is the second snapshot buffer. Scratch: best / table
// (W x H uint64) and info (W x H float4).
void sd_step(float4* seeds[2], const uint8_t* key, const float2* key_grad, const uint8_t* inc,
             const float2* inc_grad, int W, int H, int pitch, const float cam[4], const float T_key_world[12],
             const float T_inc_world[12], bool switch_keyframe, float* prefill, float2* prior,
             unsigned long long* u64, float4* info, hipStream_t s) {
  Params P;
  P.W = W;
  P.H = H;
  P.pitch = pitch;
  P.fx = cam[0];
  P.fy = cam[1];
  P.cx = cam[2];
  P.cy = cam[3];
  float key_inv[12];
  se3_inv_host(T_key_world, key_inv);
  const float* L = T_inc_world;
  for (int r = 0; r < 3; ++r) {
    const int o = 4 * r;
    for (int c = 0; c < 3; ++c) P.k2i[o + c] = (L[o] * key_inv[c] + L[o + 1] * key_inv[4 + c]) + L[o + 2] * key_inv[8 + c];
    P.k2i[o + 3] = ((L[o + 3] + L[o] * key_inv[3]) + L[o + 1] * key_inv[7]) + L[o + 2] * key_inv[11];
  }
  se3_inv_host(P.k2i, P.i2k);
  P.key = key;
  P.key_grad = key_grad;
  P.inc = inc;
  P.inc_grad = inc_grad;
  const size_t n = (size_t)W * H;
  float4* a = seeds[0];
  float4* b = seeds[1];
  hipLaunchKernelGGL(k_sd_update, dim3(blocks(n, 64)), dim3(64), 0, s, P, a);
  hipLaunchKernelGGL(k_sd_fill, dim3(blocks(n, 256)), dim3(256), 0, s, a, b, W, H, inc_grad);
  hipLaunchKernelGGL(k_sd_regularize, dim3(blocks(n, 256)), dim3(256), 0, s, b, a, W, H, 0);
  hipLaunchKernelGGL(k_sd_fill_u64, dim3(blocks(n, 256)), dim3(256), 0, s, u64, n, ~0ull);
  hipLaunchKernelGGL(k_sd_project, dim3(blocks(n, 256)), dim3(256), 0, s, P, a, u64);
  hipLaunchKernelGGL(k_sd_resolve, dim3(blocks(n, 256)), dim3(256), 0, s, u64, prefill, (int)n);
  if (prior) {
    hipLaunchKernelGGL(k_sd_fill_u64, dim3(blocks(n, 256)), dim3(256), 0, s, u64, n, ~0ull);
    hipLaunchKernelGGL(k_sd_project_prior, dim3(blocks(n, 256)), dim3(256), 0, s, P, a, u64, info);
    hipLaunchKernelGGL(k_sd_resolve_prior, dim3(blocks(n, 256)), dim3(256), 0, s, u64, info, prior, (int)n);
  }
  if (switch_keyframe) {
    hipLaunchKernelGGL(k_sd_fill_u64, dim3(blocks(n, 256)), dim3(256), 0, s, u64, n, 0ull);
    hipLaunchKernelGGL(k_sd_propagate, dim3(blocks(n, 256)), dim3(256), 0, s, P, a, u64, info);
    hipLaunchKernelGGL(k_sd_init_table, dim3(blocks(n, 256)), dim3(256), 0, s, u64, info, W, H, b);
    hipLaunchKernelGGL(k_sd_regularize, dim3(blocks(n, 256)), dim3(256), 0, s, b, a, W, H, 1);
    hipLaunchKernelGGL(k_sd_fill, dim3(blocks(n, 256)), dim3(256), 0, s, a, b, W, H, inc_grad);
    hipLaunchKernelGGL(k_sd_regularize, dim3(blocks(n, 256)), dim3(256), 0, s, b, a, W, H, 0);
  }
}

}  // namespace qm
