// qm_semidense.cpp -- TEST INFRASTRUCTURE ONLY. Scalar restatement of the
// reference's semi-dense (LSD-style) keyframe seeds and the image gradient that
// feeds them, used to check the HIP kernels (qm_ctx / k_semidense.hip). Never
// linked into or called by the product path.
//
// Follows src/seed_init.cu:26-42 (gradient), src/lsd_semidense.cu:64-239,296-753
// (initialise, propagate, project, regularise, fill holes, update with the
// epipolar search) and the drivers src/seed_matrix.cu:211-304,399-446.
//
// Pinned here (and in the kernels):
// * arithmetic: FP32 without contraction, IEEE division and sqrt; the
//   reference's double literals keep their double arithmetic ("+ 0.5",
//   "2.0 *", "search_length += 4.0", the result-variance factor);
// * texture fetches: texel centres at +0.5, clamp addressing; a fetch at a
//   texel centre is exact; a bilinear fetch uses weights quantised to 1/256
//   (CUDA's 8-bit-fraction filtering weights, frac rounded to nearest) and,
//   the images being integer grey levels, is computed exactly in integers;
// * races of the reference resolved deterministically: the regularisation and
//   hole filling read a snapshot of the seeds (every read before any write, a
//   legal schedule of the in-place kernels); the propagation table keeps the
//   largest inverse depth, ties to the lower source index; the projection
//   into the sparse map keeps the nearest depth, ties to the lower source.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr float kMinGrad = 5.0f;          // MIN_GRAIDIENT
constexpr double kMinInvDepth = 0.01;     // MIN_INV_DEPTH
constexpr double kMaxInvDepth = 2.0;      // MAX_INV_DEPTH
constexpr double kIdepthInitial = 1.0;    // IDEPTH_INITIAL
constexpr double kVarianceMax = 4.0;      // VARIANCE_MAX
constexpr float kMaxDiffConst = 40.0f * 40.0f;
constexpr float kMaxDiffGradMult = 0.5f * 0.5f;
constexpr float kRegDistVar = 0.075f * 0.075f;
constexpr int kValSumMinCreate = 24;
constexpr int kValSumMinKeep = 24;
constexpr int kValSumMinUnblacklist = 100;
constexpr int kMinBlacklist = -1;
constexpr int kMaxSearch = 100;
constexpr float kMaxErrorStereo = 1300.0f;
constexpr float kMinDistanceErrorStereo = 1.5f;
constexpr int kValidityInc = 5, kValidityDec = 5, kValidityInitial = 5;
constexpr float kValidityMax = 5.0f, kValidityMaxVariable = 250.0f;
constexpr float kSuccVarIncFac = 1.01f, kFailVarIncFac = 1.1f;

// seed: [valid, blacklist, counter, 0, idepth, variance, smooth idepth, smooth variance]
struct Seed {
  float v[8];
  bool valid() const { return v[0] > 0.0f; }
  void init_default() {
    v[0] = 1.0f; v[1] = 0.0f; v[2] = 20.0f; v[3] = 0.0f;
    v[4] = (float)kIdepthInitial; v[5] = (float)kVarianceMax; v[6] = (float)kIdepthInitial; v[7] = (float)kVarianceMax;
  }
  void init(float idepth, float var, float counter) {
    v[0] = 1.0f; v[1] = 0.0f; v[2] = counter; v[3] = 0.0f;
    v[4] = idepth; v[5] = var; v[6] = (float)kIdepthInitial; v[7] = (float)kVarianceMax;
  }
};

struct Cam {
  float fx, fy, cx, cy;
};

struct F3 {
  float x, y, z;
};

// SE3 (3x4 row-major) times a point: (R p) + t, rows summed left to right
F3 se3_mul_point(const float* T, F3 p) {
  F3 r;
  r.x = ((T[0] * p.x + T[1] * p.y) + T[2] * p.z) + T[3];
  r.y = ((T[4] * p.x + T[5] * p.y) + T[6] * p.z) + T[7];
  r.z = ((T[8] * p.x + T[9] * p.y) + T[10] * p.z) + T[11];
  return r;
}
F3 se3_rotate(const float* T, F3 p) {
  F3 r;
  r.x = (T[0] * p.x + T[1] * p.y) + T[2] * p.z;
  r.y = (T[4] * p.x + T[5] * p.y) + T[6] * p.z;
  r.z = (T[8] * p.x + T[9] * p.y) + T[10] * p.z;
  return r;
}
void se3_inv(const float* d, float* r) {
  r[0] = d[0]; r[1] = d[4]; r[2] = d[8];
  r[4] = d[1]; r[5] = d[5]; r[6] = d[9];
  r[8] = d[2]; r[9] = d[6]; r[10] = d[10];
  r[3] = ((-d[0]) * d[3] - d[4] * d[7]) - d[8] * d[11];
  r[7] = ((-d[1]) * d[3] - d[5] * d[7]) - d[9] * d[11];
  r[11] = ((-d[2]) * d[3] - d[6] * d[7]) - d[10] * d[11];
}

int trunc_i32(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (-2147483647 - 1);
  return (int)v;
}

// texel at an integer + 0.5 coordinate
inline float texel(const uint8_t* img, int W, int H, int pitch, int x, int y) {
  x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
  y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
  return (float)img[(size_t)y * pitch + x];
}

// bilinear fetch at (u, v) in texel-centre coordinates
float tex_bilinear(const uint8_t* img, int W, int H, int pitch, float u, float v) {
  float xb = u - 0.5f, yb = v - 0.5f;
  xb = std::fmin(std::fmax(xb, -2.0f), (float)W + 1.0f);  // far outside: both taps clamp alike; NaN -> -2
  yb = std::fmin(std::fmax(yb, -2.0f), (float)H + 1.0f);
  const float fx0 = std::floor(xb), fy0 = std::floor(yb);
  const int a = (int)std::rint((xb - fx0) * 256.0f), b = (int)std::rint((yb - fy0) * 256.0f);
  const int i0 = (int)fx0, j0 = (int)fy0;
  auto T = [&](int i, int j) -> int {
    i = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
    j = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
    return img[(size_t)j * pitch + i];
  };
  const int s = (256 - a) * (256 - b) * T(i0, j0) + a * (256 - b) * T(i0 + 1, j0) + (256 - a) * b * T(i0, j0 + 1) +
                a * b * T(i0 + 1, j0 + 1);
  return (float)s * (1.0f / 65536.0f);
}

inline float dot2(float ax, float ay, float bx, float by) { return ax * bx + ay * by; }

struct Frame {
  const uint8_t* img;
  const float* grad;  // W x H x 2
  int W, H, pitch;
  float gx(int x, int y) const { return grad[2 * ((size_t)y * W + x)]; }
  float gy(int x, int y) const { return grad[2 * ((size_t)y * W + x) + 1]; }
  float g2(int x, int y) const { return dot2(gx(x, y), gy(x, y), gx(x, y), gy(x, y)); }
};

bool interior(int x, int y, int W, int H) { return !(x >= W - 3 || y >= H - 3 || x <= 2 || y <= 2); }

// src/lsd_semidense.cu:529-753
float search_point(int x, int y, int W, int H, float epx, float epy, float gradient_max, float my_gradient,
                   float min_idep, float max_idep, const Cam& c, const float* k2i, const Frame& key,
                   const Frame& inc, float& result_idep, float& result_var, float& result_eplength) {
  float this_patch[5];
  for (int i = -2; i <= 2; ++i)
    this_patch[i + 2] = tex_bilinear(key.img, W, H, key.pitch, (float)((double)((float)x + (float)i * epx) + 0.5),
                                     (float)((double)((float)y + (float)i * epy) + 0.5));
  const F3 KiP = {((float)x - c.cx) / c.fx, ((float)y - c.cy) / c.fy, 1.0f};
  const F3 RKiP = se3_rotate(k2i, KiP);
  const F3 T = {k2i[3], k2i[7], k2i[11]};
  const F3 KRKiP = {c.fx * RKiP.x + c.cx * RKiP.z, c.fy * RKiP.y + c.cy * RKiP.z, RKiP.z};
  const F3 KT = {c.fx * T.x + c.cx * T.z, c.fy * T.y + c.cy * T.z, T.z};
  const F3 near_p = {KRKiP.x + KT.x * max_idep, KRKiP.y + KT.y * max_idep, KRKiP.z + KT.z * max_idep};
  const F3 far_p = {KRKiP.x + KT.x * min_idep, KRKiP.y + KT.y * min_idep, KRKiP.z + KT.z * min_idep};
  float near_u = near_p.x / near_p.z, near_v = near_p.y / near_p.z;
  float far_u = far_p.x / far_p.z, far_v = far_p.y / far_p.z;
  const float du = near_u - far_u, dv = near_v - far_v;
  float search_length = std::sqrt(dot2(du, dv, du, dv));
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
  for (int i = -2; i <= 1; ++i)
    that_patch[i + 2] = tex_bilinear(inc.img, W, H, inc.pitch, (float)((double)(far_u + (float)i * sdx) + 0.5),
                                     (float)((double)(far_v + (float)i * sdy) + 0.5));
  float best_step = -1.0f, best_score = 1e9f, pre_best_score = 1e9f, post_best_score = 1e9f;
  bool last_is_best = false;
  float best_pre = NAN, best_post = NAN;
  float e1A = NAN, e1B = NAN, e2A = NAN, e2B = NAN, e3A = NAN, e3B = NAN, e4A = NAN, e4B = NAN, e5A = NAN, e5B = NAN;
  float second_best_score = 1e9f, second_best_step = -1.0f;
  float pre_score = 1e9f, this_score = 1e9f;
  int step = 0;
  for (; (float)step <= search_length && step <= kMaxSearch; step += 1) {
    pre_score = this_score;
    this_score = 0.0f;
    const float spu = far_u + (float)step * sdx, spv = far_v + (float)step * sdy;
    if (spu <= 1.0f || spu >= (float)(W - 1) || spv <= 1.0f || spv >= (float)(H - 1)) break;
    that_patch[4] = tex_bilinear(inc.img, W, H, inc.pitch, (float)(((double)spu + 2.0 * (double)sdx) + 0.5),
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
  if (std::fabs(second_best_step - best_step) > 1.0f && kMinDistanceErrorStereo * best_score > second_best_score)
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
  for (int i = 1; i < 5; ++i) grad_along += (this_patch[i] - this_patch[i - 1]) * (this_patch[i] - this_patch[i - 1]);
  if (best_score > kMaxErrorStereo + std::sqrt(grad_along) * 20.0f) return -3.0f;
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
  result_eplength = (float)step;
  return best_score;
}

// src/lsd_semidense.cu:427-527, one keyframe pixel
void update_pixel(Seed& s, int x, int y, int W, int H, const Cam& c, const float* k2i, const float* i2k,
                  const Frame& key, const Frame& inc) {
  const float gxx = key.gx(x, y), gyy = key.gy(x, y);
  const float gradient_max = std::sqrt(dot2(gxx, gyy, gxx, gyy));
  if (gradient_max < kMinGrad) return;
  if (s.v[1] < (float)kMinBlacklist) return;
  const bool first_observe = !s.valid();
  const F3 t = {i2k[3], i2k[7], i2k[11]};
  float epx = (-c.fx) * t.x + t.z * ((float)x - c.cx);
  float epy = (-c.fy) * t.y + t.z * ((float)y - c.cy);
  const float ep_len = std::sqrt(dot2(epx, epy, epx, epy));
  if ((double)ep_len < 1.0) return;
  epx = epx / ep_len;
  epy = epy / ep_len;
  const float grad_along = std::fabs(dot2(epx, epy, gxx, gyy));
  if ((double)grad_along < 2.0) return;
  if ((double)(grad_along / gradient_max) < 0.3) return;
  float result_idep = 0.0f, result_var = 0.0f, result_eplength = 0.0f;
  // the refined bounds of a seen seed are shadowed in the reference: the full range is searched
  const float err = search_point(x, y, W, H, epx, epy, gradient_max, grad_along, (float)kMinInvDepth,
                                 (float)kMaxInvDepth, c, k2i, key, inc, result_idep, result_var, result_eplength);
  const float deff = result_idep - s.v[6];
  if (first_observe) {
    if (err == -3.0f || err == -2.0f) s.v[1] -= 1.0f;
    if (err > 0.0f && (double)result_var < kVarianceMax) s.init(result_idep, result_var, (float)kValidityInitial);
  } else {
    if (err == -2.0f) {
      s.v[2] -= (float)kValidityDec;
      s.v[2] = s.v[2] < 0.0f ? 0.0f : s.v[2];
      s.v[5] *= kFailVarIncFac;
      if ((double)s.v[5] > kVarianceMax) {
        s.v[0] = -1.0f;
        s.v[1] -= 1.0f;
      }
    } else if (err == -1.0f || err == -3.0f || err == -4.0f) {
      return;
    } else if (deff * deff > result_var + s.v[7]) {
      s.v[5] *= kFailVarIncFac;
      if ((double)s.v[5] > kVarianceMax) s.v[0] = -1.0f;
    } else {
      const float inc_var = s.v[5] * kSuccVarIncFac;
      const float new_var = (inc_var * result_var) / (inc_var + result_var);
      const float new_idepth = (s.v[4] * result_var + result_idep * inc_var) / (inc_var + result_var);
      if (new_var < s.v[5]) s.v[5] = new_var;
      s.v[4] = new_idepth;
      s.v[2] += (float)kValidityInc;
      const float max_valid = kValidityMax + gradient_max * kValidityMaxVariable / 255.0f;
      s.v[2] = s.v[2] > max_valid ? max_valid : s.v[2];
    }
  }
}

// src/lsd_semidense.cu:296-344 on a snapshot
void regularize(std::vector<Seed>& seeds, int W, int H, bool remove_occlusions) {
  const std::vector<Seed> snap = seeds;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      Seed s = snap[(size_t)y * W + x];
      if (!s.valid()) continue;
      float sum_idepth = 0.0f, sum_ivar = 0.0f, sum_valid = 0.0f;
      int occlusion = 0, not_occlusion = 0;
      for (int j = -2; j <= 2; ++j)
        for (int i = -2; i <= 2; ++i) {
          const Seed& o = snap[(size_t)(y + j) * W + (x + i)];
          if (!o.valid()) continue;
          const float diff = s.v[4] - o.v[4];
          if (diff * diff > s.v[5] + o.v[5]) {
            if (s.v[4] < o.v[4]) occlusion++;
            continue;
          }
          sum_valid += o.v[2];
          not_occlusion++;
          const float spatial = kRegDistVar * (float)(i * i + j * j);
          const float ivar = 1.0f / (o.v[5] + spatial);
          sum_idepth += ivar * o.v[4];
          sum_ivar += ivar;
        }
      if (sum_valid < (float)kValSumMinKeep) s.v[0] = -1.0f;
      if (remove_occlusions && occlusion > not_occlusion) s.v[0] = -1.0f;
      s.v[6] = sum_idepth / sum_ivar;
      s.v[7] = 1.0f / sum_ivar;
      seeds[(size_t)y * W + x] = s;
    }
}

// src/lsd_semidense.cu:345-387 on a snapshot; the gradient test reads the income frame
void fill_holes(std::vector<Seed>& seeds, int W, int H, const Frame& inc) {
  const std::vector<Seed> snap = seeds;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      if (inc.g2(x, y) < kMinGrad * kMinGrad) continue;
      Seed s = snap[(size_t)y * W + x];
      if (s.valid()) continue;
      float sum_idepth = 0.0f, sum_ivar = 0.0f, sum_valid = 0.0f;
      for (int j = -2; j <= 2; ++j)
        for (int i = -2; i <= 2; ++i) {
          const Seed& o = snap[(size_t)(y + j) * W + (x + i)];
          if (o.valid()) {
            sum_valid += o.v[2];
            const float w = 1.0f / o.v[5];
            sum_idepth += o.v[4] * w;
            sum_ivar += w;
          }
        }
      const bool create = (sum_valid > (float)kValSumMinUnblacklist) ||
                          (s.v[1] >= (float)kMinBlacklist && sum_valid > (float)kValSumMinCreate);
      if (create) {
        s.init(sum_idepth / sum_ivar, 1.0f / sum_ivar, 0.0f);
        seeds[(size_t)y * W + x] = s;
      }
    }
}

// projection of a keyframe seed into the income frame shared by the sparse-map
// projection and the propagation (src/lsd_semidense.cu:120-160,195-239)
bool project_seed(const Seed& s, int x, int y, int W, int H, const Cam& c, const float* T, const Frame& key,
                  const Frame& inc, F3& np, float& new_idepth, int& nx, int& ny) {
  F3 kp = {((float)x - c.cx) / c.fx, ((float)y - c.cy) / c.fy, 1.0f};
  kp.x = kp.x / s.v[6];
  kp.y = kp.y / s.v[6];
  kp.z = kp.z / s.v[6];
  np = se3_mul_point(T, kp);
  new_idepth = 1.0f / np.z;
  nx = trunc_i32((double)((np.x * new_idepth) * c.fx + c.cx) + 0.5);
  ny = trunc_i32((double)((np.y * new_idepth) * c.fy + c.cy) + 0.5);
  if (nx <= 2 || nx >= W - 3 || ny <= 2 || ny >= H - 3) return false;
  const float g2 = inc.g2(nx, ny);
  if (g2 < kMinGrad * kMinGrad) return false;
  const float diff = texel(key.img, W, H, key.pitch, x, y) - texel(inc.img, W, H, inc.pitch, nx, ny);
  if (diff * diff > (kMaxDiffConst + kMaxDiffGradMult * g2)) return false;
  return true;
}

// order-preserving unsigned key of a non-NaN float
uint32_t fkey(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// src/lsd_semidense.cu:181-239: seeds with validity >= 20 and smooth variance
// <= 0.02 write their depth into the sparse map; nearest wins
void project_depth(const std::vector<Seed>& seeds, int W, int H, const Cam& c, const float* k2i, const Frame& key,
                   const Frame& inc, float* out) {
  std::vector<uint64_t> best((size_t)W * H, ~0ull);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      const Seed& s = seeds[(size_t)y * W + x];
      if (!s.valid() || s.v[2] < 20.0f) continue;
      F3 np;
      float nid;
      int nx, ny;
      if (!project_seed(s, x, y, W, H, c, k2i, key, inc, np, nid, nx, ny)) continue;
      const float nd = std::sqrt((np.x * np.x + np.y * np.y) + np.z * np.z);
      if (nd != nd || (double)s.v[7] > 0.02) continue;
      uint32_t b;
      std::memcpy(&b, &nd, 4);
      const uint64_t k = ((uint64_t)b << 32) | (uint32_t)((size_t)y * W + x);
      uint64_t& t = best[(size_t)ny * W + nx];
      if (k < t) t = k;
    }
  for (size_t i = 0; i < (size_t)W * H; ++i) {
    if (best[i] == ~0ull) continue;
    const uint32_t b = (uint32_t)(best[i] >> 32);
    std::memcpy(&out[i], &b, 4);
  }
}

// src/lsd_semidense.cu:241-294: the (inverse depth, variance) prior of the
// income frame. Seeds with validity >= 10 project as above; the variance is
// scaled by (depth / z)^2 and held at >= 0.01. The reference returns when both
// values are finite (:290-291), which stores no finite prior at all; this
// restatement keeps the finite ones, as that code intends. Nearest depth wins,
// ties to the lower source index; pixels no seed reaches stay (0, 0).
void project_prior(const std::vector<Seed>& seeds, int W, int H, const Cam& c, const float* k2i, const Frame& key,
                   const Frame& inc, float* out /* H x W x 2 */) {
  std::vector<uint64_t> best((size_t)W * H, ~0ull);
  std::vector<float> val((size_t)W * H * 2, 0.0f);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      const size_t i = (size_t)y * W + x;
      const Seed& s = seeds[i];
      if (!s.valid() || s.v[2] < 10.0f) continue;
      F3 np;
      float nid;
      int nx, ny;
      if (!project_seed(s, x, y, W, H, c, k2i, key, inc, np, nid, nx, ny)) continue;
      const float nd = std::sqrt((np.x * np.x + np.y * np.y) + np.z * np.z);
      const float vs = nd / np.z;
      float var = (vs * vs) * s.v[7];
      if (!(std::isfinite(nd) && std::isfinite(var))) continue;
      if ((double)var < 0.01) var = (float)0.01;
      val[2 * i] = 1.0f / nd;
      val[2 * i + 1] = var;
      uint32_t b;
      std::memcpy(&b, &nd, 4);
      const uint64_t k = ((uint64_t)b << 32) | (uint32_t)i;
      uint64_t& t = best[(size_t)ny * W + nx];
      if (k < t) t = k;
    }
  for (size_t i = 0; i < (size_t)W * H; ++i) {
    out[2 * i] = 0.0f;
    out[2 * i + 1] = 0.0f;
    if (best[i] == ~0ull) continue;
    const uint32_t src = (uint32_t)(best[i] & 0xFFFFFFFFull);
    out[2 * i] = val[2 * (size_t)src];
    out[2 * i + 1] = val[2 * (size_t)src + 1];
  }
}

// src/lsd_semidense.cu:64-118 and src/seed_matrix.cu:247-304: propagate the old
// keyframe's seeds into the income frame, initialise the new keyframe from
// them, then regularise (removing occlusions), fill holes, regularise again
void new_keyframe(const std::vector<Seed>& old, int W, int H, const Cam& c, const float* o2n, const Frame& key,
                  const Frame& inc, std::vector<Seed>& out) {
  const size_t n = (size_t)W * H;
  std::vector<uint64_t> table(n, 0);
  std::vector<F3> info(n, F3{0.0f, 0.0f, 0.0f});
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      const Seed& s = old[(size_t)y * W + x];
      if (!s.valid()) continue;
      F3 np;
      float nid;
      int nx, ny;
      if (!project_seed(s, x, y, W, H, c, o2n, key, inc, np, nid, nx, ny)) continue;
      float r4 = nid / s.v[6];
      r4 *= r4;
      r4 *= r4;
      const float nvar = r4 * s.v[7];
      const size_t src = (size_t)y * W + x;
      info[src] = F3{nid, nvar, s.v[2]};
      // largest inverse depth wins, ties to the lower source index
      const uint64_t k = ((uint64_t)fkey(nid) << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)src);
      uint64_t& t = table[(size_t)ny * W + nx];
      if (k > t) t = k;
    }
  out.assign(n, Seed{{0, 0, 0, 0, 0, 0, 0, 0}});
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      const uint64_t k = table[(size_t)y * W + x];
      if (k == 0) continue;
      const uint32_t src = 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFu);
      if (src == 0) continue;
      const F3 in = info[src];
      out[(size_t)y * W + x].init(in.x, in.y, in.z);
    }
  regularize(out, W, H, true);
  fill_holes(out, W, H, inc);
  regularize(out, W, H, false);
}

}  // namespace

extern "C" {

// src/seed_init.cu:26-42: central differences on interior pixels, 0 elsewhere
void qmo_sd_gradient(const uint8_t* img, int W, int H, int pitch, float* grad) {
  std::memset(grad, 0, sizeof(float) * 2 * (size_t)W * H);
  for (int y = 1; y < H - 1; ++y)
    for (int x = 1; x < W - 1; ++x) {
      const float r = img[(size_t)y * pitch + x + 1], l = img[(size_t)y * pitch + x - 1];
      const float d = img[(size_t)(y + 1) * pitch + x], u = img[(size_t)(y - 1) * pitch + x];
      grad[2 * ((size_t)y * W + x)] = (float)((double)(r - l) / 2.0);
      grad[2 * ((size_t)y * W + x) + 1] = (float)((double)(d - u) / 2.0);
    }
}

// src/lsd_semidense.cu:64-81 (initial keyframe): every interior pixel with
// |gradient| >= 5 gets the default seed; seeds is W x H x 8, zeroed first
void qmo_sd_init(const float* grad, int W, int H, float* seeds) {
  std::memset(seeds, 0, sizeof(float) * 8 * (size_t)W * H);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      if (!interior(x, y, W, H)) continue;
      const float gx = grad[2 * ((size_t)y * W + x)], gy = grad[2 * ((size_t)y * W + x) + 1];
      if (gx * gx + gy * gy < kMinGrad * kMinGrad) continue;
      reinterpret_cast<Seed*>(seeds)[(size_t)y * W + x].init_default();
    }
}

// SeedMatrix::update_keyframe (src/seed_matrix.cu:399-446): update the
// keyframe seeds from the income frame, fill holes, regularise, and write the
// sparse-map pre-fill (zeroed first). Then, when switch_keyframe is set,
// SeedMatrix::create_new_keyframe (src/seed_matrix.cu:247-304) replaces the
// seeds by the propagated ones (the caller then makes the income frame the
// keyframe). T_key_world / T_inc_world are the frames' world -> camera poses.
void qmo_sd_step(float* seeds, const uint8_t* key_img, const float* key_grad, const uint8_t* inc_img,
                 const float* inc_grad, int W, int H, int pitch, float fx, float fy, float cx, float cy,
                 const float* T_key_world, const float* T_inc_world, int switch_keyframe, float* prefill,
                 float* prior /* H x W x 2 or null */) {
  const Cam c{fx, fy, cx, cy};
  const Frame key{key_img, key_grad, W, H, pitch};
  const Frame inc{inc_img, inc_grad, W, H, pitch};
  float key_inv[12], k2i[12], i2k[12];
  se3_inv(T_key_world, key_inv);
  // key_to_income = income_transform * keyframe_transform.inv() (reference SE3 product order)
  const float* L = T_inc_world;
  const float* R = key_inv;
  for (int r = 0; r < 3; ++r) {
    const int o = 4 * r;
    for (int col = 0; col < 3; ++col) k2i[o + col] = (L[o] * R[col] + L[o + 1] * R[4 + col]) + L[o + 2] * R[8 + col];
    k2i[o + 3] = ((L[o + 3] + L[o] * R[3]) + L[o + 1] * R[7]) + L[o + 2] * R[11];
  }
  se3_inv(k2i, i2k);
  const size_t n = (size_t)W * H;
  std::vector<Seed> s(n);
  std::memcpy(s.data(), seeds, sizeof(Seed) * n);
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      if (interior(x, y, W, H)) update_pixel(s[(size_t)y * W + x], x, y, W, H, c, k2i, i2k, key, inc);
  fill_holes(s, W, H, inc);
  regularize(s, W, H, false);
  std::memset(prefill, 0, sizeof(float) * n);
  project_depth(s, W, H, c, k2i, key, inc, prefill);
  if (prior) project_prior(s, W, H, c, k2i, key, inc, prior);
  if (switch_keyframe) {
    std::vector<Seed> nk;
    new_keyframe(s, W, H, c, k2i, key, inc, nk);
    s.swap(nk);
  }
  std::memcpy(seeds, s.data(), sizeof(Seed) * n);
}

// Single steps of the module, for the oracle's own pins (tests/test_oracle_pins.py).
// search_point of one key pixel: out = (result inverse depth, variance,
// epipolar length); returns the score or the -1/-2/-3 failure code.
float qmo_sd_search_point(int x, int y, int W, int H, float epx, float epy, float gradient_max, float my_gradient,
                          float min_idep, float max_idep, float fx, float fy, float cx, float cy, const float* k2i,
                          const uint8_t* key_img, const float* key_grad, const uint8_t* inc_img, const float* inc_grad,
                          float* out) {
  const Cam c{fx, fy, cx, cy};
  const Frame key{key_img, key_grad, W, H, W};
  const Frame inc{inc_img, inc_grad, W, H, W};
  float idep = 0.0f, var = 0.0f, len = 0.0f;
  const float err =
      search_point(x, y, W, H, epx, epy, gradient_max, my_gradient, min_idep, max_idep, c, k2i, key, inc, idep, var, len);
  out[0] = idep;
  out[1] = var;
  out[2] = len;
  return err;
}
// one bilinear texture fetch at (u, v) in texel-centre coordinates
float qmo_sd_tex_bilinear(const uint8_t* img, int W, int H, float u, float v) { return tex_bilinear(img, W, H, W, u, v); }
// hole filling (reads inc_grad) and regularisation of a W x H x 8 seed array, in place
void qmo_sd_fill_holes(float* seeds, int W, int H, const float* inc_grad) {
  const Frame inc{nullptr, inc_grad, W, H, W};
  std::vector<Seed> s((size_t)W * H);
  std::memcpy(s.data(), seeds, sizeof(Seed) * s.size());
  fill_holes(s, W, H, inc);
  std::memcpy(seeds, s.data(), sizeof(Seed) * s.size());
}
void qmo_sd_regularize(float* seeds, int W, int H, int remove_occlusions) {
  std::vector<Seed> s((size_t)W * H);
  std::memcpy(s.data(), seeds, sizeof(Seed) * s.size());
  regularize(s, W, H, remove_occlusions != 0);
  std::memcpy(seeds, s.data(), sizeof(Seed) * s.size());
}
// propagation of old-keyframe seeds through o2n (old -> new camera) into the new keyframe
void qmo_sd_new_keyframe(const float* old_seeds, int W, int H, float fx, float fy, float cx, float cy, const float* o2n,
                         const uint8_t* key_img, const float* key_grad, const uint8_t* inc_img, const float* inc_grad,
                         float* out_seeds) {
  const Cam c{fx, fy, cx, cy};
  const Frame key{key_img, key_grad, W, H, W};
  const Frame inc{inc_img, inc_grad, W, H, W};
  std::vector<Seed> old((size_t)W * H), out;
  std::memcpy(old.data(), old_seeds, sizeof(Seed) * old.size());
  new_keyframe(old, W, H, c, o2n, key, inc, out);
  std::memcpy(out_seeds, out.data(), sizeof(Seed) * out.size());
}

}  // extern "C"
