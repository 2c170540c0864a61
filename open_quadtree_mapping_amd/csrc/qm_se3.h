// qm_se3.h -- host-side rigid transforms as float[12] (3x4 row-major [R | t])
// with the operation order of the reference's SE3<float> product and inverse
// (include/quadmap/se3.cuh:90-107, 177-195), so relative poses are bit-identical
// to what the reference host code computes from the same absolute poses.
#pragma once

#include <cmath>
#include <cstring>

namespace qm {

inline void se3_identity(float* T) {
  std::memset(T, 0, sizeof(float) * 12);
  T[0] = T[5] = T[10] = 1.0f;
}

inline void se3_inv(const float* d, float* r) {
  float o[12];
  o[0] = d[0];
  o[1] = d[4];
  o[2] = d[8];
  o[4] = d[1];
  o[5] = d[5];
  o[6] = d[9];
  o[8] = d[2];
  o[9] = d[6];
  o[10] = d[10];
  o[3] = ((-d[0]) * d[3] - d[4] * d[7]) - d[8] * d[11];
  o[7] = ((-d[1]) * d[3] - d[5] * d[7]) - d[9] * d[11];
  o[11] = ((-d[2]) * d[3] - d[6] * d[7]) - d[10] * d[11];
  std::memcpy(r, o, sizeof(o));
}

inline void se3_mul(const float* a, const float* b, float* r) {
  float o[12];
  for (int row = 0; row < 3; ++row) {
    const int k = 4 * row;
    for (int col = 0; col < 3; ++col) o[k + col] = (a[k] * b[col] + a[k + 1] * b[4 + col]) + a[k + 2] * b[8 + col];
    o[k + 3] = ((a[k + 3] + a[k] * b[3]) + a[k + 1] * b[7]) + a[k + 2] * b[11];
  }
  std::memcpy(r, o, sizeof(o));
}

// cos of the angle between the optical axes and the distance between the
// camera centres of two world->camera poses (src/seed_matrix.cu:171-183,200-209)
inline void se3_axis_cos_baseline(const float* T_a_world, const float* T_b_world, float* cosv, float* base) {
  float A[12], B[12];
  se3_inv(T_a_world, A);
  se3_inv(T_b_world, B);
  *cosv = (A[2] * B[2] + A[6] * B[6]) + A[10] * B[10];
  const float dx = A[3] - B[3], dy = A[7] - B[7], dz = A[11] - B[11];
  *base = std::sqrt((dx * dx + dy * dy) + dz * dz);
}

}  // namespace qm
