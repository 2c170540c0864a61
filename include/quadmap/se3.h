// quadmap/se3.h -- SE3<float> for callers of quadmap::Depthmap, plain C++.
//
// The interface of this file (the constructors, data(r, c), inv(), the products
// and their operation order) follows include/quadmap/se3.cuh of the reference,
// which is part of REMODE - REgularized MOnocular Depth Estimation:
//   Copyright (C) 2014 Matia Pizzoli <matia dot pizzoli at gmail dot com>
//   Robotics and Perception Group, University of Zurich, Switzerland
//   http://rpg.ifi.uzh.ch
// REMODE is free software under the GNU General Public License, version 3 or
// any later version; this file is distributed under the same license, WITHOUT
// ANY WARRANTY, without even the implied warranty of MERCHANTABILITY or FITNESS
// FOR A PARTICULAR PURPOSE (see <http://www.gnu.org/licenses/>).
//
// Same surface as the reference's include/quadmap/se3.cuh:26-201 (identity and
// quaternion constructors, data(r, c), inv(), products with SE3 and float3,
// getTranslation) so caller code compiles unchanged; the operation order of
// inv() and operator* is the reference's. float3 here is a plain struct.
#pragma once

#include <cmath>
#include <ostream>

#ifndef QUADMAP_HAVE_FLOAT3
#define QUADMAP_HAVE_FLOAT3
struct float3 {
  float x, y, z;
};
inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline float dot(const float3& a, const float3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float length(const float3& v) { return std::sqrt(dot(v, v)); }
inline float3 operator-(const float3& a, const float3& b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
#endif

namespace quadmap {

template <typename Type, unsigned R, unsigned C>
struct Matrix {
  Type operator()(int row, int col) const { return data[row * C + col]; }
  Type& operator()(int row, int col) { return data[row * C + col]; }
  Type operator[](int i) const { return data[i]; }
  Type& operator[](int i) { return data[i]; }
  Type data[R * C];
};

template <typename Type>
struct SE3 {
  SE3() {
    for (int i = 0; i < 12; ++i) data[i] = 0;
    data(0, 0) = data(1, 1) = data(2, 2) = (Type)1;
  }
  // normalised quaternion (w, x, y, z) and translation
  SE3(Type qw, Type qx, Type qy, Type qz, Type tx, Type ty, Type tz) {
    const Type x = 2 * qx, y = 2 * qy, z = 2 * qz;
    const Type wx = x * qw, wy = y * qw, wz = z * qw;
    const Type xx = x * qx, xy = y * qx, xz = z * qx;
    const Type yy = y * qy, yz = z * qy, zz = z * qz;
    data(0, 0) = 1 - (yy + zz);
    data(0, 1) = xy - wz;
    data(0, 2) = xz + wy;
    data(1, 0) = xy + wz;
    data(1, 1) = 1 - (xx + zz);
    data(1, 2) = yz - wx;
    data(2, 0) = xz - wy;
    data(2, 1) = yz + wx;
    data(2, 2) = 1 - (xx + yy);
    data(0, 3) = tx;
    data(1, 3) = ty;
    data(2, 3) = tz;
  }
  // r: row-major rotation, t: translation
  SE3(const Type* r, const Type* t) {
    data[0] = r[0]; data[1] = r[1]; data[2] = r[2]; data[3] = t[0];
    data[4] = r[3]; data[5] = r[4]; data[6] = r[5]; data[7] = t[1];
    data[8] = r[6]; data[9] = r[7]; data[10] = r[8]; data[11] = t[2];
  }
  SE3<Type> inv() const {
    SE3<Type> r;
    r.data[0] = data[0]; r.data[1] = data[4]; r.data[2] = data[8];
    r.data[4] = data[1]; r.data[5] = data[5]; r.data[6] = data[9];
    r.data[8] = data[2]; r.data[9] = data[6]; r.data[10] = data[10];
    r.data[3] = -data[0] * data[3] - data[4] * data[7] - data[8] * data[11];
    r.data[7] = -data[1] * data[3] - data[5] * data[7] - data[9] * data[11];
    r.data[11] = -data[2] * data[3] - data[6] * data[7] - data[10] * data[11];
    return r;
  }
  Type operator()(int r, int c) const { return data(r, c); }
  Type& operator()(int r, int c) { return data(r, c); }
  float3 rotate(const float3& p) const {
    return make_float3(data(0, 0) * p.x + data(0, 1) * p.y + data(0, 2) * p.z,
                       data(1, 0) * p.x + data(1, 1) * p.y + data(1, 2) * p.z,
                       data(2, 0) * p.x + data(2, 1) * p.y + data(2, 2) * p.z);
  }
  float3 translate(const float3& p) const { return make_float3(p.x + data(0, 3), p.y + data(1, 3), p.z + data(2, 3)); }
  float3 getTranslation() const { return make_float3(data(0, 3), data(1, 3), data(2, 3)); }
  float3 getinvTranslation() const { return inv().getTranslation(); }
  friend std::ostream& operator<<(std::ostream& out, const SE3& m) {
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 4; ++c) out << m.data(r, c) << " ";
      out << "\n";
    }
    return out;
  }
  Matrix<Type, 3, 4> data;
};

template <typename Type>
inline SE3<Type> operator*(const SE3<Type>& l, const SE3<Type>& r) {
  SE3<Type> o;
  for (int row = 0; row < 3; ++row) {
    const int k = 4 * row;
    for (int col = 0; col < 3; ++col) o.data[k + col] = l.data[k] * r.data[col] + l.data[k + 1] * r.data[4 + col] + l.data[k + 2] * r.data[8 + col];
    o.data[k + 3] = l.data[k + 3] + l.data[k] * r.data[3] + l.data[k + 1] * r.data[7] + l.data[k + 2] * r.data[11];
  }
  return o;
}

inline float3 operator*(const SE3<float>& se3, const float3& p) { return se3.translate(se3.rotate(p)); }

}  // namespace quadmap
