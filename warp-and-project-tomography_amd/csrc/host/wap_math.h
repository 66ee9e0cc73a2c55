// Host math types of the drop-in class API: column-major float 4x4 matrices, 4-vectors
// and axis-angle quaternions with the reference's operation order, so matrices built
// through these classes are bit-identical to the reference's (the projector's sample
// sets depend on every bit of them).
//   Matrix4x4   Project/Matrix4x4.{h,cpp} (products :66-84, cofactor Inverse :80-135)
//   Vector4     Project/Vector4.h (Normalize :125-131)
//   Quaternion  Project/Quaternion.{h,cpp} (axis-angle ctor :7-40, ToMatrix :46-70)
#pragma once

#include <math.h>
#include <string.h>

class Vector4 {
 public:
  float ele[4];
  Vector4() { ele[0] = ele[1] = ele[2] = 0.0f, ele[3] = 1.0f; }
  Vector4(float x, float y, float z, float w = 1.0f) {
    ele[0] = x;
    ele[1] = y;
    ele[2] = z;
    ele[3] = w;
  }
  float& operator[](int i) { return ele[i]; }
  float operator[](int i) const { return ele[i]; }
  Vector4& operator*=(float s) {
    for (float& e : ele) e *= s;
    return *this;
  }
  void Normalize() {
    const float inv = 1.0f / sqrtf(ele[0] * ele[0] + ele[1] * ele[1] + ele[2] * ele[2]);
    ele[0] *= inv;
    ele[1] *= inv;
    ele[2] *= inv;
  }
  float Length() const { return sqrtf(ele[0] * ele[0] + ele[1] * ele[1] + ele[2] * ele[2]); }
};

class Matrix4x4 {
 public:
  float ele[16];
  Matrix4x4() { IdentityMatrix(); }
  void IdentityMatrix() {
    for (int i = 0; i < 16; ++i) ele[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  }
  operator float*() { return ele; }
  operator const float*() const { return ele; }
  static Matrix4x4 TranslateMatrix(float tx, float ty, float tz) {
    Matrix4x4 m;
    m.ele[12] = tx;
    m.ele[13] = ty;
    m.ele[14] = tz;
    return m;
  }
  static Matrix4x4 ScaleMatrix(float sx, float sy, float sz) {
    Matrix4x4 m;
    m.ele[0] = sx;
    m.ele[5] = sy;
    m.ele[10] = sz;
    return m;
  }
  void Transpose() {
    for (int r = 0; r < 4; ++r)
      for (int c = r + 1; c < 4; ++c) {
        const float t = ele[4 * c + r];
        ele[4 * c + r] = ele[4 * r + c];
        ele[4 * r + c] = t;
      }
  }
  // float cofactor inverse; returns the determinant (0: matrix left unchanged)
  float Inverse();
};

inline Matrix4x4 operator*(const Matrix4x4& A, const Matrix4x4& B) {
  Matrix4x4 r;
  for (int c = 0; c < 4; ++c)
    for (int row = 0; row < 4; ++row)
      r.ele[4 * c + row] = A.ele[row] * B.ele[4 * c] + A.ele[4 + row] * B.ele[4 * c + 1] +
                           A.ele[8 + row] * B.ele[4 * c + 2] + A.ele[12 + row] * B.ele[4 * c + 3];
  return r;
}

inline Vector4 operator*(const Matrix4x4& M, const Vector4& v) {
  Vector4 r;
  for (int row = 0; row < 4; ++row)
    r.ele[row] = M.ele[row] * v.ele[0] + M.ele[4 + row] * v.ele[1] + M.ele[8 + row] * v.ele[2] +
                 M.ele[12 + row] * v.ele[3];
  return r;
}

class Quaternion {
 public:
  float x, y, z, w;
  Quaternion() { Identity(); }
  Quaternion(float ax, float ay, float az, float angle) {
    const float len = sqrtf(ax * ax + ay * ay + az * az);
    if (0.0f == len) {
      Identity();
      return;
    }
    const float li = 1.0f / len;
    const float half = angle / 2;
    w = cosf(half);
    const float s = sinf(half);
    x = (ax * li) * s;
    y = (ay * li) * s;
    z = (az * li) * s;
  }
  void Identity() {
    x = y = z = 0.0f;
    w = 1.0f;
  }
  Matrix4x4 ToMatrix() const;
};

inline Quaternion operator*(const Quaternion& p, const Quaternion& q) {
  Quaternion r;
  r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
  r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
  r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
  r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
  return r;
}
