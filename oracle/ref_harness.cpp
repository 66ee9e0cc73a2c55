// Test infrastructure only: a C-ABI wrapper around the portable pieces of the
// reference's own source files, compiled in place from the reference tree by
// oracle/Makefile into oracle/_ref/libwapref.so. It is used to validate the oracle
// restatement (oracle/wap_oracle.cpp) and to write the golden fixtures under
// tests/golden/. Nothing in the product path loads it.
//
// Reference pieces called here (file:line, paths relative to the reference root):
//   ProduceRay               Project/TypicalRayProducer.core.h:32-69
//   CastingRay               Project/RayCaster.core.h:42-188
//   T_VolumeAccess           Project/SingleScalarVolume.core.h:37-328 (trilinear + scatter)
//   advection::advect        Project/advection.h:117-146
//   L1TVOpticalFlowNonlinear3D / _MultiScale3D   Project/optical_flow.h:87-995, 1002-1126
//   CImg get_blur / get_resize / get_gradient / cut   Project/CImg.h (1.8.0)
//   Matrix4x4 / Quaternion   Project/Matrix4x4.{h,cpp}, Project/Quaternion.{h,cpp}
// The pixel loop in ref_project_view follows GenerateImage_MT
// (Project/TypicalRayProducer.core.h:168-224); that function itself needs the
// TypicalRayProducer class body from TypicalRayProducer.cpp, which includes GL and
// Windows headers this image does not have, so the loop is restated here around the
// reference's own ProduceRay/CastingRay.

#include <functional>
#include <limits>
#include <iostream>
#include <sstream>
#include <vector>
#include <thread>
#include <algorithm>
#include <cstring>
#include <cmath>
#include <ctime>

#define cimg_display 0
#include "CImg.h"
using namespace cimg_library;
using namespace std;
typedef CImg<float> image;
typedef CImgList<float> imagelist;

#include "../Shared/utilities/vec3.h"
#include "advection.h"
#include "optical_flow.h"
#include "VolumeData.h"
#include "TypicalRayProducer.h"
#include "RayCaster.h"
#include "Quaternion.h"

#ifndef max
#define max(a, b) (((a) > (b)) ? (a) : (b))
#endif
#ifndef min
#define min(a, b) (((a) < (b)) ? (a) : (b))
#endif

// GenerateImage_MT (defined in TypicalRayProducer.core.h) needs class methods from
// the unbuildable TypicalRayProducer.cpp; hidden visibility lets --gc-sections drop it.
#pragma GCC visibility push(hidden)
#include "SingleScalarVolume.core.h"
#include "RayCaster.core.h"
#include "TypicalRayProducer.core.h"
#pragma GCC visibility pop

#undef max
#undef min

namespace {

// Silence the reference's progress printing while a call runs.
struct MuteCout {
  std::streambuf* old;
  std::ostringstream sink;
  MuteCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
  ~MuteCout() { std::cout.rdbuf(old); }
};

image wrap(const float* p, int w, int h, int d, int c) {
  image im(w, h, d, c);
  std::memcpy(im.data(), p, sizeof(float) * (size_t)w * h * d * c);
  return im;
}

void unwrap(const image& im, float* out) {
  std::memcpy(out, im.data(), sizeof(float) * im.size());
}

Matrix4x4 load(const float* m) {
  Matrix4x4 r;
  for (int i = 0; i < 16; ++i) r.ele[i] = m[i];
  return r;
}

void store(const Matrix4x4& m, float* out) {
  for (int i = 0; i < 16; ++i) out[i] = m.ele[i];
}

void set_projector_state(const float* vol, const int* dims, const float* P2M,
                         const float* M2V, float sd, bool scatter, float* contrib,
                         float* weightsum, float* corr) {
  P2MMatrix = load(P2M);
  M2VMatrix = load(M2V);
  Volume_Dims[0] = dims[0];
  Volume_Dims[1] = dims[1];
  Volume_Dims[2] = dims[2];
  CroppingBounds[0] = CroppingBounds[2] = CroppingBounds[4] = 0.0f;
  CroppingBounds[1] = (float)(dims[0] - 1);
  CroppingBounds[3] = (float)(dims[1] - 1);
  CroppingBounds[5] = (float)(dims[2] - 1);
  EyePos[0] = P2MMatrix[8] / P2MMatrix[11];
  EyePos[1] = P2MMatrix[9] / P2MMatrix[11];
  EyePos[2] = P2MMatrix[10] / P2MMatrix[11];
  EyePos[3] = 1.0f;
  sampleDistance = sd;
  bsize[0] = dims[0];
  bsize[1] = dims[1];
  bsize[2] = dims[2];
  static T_VolumeAccess* access = nullptr;
  delete access;
  float spacings[3] = {1.0f, 1.0f, 1.0f};
  access = new T_VolumeAccess(wrap(vol, dims[0], dims[1], dims[2], 1), dims, spacings,
                              scatter, contrib, weightsum, corr);
  va = access;
}

}  // namespace

extern "C" {

int ref_version() { return 1; }

// Forward projection of one view: the GenerateImage_MT pixel loop over the
// reference's ProduceRay + CastingRay (cone beam through P2M / eye position).
void ref_project_view(const float* vol, const int* dims, const float* P2M,
                      const float* M2V, float sd, int nu, int nv, float* proj,
                      float* raylen, float* weight) {
  set_projector_state(vol, dims, P2M, M2V, sd, false, nullptr, nullptr, nullptr);
  for (int j = 0; j < nv; ++j) {
    for (int i = 0; i < nu; ++i) {
      const int k = i + j * nu;
      Vector4 voxel_zero, voxel_unit, dir;
      float tnear, tfar;
      ProduceRay((float)i, (float)j, voxel_zero, voxel_unit, dir, tnear, tfar);
      double value = 0.0;
      float sw = 0.0f;
      raylen[k] = 0.0f;
      if (tnear < tfar) {
        raylen[k] = tfar - tnear;
        value = CastingRay(voxel_zero, voxel_unit, dir, tnear, tfar, sw, k);
      }
      weight[k] = sw;
      proj[k] = (float)value;
    }
  }
}

// Ray setup only (ProduceRay): origin, direction, tnear, tfar for every pixel.
void ref_produce_rays(const int* dims, const float* P2M, const float* M2V, int nu, int nv,
                      float* zero4, float* unit4, float* tn, float* tf) {
  std::vector<float> vol((size_t)dims[0] * dims[1] * dims[2], 0.0f);
  set_projector_state(vol.data(), dims, P2M, M2V, 1.0f, false, nullptr, nullptr, nullptr);
  for (int j = 0; j < nv; ++j) {
    for (int i = 0; i < nu; ++i) {
      const int k = i + j * nu;
      Vector4 voxel_zero, voxel_unit, dir;
      float tnear, tfar;
      ProduceRay((float)i, (float)j, voxel_zero, voxel_unit, dir, tnear, tfar);
      for (int c = 0; c < 4; ++c) {
        zero4[4 * k + c] = voxel_zero[c];
        unit4[4 * k + c] = voxel_unit[c];
      }
      tn[k] = tnear;
      tf[k] = tfar;
    }
  }
}

// CastingRay on explicit rays (voxel-space origin/direction, entry/exit).
void ref_cast_rays(const float* vol, const int* dims, float sd, int n, const float* zero4,
                   const float* unit4, const float* tn, const float* tf, float* out,
                   float* weight) {
  float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  set_projector_state(vol, dims, I, I, sd, false, nullptr, nullptr, nullptr);
  for (int k = 0; k < n; ++k) {
    Vector4 z(zero4[4 * k], zero4[4 * k + 1], zero4[4 * k + 2], zero4[4 * k + 3]);
    Vector4 u(unit4[4 * k], unit4[4 * k + 1], unit4[4 * k + 2], unit4[4 * k + 3]);
    float sw = 0.0f;
    double v = 0.0;
    if (tn[k] < tf[k]) v = CastingRay(z, u, u, tn[k], tf[k], sw, k);
    out[k] = (float)v;
    weight[k] = sw;
  }
}

// The dormant ray-driven scatter (backprojection mode of T_VolumeAccess) on
// explicit rays: contrib[v] += corr[ray]*w, wsum[v] += w for every trilinear corner.
// The reference does not clamp the +1 corner index (those corners always carry a
// zero weight, since positions are clamped to [0, dim-1]), so the scatter runs into
// padded scratch buffers and the first W*H*D entries are copied out.
void ref_scatter_rays(const int* dims, float sd, int n, const float* zero4, const float* unit4,
                      const float* tn, const float* tf, const float* corr, float* contrib,
                      float* wsum) {
  const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
  const size_t pad = (size_t)dims[0] * dims[1] + dims[0] + 2;
  std::vector<float> vol(nvox, 0.0f), c(nvox + pad, 0.0f), w(nvox + pad, 0.0f);
  float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  set_projector_state(vol.data(), dims, I, I, sd, true, c.data(), w.data(),
                      const_cast<float*>(corr));
  for (int k = 0; k < n; ++k) {
    Vector4 z(zero4[4 * k], zero4[4 * k + 1], zero4[4 * k + 2], zero4[4 * k + 3]);
    Vector4 u(unit4[4 * k], unit4[4 * k + 1], unit4[4 * k + 2], unit4[4 * k + 3]);
    float sw = 0.0f;
    if (tn[k] < tf[k]) CastingRay(z, u, u, tn[k], tf[k], sw, k);
  }
  std::memcpy(contrib, c.data(), sizeof(float) * nvox);
  std::memcpy(wsum, w.data(), sizeof(float) * nvox);
}

void ref_advect(const int* dims, const float* aabb, const float* vel, const float* src, int nc,
                float dt, float* out) {
  MuteCout mute;
  image V = wrap(vel, dims[0], dims[1], dims[2], 3);
  image S = wrap(src, dims[0], dims[1], dims[2], nc);
  image R = advection::advect<float, image>(dims, aabb, V, S, dt);
  unwrap(R, out);
}

void ref_blur(const float* src, const int* dims, int nc, float sigma, float* out) {
  image S = wrap(src, dims[0], dims[1], dims[2], nc);
  unwrap(S.get_blur(sigma), out);
}

void ref_resize(const float* src, const int* dims, int nc, const int* ndims, int interp,
                float* out) {
  image S = wrap(src, dims[0], dims[1], dims[2], nc);
  unwrap(S.get_resize(ndims[0], ndims[1], ndims[2], nc, interp), out);
}

void ref_gradient(const float* src, const int* dims, float* out3) {
  image S = wrap(src, dims[0], dims[1], dims[2], 1);
  imagelist g = S.get_gradient("xyz", 0);
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  for (int c = 0; c < 3; ++c) std::memcpy(out3 + c * n, g[c].data(), sizeof(float) * n);
}

void ref_cut(float* data, const int* dims, float lo, float hi) {
  image S = wrap(data, dims[0], dims[1], dims[2], 1);
  S.cut(lo, hi);
  unwrap(S, data);
}

// Single-scale solver; V (3 ch) and Y1..Y3 (3 ch), Y4 (1 ch) are read and written.
void ref_flow_single(const int* dims, const float* I1, const float* I2, float* V, float* Y1,
                     float* Y2, float* Y3, float* Y4, float tol, float lambda, int max_iter,
                     int nwarps, const float* aabb, const int* roi) {
  MuteCout mute;
  int d[3] = {dims[0], dims[1], dims[2]};
  int r[6] = {roi[0], roi[1], roi[2], roi[3], roi[4], roi[5]};
  float bb[6];
  for (int i = 0; i < 6; ++i) bb[i] = aabb[i];
  image A = wrap(I1, d[0], d[1], d[2], 1), B = wrap(I2, d[0], d[1], d[2], 1);
  image v = wrap(V, d[0], d[1], d[2], 3);
  image y1 = wrap(Y1, d[0], d[1], d[2], 3), y2 = wrap(Y2, d[0], d[1], d[2], 3),
        y3 = wrap(Y3, d[0], d[1], d[2], 3), y4 = wrap(Y4, d[0], d[1], d[2], 1);
  optical_flow::L1TVOpticalFlowNonlinear3D<float, image>(d, A, B, v, y1, y2, y3, y4, tol,
                                                          lambda, max_iter, 1, nwarps, bb,
                                                          1.0f, r);
  unwrap(v, V);
  unwrap(y1, Y1);
  unwrap(y2, Y2);
  unwrap(y3, Y3);
  unwrap(y4, Y4);
}

// Multi-scale driver. V_init may be null (an empty flow, as on the first outer
// iteration of the reference loop). Returns the finest flow in V_out (3 ch).
void ref_flow_multiscale(const int* dims, const float* I1, const float* I2,
                         const float* V_init, float tol, float lambda, int max_iter,
                         int nwarps, float eta, float sigma, int scales, const int* cut,
                         float* V_out) {
  MuteCout mute;
  int d[3] = {dims[0], dims[1], dims[2]};
  int c[8];
  for (int i = 0; i < 8; ++i) c[i] = cut[i];
  image A = wrap(I1, d[0], d[1], d[2], 1), B = wrap(I2, d[0], d[1], d[2], 1);
  image v;
  if (V_init) v = wrap(V_init, d[0], d[1], d[2], 3);
  image y1(d[0], d[1], d[2], 3, 0.0f), y2(d[0], d[1], d[2], 3, 0.0f),
      y3(d[0], d[1], d[2], 3, 0.0f), y4(d[0], d[1], d[2], 1, 0.0f);
  image r = optical_flow::L1TVOpticalFlowNonlinear_MultiScale3D<float, image>(
      d, A, B, v, y1, y2, y3, y4, tol, lambda, max_iter, 1, nwarps, eta, sigma, scales, 1.0f,
      c);
  unwrap(r, V_out);
}

void ref_matrix_mul(const float* A, const float* B, float* out) {
  store(load(A) * load(B), out);
}

float ref_matrix_inverse(const float* A, float* out) {
  Matrix4x4 m = load(A);
  float det = m.Inverse();
  store(m, out);
  return det;
}

void ref_matrix_vec(const float* A, const float* v, float* out) {
  Vector4 r = load(A) * Vector4(v[0], v[1], v[2], v[3]);
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void ref_quaternion_matrix(float ax, float ay, float az, float angle, float* out) {
  Quaternion q(ax, ay, az, angle);
  Quaternion id;
  store((q * id).ToMatrix(), out);
}

void ref_translate_matrix(float tx, float ty, float tz, float* out) {
  store(Matrix4x4::TranslateMatrix(tx, ty, tz), out);
}

void ref_scale_matrix(float sx, float sy, float sz, float* out) {
  store(Matrix4x4::ScaleMatrix(sx, sy, sz), out);
}

}  // extern "C"
