// Host-side scan geometry and loop bookkeeping of the wap library.
//
// Per-view matrices are built in float with the reference's operation order, because
// the projector's sample sets (and so the projections) depend on every bit of them:
//   rotation  PSART::RotateVolume                Project/WaP_Tomo.cpp:1764-1797
//             Quaternion(axis, angle), ToMatrix   Project/Quaternion.cpp:7-40, Quaternion.h:46-86
//   camera    Scene::_calculateCamera/_calculateMatrix   Project/Scene.cpp:65-127, 163-189
//   volume    TypicalRayProducer::_calculateMatrix      Project/TypicalRayProducer.cpp:157-232
//   inverse   Matrix4x4::Inverse (float cofactors)      Project/Matrix4x4.cpp:80-135
//   eye       P2M column 2 / P2M[11]                    Project/TypicalRayProducer.core.h:255-258
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/wap.h"
#include "host/wap_math.h"

extern "C" int wap_view_geometry(const WapScanGeometry* g, double angle_deg, WapView* out) {
  if (!g || !out) return WAP_E_ARG;
  if (g->vol_dims[0] < 2 || g->vol_dims[1] < 2 || g->vol_dims[2] < 2 || g->det_u < 2 ||
      g->det_v < 2 || !(g->vol_spacing > 0.0f) || !(g->det_pitch > 0.0))
    return WAP_E_ARG;
  // rotation about +y (counter-clockwise) or -y; the angle is in double, then float
  const double angle = angle_deg * 3.1415926 / 180;
  Vector4 axis(0.0f, g->counterclockwise ? 1.0f : -1.0f, 0.0f, 0.0f);
  const double alen = axis.Length();
  axis *= (float)(1.0f / alen);
  const Matrix4x4 R = (Quaternion(axis[0], axis[1], axis[2], (float)(alen * angle)) *
                       Quaternion())
                          .ToMatrix();
  // view = T(tx, ty, -sid + tz) * S(1) * R, perspective with near = sdd, far = 10000
  const float zbase = -1 * g->sid;
  const Matrix4x4 view = Matrix4x4::TranslateMatrix(g->translate[0], g->translate[1],
                                                    zbase + g->translate[2]) *
                         Matrix4x4::ScaleMatrix(1.0f, 1.0f, 1.0f) * R;
  const double hw = (g->det_u - 1) * g->det_pitch * 0.5f;
  const double hh = (g->det_v - 1) * g->det_pitch * 0.5f;
  const float nearp = g->sdd, farp = 10000.0f;
  Matrix4x4 P;
  memset(P.ele, 0, sizeof(P.ele));
  P.ele[0] = (float)(nearp / hw);
  P.ele[5] = (float)(nearp / hh);
  P.ele[10] = -(farp + nearp) / (farp - nearp);
  P.ele[11] = -1.0f;
  P.ele[14] = -2.0f * farp * nearp / (farp - nearp);
  const float sp = g->vol_spacing;
  const float size[3] = {(g->vol_dims[0] - 1) * sp, (g->vol_dims[1] - 1) * sp,
                         (g->vol_dims[2] - 1) * sp};
  const Matrix4x4 MV =
      view * Matrix4x4::TranslateMatrix(-size[0] * 0.5f, -size[1] * 0.5f, -size[2] * 0.5f);
  Matrix4x4 M2V;
  M2V.ele[0] = M2V.ele[5] = M2V.ele[10] = 1.0f / sp;
  Matrix4x4 V2M = M2V;
  V2M.Inverse();
  Matrix4x4 toPix;
  toPix.ele[0] = 0.5f * (g->det_u - 1);
  toPix.ele[5] = 0.5f * (g->det_v - 1);
  toPix.ele[10] = 0.5f;
  toPix.ele[12] = toPix.ele[0];
  toPix.ele[13] = toPix.ele[5];
  toPix.ele[14] = toPix.ele[10];
  memset(out, 0, sizeof(*out));
  out->half_w = hw;
  out->half_h = hh;
  out->ds = g->det_pitch;
  out->sdd = g->sdd;
  out->parallel = g->parallel ? 1 : 0;
  memcpy(out->m2v, M2V.ele, sizeof(M2V.ele));
  const Matrix4x4 A = MV * V2M;
  memcpy(out->mvv2m, A.ele, sizeof(A.ele));
  if (!g->parallel) {
    Matrix4x4 P2M = toPix * P * MV;
    P2M.Inverse();
    memcpy(out->p2m, P2M.ele, sizeof(P2M.ele));
    out->eye[0] = P2M.ele[8] / P2M.ele[11];
    out->eye[1] = P2M.ele[9] / P2M.ele[11];
    out->eye[2] = P2M.ele[10] / P2M.ele[11];
    out->eye[3] = 1.0f;
  } else {
    // detector pixel (i, j) sits at view (i*ds - hw, j*ds - hh, 0); rays run along -z_view
    Matrix4x4 pix;
    pix.ele[0] = pix.ele[5] = (float)g->det_pitch;
    pix.ele[12] = (float)-hw;
    pix.ele[13] = (float)-hh;
    Matrix4x4 iMV = MV;
    iMV.Inverse();
    const Matrix4x4 P2M = iMV * pix;
    memcpy(out->p2m, P2M.ele, sizeof(P2M.ele));
    const Vector4 d = iMV * Vector4(0.0f, 0.0f, -1.0f, 0.0f);
    const float li = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    out->dir[0] = d[0] * li;
    out->dir[1] = d[1] * li;
    out->dir[2] = d[2] * li;
    out->dir[3] = 0.0f;
  }
  return WAP_OK;
}

extern "C" int wap_view_order(int first, int last, int* order) {
  if (!order || last < first) return WAP_E_ARG;
  const int num = last - first + 1;
  for (int i = 0; 2 * i < num; ++i) order[2 * i] = first + i;
  for (int i = 0; 2 * i + 1 < num; ++i) order[2 * i + 1] = last - i;
  return WAP_OK;
}

extern "C" int wap_frame_windows(int nframes, int total_views, float base, float* ftime,
                                 int* start, int* end) {
  if (nframes <= 0 || total_views <= 0 || !ftime || !start || !end) return WAP_E_ARG;
  const float unit = (float)((float)total_views / (float)nframes);
  for (int i = 0; i < nframes; ++i) {
    ftime[i] = unit * i;
    start[i] = std::max(0, (int)(ftime[i] - base));
    end[i] = std::min((int)(ftime[i] + base), total_views - 1);
  }
  return WAP_OK;
}

extern "C" int wap_low_discrepancy_angles(int nproj, int nseg, int base, int zero_every,
                                          double* angles) {
  if (nproj < 0 || nseg < 1 || base < 2 || zero_every < 0 || (nproj > 0 && !angles))
    return WAP_E_ARG;
  const double step = 360.0;
  int n = 0;
  for (long long idx = 0; n < nproj; ++idx) {
    // reference view at 0 degrees each time the count reaches a multiple of zero_every
    if (zero_every > 0 && n > 0 && n % zero_every == 0) angles[n++] = 0.0;
    // radical inverse of idx in the given base, digits taken least significant first and
    // summed in that order (the order fixes the double result)
    double frac = 0.0, weight = 1.0 / base;
    for (long long rest = idx; rest != 0; rest /= base) {
      frac = frac + (double)(rest % base) * weight;
      weight = weight / base;
    }
    const double first = step * frac / nseg;
    for (int seg = 0; seg < nseg && n < nproj; ++seg) angles[n++] = first + (double)seg * step / nseg;
  }
  return WAP_OK;
}

extern "C" int wap_roi_mask(const int* dims, const int* r, float* mask) {
  if (!dims || !r || !mask) return WAP_E_ARG;
  const int W = dims[0], H = dims[1], D = dims[2];
  for (int k = 0; k < D; ++k)
    for (int j = 0; j < H; ++j)
      for (int i = 0; i < W; ++i) {
        const bool lowx = i < r[6], highx = i > W - r[6];
        const bool lowz = k < r[7], highz = k > D - r[7];
        bool keep = !((lowx || highx) && (lowz || highz));
        // the same corner rule with the roles of x and z swapped
        if ((k < r[6] || k > D - r[6]) && (i < r[7] || i > W - r[7])) keep = false;
        if (i < r[0] || i > W - r[1] || j < r[2] || j > H - r[3] || k < r[4] || k > D - r[5])
          keep = false;
        mask[(size_t)i + (size_t)W * ((size_t)j + (size_t)H * k)] = keep ? 1.0f : 0.0f;
      }
  return WAP_OK;
}
