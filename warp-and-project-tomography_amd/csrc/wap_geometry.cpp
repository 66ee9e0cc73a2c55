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

namespace {

struct Mat {
  float e[16];
};

Mat ident() {
  Mat m;
  for (int i = 0; i < 16; ++i) m.e[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  return m;
}

Mat mul(const Mat& A, const Mat& B) {
  Mat r;
  for (int c = 0; c < 4; ++c)
    for (int row = 0; row < 4; ++row)
      r.e[4 * c + row] = A.e[row] * B.e[4 * c] + A.e[4 + row] * B.e[4 * c + 1] +
                         A.e[8 + row] * B.e[4 * c + 2] + A.e[12 + row] * B.e[4 * c + 3];
  return r;
}

float cof3(float a, float b, float c, float d, float e, float f, float g, float h, float i) {
  return a * e * i + b * f * g + c * d * h - c * e * g - a * f * h - b * d * i;
}

Mat inverse(const Mat& m) {
  const float* s = m.e;
  const float a1 = s[0], b1 = s[1], c1 = s[2], d1 = s[3], a2 = s[4], b2 = s[5], c2 = s[6],
              d2 = s[7], a3 = s[8], b3 = s[9], c3 = s[10], d3 = s[11], a4 = s[12], b4 = s[13],
              c4 = s[14], d4 = s[15];
  const float k1 = cof3(b2, b3, b4, c2, c3, c4, d2, d3, d4);
  const float k2 = cof3(a2, a3, a4, c2, c3, c4, d2, d3, d4);
  const float k3 = cof3(a2, a3, a4, b2, b3, b4, d2, d3, d4);
  const float k4 = cof3(a2, a3, a4, b2, b3, b4, c2, c3, c4);
  const float det = a1 * k1 - b1 * k2 + c1 * k3 - d1 * k4;
  Mat r = m;
  if (det == 0.0f) return r;
  const float q = 1.0f / det;
  float* o = r.e;
  o[0] = k1 * q;
  o[4] = -k2 * q;
  o[8] = k3 * q;
  o[12] = -k4 * q;
  o[1] = -cof3(b1, b3, b4, c1, c3, c4, d1, d3, d4) * q;
  o[5] = cof3(a1, a3, a4, c1, c3, c4, d1, d3, d4) * q;
  o[9] = -cof3(a1, a3, a4, b1, b3, b4, d1, d3, d4) * q;
  o[13] = cof3(a1, a3, a4, b1, b3, b4, c1, c3, c4) * q;
  o[2] = cof3(b1, b2, b4, c1, c2, c4, d1, d2, d4) * q;
  o[6] = -cof3(a1, a2, a4, c1, c2, c4, d1, d2, d4) * q;
  o[10] = cof3(a1, a2, a4, b1, b2, b4, d1, d2, d4) * q;
  o[14] = -cof3(a1, a2, a4, b1, b2, b4, c1, c2, c4) * q;
  o[3] = -cof3(b1, b2, b3, c1, c2, c3, d1, d2, d3) * q;
  o[7] = cof3(a1, a2, a3, c1, c2, c3, d1, d2, d3) * q;
  o[11] = -cof3(a1, a2, a3, b1, b2, b3, d1, d2, d3) * q;
  o[15] = cof3(a1, a2, a3, b1, b2, b3, c1, c2, c3) * q;
  return r;
}

Mat translation(float x, float y, float z) {
  Mat m = ident();
  m.e[12] = x;
  m.e[13] = y;
  m.e[14] = z;
  return m;
}

// axis-angle quaternion composed with the identity rotation, as a matrix
Mat rotation(float ax, float ay, float az, float angle) {
  float x = 0.0f, y = 0.0f, z = 0.0f, w = 1.0f;
  const float len = sqrtf(ax * ax + ay * ay + az * az);
  if (len != 0.0f) {
    const float li = 1.0f / len;
    const float half = angle / 2;
    const float sn = sinf(half);
    w = cosf(half);
    x = (ax * li) * sn;
    y = (ay * li) * sn;
    z = (az * li) * sn;
  }
  // (x, y, z, w) * (0, 0, 0, 1)
  const float rw = w * 1.0f - x * 0.0f - y * 0.0f - z * 0.0f;
  const float rx = w * 0.0f + x * 1.0f + y * 0.0f - z * 0.0f;
  const float ry = w * 0.0f + y * 1.0f + z * 0.0f - x * 0.0f;
  const float rz = w * 0.0f + z * 1.0f + x * 0.0f - y * 0.0f;
  x = rx;
  y = ry;
  z = rz;
  w = rw;
  Mat m;
  float* e = m.e;
  e[0] = 1.0f - 2.0f * (y * y + z * z);
  e[1] = 2.0f * (x * y + z * w);
  e[2] = 2.0f * (z * x - y * w);
  e[3] = 0.0f;
  e[4] = 2.0f * (x * y - z * w);
  e[5] = 1.0f - 2.0f * (z * z + x * x);
  e[6] = 2.0f * (y * z + x * w);
  e[7] = 0.0f;
  e[8] = 2.0f * (z * x + y * w);
  e[9] = 2.0f * (y * z - x * w);
  e[10] = 1.0f - 2.0f * (x * x + y * y);
  e[11] = 0.0f;
  e[12] = e[13] = e[14] = 0.0f;
  e[15] = 1.0f;
  return m;
}

}  // namespace

extern "C" int wap_view_geometry(const WapScanGeometry* g, double angle_deg, WapView* out) {
  if (!g || !out) return WAP_E_ARG;
  if (g->vol_dims[0] < 2 || g->vol_dims[1] < 2 || g->vol_dims[2] < 2 || g->det_u < 2 ||
      g->det_v < 2 || !(g->vol_spacing > 0.0f) || !(g->det_pitch > 0.0))
    return WAP_E_ARG;
  // rotation about +y (counter-clockwise) or -y, angle in double, then to float
  const double angle = angle_deg * 3.1415926 / 180;
  float axis_y = g->counterclockwise ? 1.0f : -1.0f;
  const double alen = (double)sqrtf(axis_y * axis_y);
  const float unit = (float)(1.0f / alen);
  const Mat R = rotation(0.0f * unit, axis_y * unit, 0.0f * unit, (float)(alen * angle));
  // view = T(tx, ty, -sid + tz) * S(1) * R
  const float zbase = -1 * g->sid;
  Mat S = ident();
  const Mat view =
      mul(mul(translation(g->translate[0], g->translate[1], zbase + g->translate[2]), S), R);
  const double hw = (g->det_u - 1) * g->det_pitch * 0.5f;
  const double hh = (g->det_v - 1) * g->det_pitch * 0.5f;
  const float nearp = g->sdd, farp = 10000.0f;
  Mat P;
  memset(P.e, 0, sizeof(P.e));
  P.e[0] = (float)(nearp / hw);
  P.e[5] = (float)(nearp / hh);
  P.e[10] = -(farp + nearp) / (farp - nearp);
  P.e[11] = -1.0f;
  P.e[14] = -2.0f * farp * nearp / (farp - nearp);
  const float sp = g->vol_spacing;
  const float size[3] = {(g->vol_dims[0] - 1) * sp, (g->vol_dims[1] - 1) * sp,
                         (g->vol_dims[2] - 1) * sp};
  const Mat MV = mul(view, translation(-size[0] * 0.5f, -size[1] * 0.5f, -size[2] * 0.5f));
  Mat M2V = ident();
  M2V.e[0] = M2V.e[5] = M2V.e[10] = 1.0f / sp;
  const Mat V2M = inverse(M2V);
  Mat toPix = ident();
  toPix.e[0] = 0.5f * (g->det_u - 1);
  toPix.e[5] = 0.5f * (g->det_v - 1);
  toPix.e[10] = 0.5f;
  toPix.e[12] = toPix.e[0];
  toPix.e[13] = toPix.e[5];
  toPix.e[14] = toPix.e[10];
  memset(out, 0, sizeof(*out));
  out->half_w = hw;
  out->half_h = hh;
  out->ds = g->det_pitch;
  out->sdd = g->sdd;
  out->parallel = g->parallel ? 1 : 0;
  memcpy(out->m2v, M2V.e, sizeof(M2V.e));
  const Mat A = mul(MV, V2M);
  memcpy(out->mvv2m, A.e, sizeof(A.e));
  if (!g->parallel) {
    const Mat P2M = inverse(mul(mul(toPix, P), MV));
    memcpy(out->p2m, P2M.e, sizeof(P2M.e));
    out->eye[0] = P2M.e[8] / P2M.e[11];
    out->eye[1] = P2M.e[9] / P2M.e[11];
    out->eye[2] = P2M.e[10] / P2M.e[11];
    out->eye[3] = 1.0f;
  } else {
    // detector pixel (i, j) sits at view (i*ds - hw, j*ds - hh, 0); rays run along -z_view
    Mat pix = ident();
    pix.e[0] = pix.e[5] = (float)g->det_pitch;
    pix.e[12] = (float)-hw;
    pix.e[13] = (float)-hh;
    const Mat iMV = inverse(MV);
    const Mat P2M = mul(iMV, pix);
    memcpy(out->p2m, P2M.e, sizeof(P2M.e));
    float d[4];
    for (int r = 0; r < 4; ++r)
      d[r] = iMV.e[r] * 0.0f + iMV.e[4 + r] * 0.0f + iMV.e[8 + r] * -1.0f + iMV.e[12 + r] * 0.0f;
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
