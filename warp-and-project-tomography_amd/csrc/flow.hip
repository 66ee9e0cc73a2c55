// L1 optical-flow primal-dual iteration for gfx950.
//
// Reference semantics: optical_flow::L1TVOpticalFlowNonlinear3D
// (Project/optical_flow.h:87-995) with the finite differences of Project/tools.cpp:247-314.
// Per warp: u = central gradient of the warped second state, ut = I2w - I1 - u.v over
// the ROI, tau_c = 1/(4 + |u_c|), sigma2 = 1/(|ux|+|uy|+|uz|), then iterations of
//   primal (ROI):  Kty_c = -(Dx- y_c1 + Dy- y_c2 + Dz- y_c3) + u_c*y4;  v_c -= tau_c*Kty_c
//   dual (all):    y_cd  = (y_cd + s1*(2*Dd+ v_c - Dd+ v_c_prev)) * lam/(lam + 2*s1)
//                  y4    = clip(y4 + sigma2*(2*u.v - u.v_prev + ut), -1, 1)
// with a residual every 50 iterations deciding whether to go on.
//
// One kernel per iteration (k_flow_pd) reads v (3 planes), y (10), u (3), ut (1) and
// writes v and y (both double-buffered): 30 floats = 120 B per voxel. The reference
// keeps Kx/KxOld/Kty arrays besides (about 70 planes); here the previous iteration's
// K*x terms are recomputed from the previous v and the neighbours' primal step from the
// old y, which gives the same floats because the operations are identical. tau and
// sigma2 are recomputed from u. The run/stop decision lives on the device (ctl[0]) so a
// whole warp of iterations can be queued (or captured in a graph) without host round
// trips.
#include "wap_device.h"

namespace {

struct FlowGrid {
  int W, H, D;
  int r0, r1, r2, r3, r4, r5;  // ROI margins
};

WAP_DEV float my_abs(float x) { return x > 0.0f ? x : -x; }

// Backward difference with the reference's boundary rule (i==0: y[i]; i==n-1: -y[i-1]).
WAP_DEV float dminus(const float* __restrict__ a, size_t t, int c, int n, size_t stride) {
  if (c > 0 && c < n - 1) return a[t] - a[t - stride];
  if (c == 0) return a[t];
  return -a[t - stride];
}

template <int BS>
WAP_DEV double block_sum(double v) {
  __shared__ double red[BS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; ++w) s += red[w];
  return s;
}

// u and ut over the ROI from the warped state (central differences, clamped borders).
__global__ __launch_bounds__(256) void k_flow_prep(FlowGrid G, const float* __restrict__ warped,
                                                   const float* __restrict__ I1,
                                                   const float* __restrict__ v,
                                                   float* __restrict__ u,
                                                   float* __restrict__ ut) {
  const int bx = G.W - G.r0 - G.r1, by = G.H - G.r2 - G.r3, bz = G.D - G.r4 - G.r5;
  const size_t nb = (size_t)bx * by * bz;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  const int i = G.r0 + (int)(q % bx), j = G.r2 + (int)((q / bx) % by),
            k = G.r4 + (int)(q / ((size_t)bx * by));
  const size_t n = (size_t)G.W * G.H * G.D;
  const size_t t = vidx(i, j, k, G.W, G.H);
  auto at = [&](int a, int b, int c) {
    return warped[vidx(iclamp(a, 0, G.W - 1), iclamp(b, 0, G.H - 1), iclamp(c, 0, G.D - 1),
                       G.W, G.H)];
  };
  const float gx = (at(i + 1, j, k) - at(i - 1, j, k)) / 2.0f;
  const float gy = (at(i, j + 1, k) - at(i, j - 1, k)) / 2.0f;
  const float gz = (at(i, j, k + 1) - at(i, j, k - 1)) / 2.0f;
  u[t] = gx;
  u[n + t] = gy;
  u[2 * n + t] = gz;
  ut[t] = warped[t] - I1[t] - gx * v[t] - gy * v[n + t] - gz * v[2 * n + t];
}

// u.v of the outgoing warp, needed as the previous K*x on the first iteration of the
// next warp (u changes between warps, v does not).
__global__ void k_flow_kx4(size_t n, const float* __restrict__ u, const float* __restrict__ v,
                           float* __restrict__ kx4) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  kx4[t] = u[t] * v[t] + u[n + t] * v[n + t] + u[2 * n + t] * v[2 * n + t];
}

WAP_DEV bool in_roi(const FlowGrid& G, int i, int j, int k) {
  return i >= G.r0 && i < G.W - G.r1 && j >= G.r2 && j < G.H - G.r3 && k >= G.r4 &&
         k < G.D - G.r5;
}

// Primal update of flow component c at voxel (i, j, k) = linear index s:
// inside the ROI  v - tau*K^T y  with  K^T y = -(Dx- y_c1 + Dy- y_c2 + Dz- y_c3) + u_c*y4,
// outside it the unchanged v. K receives K^T y (inside the ROI).
WAP_DEV float primal_at(const FlowGrid& G, const float* __restrict__ vold,
                        const float* __restrict__ y, const float* __restrict__ u, size_t n, int c,
                        int i, int j, int k, size_t s, float* K, float* tau) {
  const float vo = vold[c * n + s];
  if (!in_roi(G, i, j, k)) return vo;
  const size_t sy = (size_t)G.W, sz = (size_t)G.W * G.H;
  const float uc = u[c * n + s];
  const float kt = -1.0f * dminus(y + (3 * c + 0) * n, s, i, G.W, 1) -
                   1.0f * dminus(y + (3 * c + 1) * n, s, j, G.H, sy) -
                   1.0f * dminus(y + (3 * c + 2) * n, s, k, G.D, sz) + uc * y[9 * n + s];
  const float ta = 1.0f / (4.0f / 1.0f + my_abs(uc));
  if (K) *K = kt;
  if (tau) *tau = ta;
  return vo - ta * kt;
}

// One fused primal-dual iteration. A workgroup owns a TX x TY column tile and marches
// it through ZC planes. For every plane it computes the new v (the primal step) once
// for the tile plus a one-voxel +x/+y halo into LDS, keeping the previous plane's, so
// the dual step at (x, y, z) finds the new v at itself and its +x/+y/+z neighbours in
// LDS. v and y are each read once and written once per iteration from HBM (both
// double-buffered); u and ut are read once.
// mode 0: previous K*x = 0 (first iteration of the solve); mode 1: previous u.v from
// kx4prev (first iteration of a later warp); mode 2: normal.
// resid: per-workgroup FP64 sums of the primal (ROI) and dual residual terms;
// store_kty: keep K^T y (ROI) for the next iteration's primal residual.
template <int TX, int TY, int ZC>
__global__ __launch_bounds__(TX * TY) void k_flow_pd(
    FlowGrid G, const float* __restrict__ vold, float* __restrict__ vnew,
    const float* __restrict__ yold, float* __restrict__ ynew, const float* __restrict__ u,
    const float* __restrict__ ut, const float* __restrict__ kx4prev, float* __restrict__ kty,
    int mode, int resid, int store_kty, float sigma1, float lambda, double* __restrict__ ppart,
    double* __restrict__ dpart, int* __restrict__ ctl, int iter) {
  constexpr int PX = TX + 1, PY = TY + 1, NP = PX * PY;
  __shared__ float pl[2][3][NP];
  if (ctl[0] == 0) return;
  const size_t n = (size_t)G.W * G.H * G.D;
  const int tilesX = (G.W + TX - 1) / TX, tilesY = (G.H + TY - 1) / TY;
  const int bid = (int)blockIdx.x;
  const int x0 = (bid % tilesX) * TX;
  const int y0 = ((bid / tilesX) % tilesY) * TY;
  const int z0 = (bid / (tilesX * tilesY)) * ZC;
  const int z1 = min(z0 + ZC, G.D);
  const int tx = (int)threadIdx.x % TX, ty = (int)threadIdx.x / TX;
  const int i = x0 + tx, j = y0 + ty;
  const bool own = i < G.W && j < G.H;
  if (bid == 0 && threadIdx.x == 0) ctl[1] = iter;
  const float shrink_den = lambda + 2.0f * sigma1;
  double pterm = 0.0, dterm = 0.0;
  // primal of the plane k (tile + halo) into slot b
  auto fill = [&](int b, int k) {
    for (int q = (int)threadIdx.x; q < NP; q += TX * TY) {
      const int qi = x0 + q % PX, qj = y0 + q / PX;
      if (qi < G.W && qj < G.H) {
        const size_t s = vidx(qi, qj, k, G.W, G.H);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          pl[b][c][q] = primal_at(G, vold, yold, u, n, c, qi, qj, k, s, nullptr, nullptr);
      }
    }
  };
  int cur = 0;
  fill(cur, z0);
  for (int k = z0; k < z1; ++k) {
    if (k + 1 < G.D) fill(cur ^ 1, k + 1);
    __syncthreads();
    if (own) {
      const size_t t = vidx(i, j, k, G.W, G.H);
      const size_t off[3] = {1, (size_t)G.W, (size_t)G.W * G.H};
      const bool has[3] = {i < G.W - 1, j < G.H - 1, k < G.D - 1};
      const int q0 = tx + PX * ty;
      const float ux = u[t], uy = u[n + t], uz = u[2 * n + t];
      const bool roi_t = in_roi(G, i, j, k);
      float vn[3], vo[3];
      float pt = 0.0f, dt = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        vo[c] = vold[c * n + t];
        vn[c] = pl[cur][c][q0];
        vnew[c * n + t] = vn[c];
        if (roi_t && (resid || store_kty)) {
          float K0 = 0.0f, tau0 = 1.0f;
          primal_at(G, vold, yold, u, n, c, i, j, k, t, &K0, &tau0);
          if (resid) pt += fabsf((vo[c] - vn[c]) / tau0 - kty[c * n + t] + K0);
          if (store_kty) kty[c * n + t] = K0;
        }
        const float vnb[3] = {pl[cur][c][q0 + 1], pl[cur][c][q0 + PX], pl[cur ^ 1][c][q0]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          float kx = 0.0f, ko = 0.0f;
          if (has[a]) {
            kx = 1.0f * (vnb[a] - vn[c]);
            if (mode != 0) ko = 1.0f * (vold[c * n + t + off[a]] - vo[c]);
          }
          const size_t p = (size_t)(3 * c + a) * n + t;
          const float yo = yold[p];
          const float yt = yo + sigma1 * (kx + kx - ko);
          const float yn = yt * lambda / shrink_den;
          ynew[p] = yn;
          if (resid) dt += fabsf((yo - yn) / sigma1 - ko + kx);
        }
      }
      const float k4 = ux * vn[0] + uy * vn[1] + uz * vn[2];
      const float k4o =
          mode == 0 ? 0.0f : (mode == 1 ? kx4prev[t] : ux * vo[0] + uy * vo[1] + uz * vo[2]);
      const float s2 = 1.0f / (fabsf(ux) + fabsf(uy) + fabsf(uz));
      const float y4o = yold[9 * n + t];
      const float y4t = y4o + s2 * (k4 + k4 - k4o + ut[t]);
      const float m1 = (y4t < 1.0f) ? y4t : 1.0f;
      const float y4n = (-1.0f < m1) ? m1 : -1.0f;
      ynew[9 * n + t] = y4n;
      if (resid) {
        dt += fabsf((y4o - y4n) / s2 - k4o + k4);
        pterm += (double)pt;
        dterm += (double)dt;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  if (resid) {
    const double ps = block_sum<TX * TY>(pterm);
    __syncthreads();
    const double ds = block_sum<TX * TY>(dterm);
    if (threadIdx.x == 0) {
      ppart[blockIdx.x] = ps;
      dpart[blockIdx.x] = ds;
    }
  }
}

// Two-kernel form of one iteration (the primal over the ROI, then the dual over every
// voxel): more HBM traffic than the fused form but no recomputation; kept as a tuning
// variant.
// primal step over the ROI: vdst = vsrc - tau*K^T y  (vdst holds vsrc outside the ROI)
__global__ __launch_bounds__(256) void k_flow_primal(
    FlowGrid G, const float* __restrict__ vsrc, float* __restrict__ vdst,
    const float* __restrict__ y, const float* __restrict__ u, float* __restrict__ kty,
    int store_kty, int resid, double* __restrict__ ppart, const int* __restrict__ ctl) {
  if (ctl[0] == 0) return;
  const int bx = G.W - G.r0 - G.r1, by = G.H - G.r2 - G.r3, bz = G.D - G.r4 - G.r5;
  const size_t nb = (size_t)bx * by * bz;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double pterm = 0.0;
  if (q < nb) {
    const int i = G.r0 + (int)(q % bx), j = G.r2 + (int)((q / bx) % by),
              k = G.r4 + (int)(q / ((size_t)bx * by));
    const size_t n = (size_t)G.W * G.H * G.D;
    const size_t t = vidx(i, j, k, G.W, G.H);
    const size_t sy = (size_t)G.W, sz = (size_t)G.W * G.H;
    const float y4 = y[9 * n + t];
    float term = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float uc = u[c * n + t];
      const float K = -1.0f * dminus(y + (3 * c + 0) * n, t, i, G.W, 1) -
                      1.0f * dminus(y + (3 * c + 1) * n, t, j, G.H, sy) -
                      1.0f * dminus(y + (3 * c + 2) * n, t, k, G.D, sz) + uc * y4;
      const float tau = 1.0f / (4.0f / 1.0f + my_abs(uc));
      const float vo = vsrc[c * n + t];
      const float vn = vo - tau * K;
      vdst[c * n + t] = vn;
      if (resid) term += fabsf((vo - vn) / tau - kty[c * n + t] + K);
      if (store_kty) kty[c * n + t] = K;
    }
    pterm = (double)term;
  }
  if (resid) {
    const double s = block_sum<256>(pterm);
    if (threadIdx.x == 0) ppart[blockIdx.x] = s;
  }
}

// dual step over every voxel; mode 0: previous K*x = 0 (first iteration of the solve),
// mode 1: previous u.v from kx4prev (first iteration of a later warp), mode 2: normal
__global__ __launch_bounds__(256) void k_flow_dual(
    FlowGrid G, const float* __restrict__ vold, const float* __restrict__ vnew,
    const float* __restrict__ y, const float* __restrict__ u, const float* __restrict__ ut,
    const float* __restrict__ kx4prev, int mode, int resid, float sigma1, float lambda,
    double* __restrict__ dpart, int* __restrict__ ctl, int iter, float* __restrict__ yout) {
  if (ctl[0] == 0) return;
  const size_t n = (size_t)G.W * G.H * G.D;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dterm = 0.0;
  if (t < n) {
    if (t == 0) ctl[1] = iter;
    const int i = (int)(t % G.W), j = (int)((t / G.W) % G.H), k = (int)(t / ((size_t)G.W * G.H));
    const size_t off[3] = {1, (size_t)G.W, (size_t)G.W * G.H};
    const bool has[3] = {i < G.W - 1, j < G.H - 1, k < G.D - 1};
    const float ux = u[t], uy = u[n + t], uz = u[2 * n + t];
    float vn[3], vo[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      vn[c] = vnew[c * n + t];
      vo[c] = vold[c * n + t];
    }
    const float shrink_den = lambda + 2.0f * sigma1;
    float term = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float kx = has[a] ? 1.0f * (vnew[c * n + t + off[a]] - vn[c]) : 0.0f;
        const float ko =
            mode == 0 ? 0.0f : (has[a] ? 1.0f * (vold[c * n + t + off[a]] - vo[c]) : 0.0f);
        const size_t p = (size_t)(3 * c + a) * n + t;
        const float yo = y[p];
        const float yt = yo + sigma1 * (kx + kx - ko);
        const float yn = yt * lambda / shrink_den;
        yout[p] = yn;
        if (resid) term += fabsf((yo - yn) / sigma1 - ko + kx);
      }
    }
    const float k4 = ux * vn[0] + uy * vn[1] + uz * vn[2];
    const float k4o =
        mode == 0 ? 0.0f : (mode == 1 ? kx4prev[t] : ux * vo[0] + uy * vo[1] + uz * vo[2]);
    const float s2 = 1.0f / (fabsf(ux) + fabsf(uy) + fabsf(uz));
    const float y4o = y[9 * n + t];
    const float y4t = y4o + s2 * (k4 + k4 - k4o + ut[t]);
    const float m1 = (y4t < 1.0f) ? y4t : 1.0f;
    const float y4n = (-1.0f < m1) ? m1 : -1.0f;
    yout[9 * n + t] = y4n;
    if (resid) {
      term += fabsf((y4o - y4n) / s2 - k4o + k4);
      dterm = (double)term;
    }
  }
  if (resid) {
    const double s = block_sum<256>(dterm);
    if (threadIdx.x == 0) dpart[blockIdx.x] = s;
  }
}

// err = (d*d + p*p)/nPx in float from the float-rounded sums; stop when !(err > tol)
__global__ __launch_bounds__(256) void k_flow_resid(const double* __restrict__ ppart, int np,
                                                    const double* __restrict__ dpart, int nd,
                                                    int npx, float tol, int* __restrict__ ctl,
                                                    float* __restrict__ errout) {
  if (ctl[0] == 0) return;
  double ps = 0.0, ds = 0.0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) ps += ppart[q];
  for (int q = threadIdx.x; q < nd; q += blockDim.x) ds += dpart[q];
  const double P = block_sum<256>(ps);
  __syncthreads();
  const double Dd = block_sum<256>(ds);
  if (threadIdx.x == 0) {
    const float p = (float)P, d = (float)Dd;
    const float err = (d * d + p * p) / (float)npx;
    if (errout) *errout = err;
    if (!(err > tol)) ctl[0] = 0;
  }
}

// After a warp: the host alternated the v and y buffers for every queued iteration;
// if the device stopped early with the other parity, move the live state into the
// buffers the host treats as current.
__global__ void k_flow_settle(size_t n, int host_last, float* __restrict__ v_host,
                              const float* __restrict__ v_other, float* __restrict__ y_host,
                              const float* __restrict__ y_other, const int* __restrict__ ctl) {
  const int dev_last = ctl[1];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (!((dev_last ^ host_last) & 1) || t >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) v_host[c * n + t] = v_other[c * n + t];
#pragma unroll
  for (int c = 0; c < 10; ++c) y_host[c * n + t] = y_other[c * n + t];
}

__global__ void k_zero_tail(double* __restrict__ p, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

__global__ void k_flow_arm(int* ctl) {
  ctl[0] = 1;
  ctl[1] = 0;
}

}  // namespace

static FlowGrid mkgrid(const int* dims, const int* roi) {
  FlowGrid G;
  G.W = dims[0];
  G.H = dims[1];
  G.D = dims[2];
  G.r0 = roi[0];
  G.r1 = roi[1];
  G.r2 = roi[2];
  G.r3 = roi[3];
  G.r4 = roi[4];
  G.r5 = roi[5];
  return G;
}

static size_t roi_count(const int* dims, const int* roi) {
  const long bx = dims[0] - roi[0] - roi[1], by = dims[1] - roi[2] - roi[3],
             bz = dims[2] - roi[4] - roi[5];
  if (bx <= 0 || by <= 0 || bz <= 0) return 0;
  return (size_t)bx * by * bz;
}

// Tuning variants of one iteration: 0 = two kernels (primal, dual); 1..5 = the fused
// z-marching kernel with tiles (TX, TY, ZC) below. Narrow grids (W < 64) use 16x16
// tiles with the same ZC.
struct PdShape {
  int tx, ty, zc;
};
static const PdShape kShapes[6] = {{0, 0, 0}, {64, 4, 2}, {64, 4, 4}, {64, 4, 8},
                                   {32, 8, 4}, {64, 4, 16}};

static PdShape shape_of(int variant, const int* dims) {
  PdShape p = kShapes[(variant >= 1 && variant <= 5) ? variant : 2];
  if (dims[0] < 64) {
    p.tx = 16;
    p.ty = 16;
  }
  return p;
}

size_t wapk_flow_blocks(int variant, const int* dims) {
  if (variant == 0) return ((size_t)dims[0] * dims[1] * dims[2] + 255) / 256;
  const PdShape p = shape_of(variant, dims);
  return (size_t)((dims[0] + p.tx - 1) / p.tx) * ((dims[1] + p.ty - 1) / p.ty) *
         ((dims[2] + p.zc - 1) / p.zc);
}

hipError_t wapk_flow_prep(const int* dims, const int* roi, const float* warped, const float* I1,
                          const float* v, float* u, float* ut, hipStream_t s) {
  const size_t nb = roi_count(dims, roi);
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_flow_prep, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s,
                     mkgrid(dims, roi), warped, I1, v, u, ut);
  return hipGetLastError();
}

hipError_t wapk_flow_kx4(const int* dims, const float* u, const float* v, float* kx4,
                         hipStream_t s) {
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(k_flow_kx4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, u, v,
                     kx4);
  return hipGetLastError();
}

#define WAP_PD_LAUNCH(TX, TY, ZC)                                                        \
  hipLaunchKernelGGL((k_flow_pd<TX, TY, ZC>), dim3(nb), dim3(256), 0, s, G, vold, vnew, yold, \
                     ynew, u, ut, kx4prev, kty, mode, resid, store_kty, sigma1, lambda, ppart,  \
                     dpart, ctl, iter)

hipError_t wapk_flow_pd(int variant, const int* dims, const int* roi, const float* vold,
                        float* vnew, const float* yold, float* ynew, const float* u,
                        const float* ut, const float* kx4prev, float* kty, int mode, int resid,
                        int store_kty, float sigma1, float lambda, double* ppart, double* dpart,
                        int* ctl, int iter, hipStream_t s) {
  const FlowGrid G = mkgrid(dims, roi);
  const unsigned nb = (unsigned)wapk_flow_blocks(variant, dims);
  if (variant == 0) {
    const size_t nr = roi_count(dims, roi);
    if (nr)
      hipLaunchKernelGGL(k_flow_primal, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, G,
                         vold, vnew, yold, u, kty, store_kty, resid, ppart, ctl);
    if (resid) {
      // the primal sums sit in the first ceil(nr/256) slots; clear the rest
      const size_t used = (nr + 255) / 256;
      if (nb > used)
        hipLaunchKernelGGL(k_zero_tail, dim3((unsigned)((nb - used + 255) / 256)), dim3(256), 0,
                           s, ppart + used, nb - used);
    }
    hipLaunchKernelGGL(k_flow_dual, dim3(nb), dim3(256), 0, s, G, vold, vnew, yold, u, ut,
                       kx4prev, mode, resid, sigma1, lambda, dpart, ctl, iter, ynew);
    return hipGetLastError();
  }
  const PdShape p = shape_of(variant, dims);
  if (p.tx == 16) {
    if (p.zc == 2) WAP_PD_LAUNCH(16, 16, 2);
    else if (p.zc == 4) WAP_PD_LAUNCH(16, 16, 4);
    else if (p.zc == 8) WAP_PD_LAUNCH(16, 16, 8);
    else WAP_PD_LAUNCH(16, 16, 16);
  } else if (p.tx == 32) {
    WAP_PD_LAUNCH(32, 8, 4);
  } else {
    if (p.zc == 2) WAP_PD_LAUNCH(64, 4, 2);
    else if (p.zc == 4) WAP_PD_LAUNCH(64, 4, 4);
    else if (p.zc == 8) WAP_PD_LAUNCH(64, 4, 8);
    else WAP_PD_LAUNCH(64, 4, 16);
  }
  return hipGetLastError();
}

hipError_t wapk_flow_resid(const double* ppart, int np, const double* dpart, int nd, int npx,
                           float tol, int* ctl, float* errout, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_resid, dim3(1), dim3(256), 0, s, ppart, np, dpart, nd, npx, tol, ctl,
                     errout);
  return hipGetLastError();
}

hipError_t wapk_flow_settle(size_t n, int host_last, float* v_host, const float* v_other,
                            float* y_host, const float* y_other, const int* ctl, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_settle, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                     host_last, v_host, v_other, y_host, y_other, ctl);
  return hipGetLastError();
}

hipError_t wapk_flow_arm(int* ctl, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_arm, dim3(1), dim3(1), 0, s, ctl);
  return hipGetLastError();
}
