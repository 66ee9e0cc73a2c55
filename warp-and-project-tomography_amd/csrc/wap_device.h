// Device-side helpers shared by the gfx950 kernels of the warp-and-project path.
//
// Every kernel file is compiled with -ffp-contract=off and correctly rounded f32
// division/sqrt, so a*b+c is evaluated as two rounded operations exactly like the
// reference's x86 float code; this is what makes ray sample sets, advection clamps and
// the flow iteration reproduce the reference bit for bit instead of to a tolerance.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wap_types.h"

#define WAP_DEV __device__ __forceinline__

// Linear voxel index of (x, y, z) in a planar x-fastest W*H*D grid.
WAP_DEV size_t vidx(int x, int y, int z, int W, int H) {
  return (size_t)x + (size_t)W * ((size_t)y + (size_t)H * (size_t)z);
}

WAP_DEV int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The reference's min/max macros ((a)<(b)?(a):(b)) / ((a)>(b)?(a):(b)): their NaN
// behaviour (the second operand wins) matters for ray entry/exit parameters.
WAP_DEV float mmin(float a, float b) { return (a < b) ? a : b; }
WAP_DEV float mmax(float a, float b) { return (a > b) ? a : b; }

// Column-major 4x4 times (x, y, z, w), summed left to right as Matrix4x4 * Vector4.
WAP_DEV void mat4_apply(const float* M, float x, float y, float z, float w, float* o) {
#pragma unroll
  for (int r = 0; r < 4; ++r) o[r] = M[r] * x + M[4 + r] * y + M[8 + r] * z + M[12 + r] * w;
}

// Map a linear block id so that consecutive logical tiles land on the same XCD (the
// hardware deals workgroups round-robin over the 8 XCDs; neighbouring tiles share
// volume data and should share an L2).
WAP_DEV unsigned xcd_remap(unsigned bid, unsigned nblocks) {
  const unsigned nx = 8;
  const unsigned per = nblocks / nx;
  if (per == 0 || bid >= per * nx) return bid;
  return (bid % nx) * per + bid / nx;
}

// Volume fetch with clamp-to-edge indices.
WAP_DEV float vfetch(const float* __restrict__ vol, int x, int y, int z, int W, int H, int D) {
  x = iclamp(x, 0, W - 1);
  y = iclamp(y, 0, H - 1);
  z = iclamp(z, 0, D - 1);
  return vol[vidx(x, y, z, W, H)];
}

// Storage element types of the precision form: values are computed in float and stored
// as float, IEEE half (round to nearest even) or bfloat16 (round to nearest even, NaN
// kept quiet).
struct Bf16 {
  unsigned short b;
};
WAP_DEV float to_f(float x) { return x; }
WAP_DEV float to_f(_Float16 x) { return (float)x; }
WAP_DEV float to_f(Bf16 x) { return __uint_as_float((unsigned)x.b << 16); }
template <class S>
WAP_DEV S from_f(float v);
template <>
WAP_DEV float from_f<float>(float v) {
  return v;
}
template <>
WAP_DEV _Float16 from_f<_Float16>(float v) {
  return (_Float16)v;
}
template <>
WAP_DEV Bf16 from_f<Bf16>(float v) {
  unsigned u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return Bf16{(unsigned short)((u >> 16) | 0x40u)};
  u += 0x7fffu + ((u >> 16) & 1u);
  return Bf16{(unsigned short)(u >> 16)};
}

// Element access at a float-scaled byte offset t4 (= 4 x element index) from a uniform
// plane base: the offset of an S element is t4 * sizeof(S) / 4.
template <class S>
WAP_DEV float lds(const S* __restrict__ base, unsigned t4) {
  return to_f(*(const S*)((const char*)base + (sizeof(S) == 4 ? t4 : t4 >> 1)));
}
template <class S>
WAP_DEV void sts(S* __restrict__ base, unsigned t4, float v) {
  *(S*)((char*)base + (sizeof(S) == 4 ? t4 : t4 >> 1)) = from_f<S>(v);
}
