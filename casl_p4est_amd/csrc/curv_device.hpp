// Per-sample bodies of the learned-curvature input normalisation and hybrid
// blend, shared by the separate kernels (curv.hip, curv_grid.hip) and the
// fused network kernel (mlp.hip), so both give the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include "stencil_math.hpp"

namespace casl {
namespace dev {

// Network inputs of a 2D sample (kml::Curvature::_computeHybridHK,
// src/my_p4est_curvature_ml.cpp:483-520): negate every feature when the
// numerical h*kappa is positive, rotate the stencil by multiples of 90 degrees
// so that the gradient at the centre points into the first quadrant
// (rotateStencilToFirstQuadrant, :373-398; rotateStencil90, :327-353; a zero
// gradient component counts as +EPS), and emit it twice: as is (o[0..9]) and
// reflected about y = x (o[10..19], reflectStencil_yEqx :356-370).  The
// gradient is the central difference of the phi stencil (the reference
// carries normal stencils for it; only their direction matters here).
__device__ __forceinline__ void curv_normalize2(const float *__restrict__ in, float *__restrict__ o) {
  float f[10], g[9];
#pragma unroll
  for (int q = 0; q < 10; ++q) f[q] = in[q];
  const float sg = f[9] > 0.f ? -1.f : 1.f;
#pragma unroll
  for (int q = 0; q < 10; ++q) f[q] *= sg;
  float gx = f[7] - f[1], gy = f[5] - f[3];
  if (gx == 0.f) gx = float(EPS);
  if (gy == 0.f) gy = float(EPS);
  // index maps of the rotated stencil: new[q] = old[map[q]]
  constexpr int CCW[9] = {2, 5, 8, 1, 4, 7, 0, 3, 6};  // +90 degrees
  constexpr int CW[9] = {6, 3, 0, 7, 4, 1, 8, 5, 2};   // -90 degrees
  constexpr int HALF[9] = {8, 7, 6, 5, 4, 3, 2, 1, 0};  // 180 degrees
  const int quad = gy > 0.f ? (gx > 0.f ? 1 : 2) : (gx < 0.f ? 3 : 4);
#pragma unroll
  for (int q = 0; q < 9; ++q) g[q] = f[q];
  if (quad == 2) {
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = g[CW[q]];
  } else if (quad == 3) {
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = g[HALF[q]];
  } else if (quad == 4) {
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = g[CCW[q]];
  }
#pragma unroll
  for (int q = 0; q < 10; ++q) o[q] = f[q];
  constexpr int REFL[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};  // swap 1-3, 2-6, 5-7
#pragma unroll
  for (int q = 0; q < 9; ++q) o[10 + q] = f[REFL[q]];
  o[19] = f[9];
}

// Hybrid h*kappa of a 2D sample (src/my_p4est_curvature_ml.cpp:483-556): the
// numerical value where |ihk| < lo; elsewhere the mean of the two network
// outputs (plus the normalised numerical value -|ihk| when the network
// predicts a correction), blended linearly towards -|ihk| for |ihk| <= hi,
// and given the sign of ihk.
__device__ __forceinline__ float curv_hybrid2(float ihk, float p0, float p1, float lo, float hi, int residual) {
  const float a = fabsf(ihk);
  if (a < lo) return ihk;
  float p = 0.5f * (p0 + p1);
  if (residual) p += -a;
  if (a <= hi) {
    const float lam = (hi - a) / (hi - lo);
    p = (1.f - lam) * p + lam * -a;
  }
  return ihk > 0.f ? fabsf(p) : -fabsf(p);
}

// position of offsets (i, j, k) in {-1, 0, 1}^3, x slowest
__host__ __device__ constexpr int curv_q3(int i, int j, int k) { return 9 * (i + 1) + 3 * (j + 1) + (k + 1); }

// 3D network inputs of a sample ([28] -> [28]): negate when ihk > 0, reflect
// every axis along which the centred difference of phi is negative, order the
// axes by decreasing difference (ties to the lower axis); then the same 27
// values in a basis where curvature is not buried under the linear part (BF16
// keeps 8 bits of each input): the centre value, the three half centred
// differences g at the -e_a slots, and everywhere else the deviation from the
// plane f0 + o.g.
__device__ __forceinline__ void curv_normalize3(const float *__restrict__ in, float *__restrict__ o) {
  float f[28];
#pragma unroll
  for (int q = 0; q < 28; ++q) f[q] = in[q];
  const float sg = f[27] > 0.f ? -1.f : 1.f;
#pragma unroll
  for (int q = 0; q < 28; ++q) f[q] *= sg;
  float g[3] = {f[curv_q3(1, 0, 0)] - f[curv_q3(-1, 0, 0)], f[curv_q3(0, 1, 0)] - f[curv_q3(0, -1, 0)],
                f[curv_q3(0, 0, 1)] - f[curv_q3(0, 0, -1)]};
  int r[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (g[a] == 0.f) g[a] = float(EPS);
    r[a] = g[a] < 0.f ? -1 : 1;
    g[a] = fabsf(g[a]);
  }
  // axes by decreasing |difference|, ties to the lower axis
  int p[3] = {0, 1, 2};
  if (g[p[1]] > g[p[0]]) {
    const int x = p[0];
    p[0] = p[1];
    p[1] = x;
  }
  if (g[p[2]] > g[p[1]]) {
    const int x = p[1];
    p[1] = p[2];
    p[2] = x;
    if (g[p[1]] > g[p[0]]) {
      const int y = p[0];
      p[0] = p[1];
      p[1] = y;
    }
  }
  float v[27];
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        // new axis b reads old axis p[b]
        int c[3];
        c[p[0]] = i * r[p[0]];
        c[p[1]] = j * r[p[1]];
        c[p[2]] = k * r[p[2]];
        v[curv_q3(i, j, k)] = f[curv_q3(c[0], c[1], c[2])];
      }
  const float f0 = v[curv_q3(0, 0, 0)];
  const float gx = 0.5f * (v[curv_q3(1, 0, 0)] - v[curv_q3(-1, 0, 0)]);
  const float gy = 0.5f * (v[curv_q3(0, 1, 0)] - v[curv_q3(0, -1, 0)]);
  const float gz = 0.5f * (v[curv_q3(0, 0, 1)] - v[curv_q3(0, 0, -1)]);
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        const int q = curv_q3(i, j, k);
        o[q] = v[q] - (f0 + (float(i) * gx + float(j) * gy + float(k) * gz));
      }
  o[curv_q3(0, 0, 0)] = f0;
  o[curv_q3(-1, 0, 0)] = gx;
  o[curv_q3(0, -1, 0)] = gy;
  o[curv_q3(0, 0, -1)] = gz;
  o[27] = f[27];
}

// Hybrid h*kappa of a 3D sample: one network output per sample.  With a
// residual network, corrections no larger than `dead` are dropped: on a
// well-resolved surface the numerical value is more accurate than the
// network's noise floor.
__device__ __forceinline__ float curv_hybrid3(float ihk, float pred, float lo, float hi, int residual, float dead) {
  const float a = fabsf(ihk);
  if (a < lo) return ihk;
  float p = pred;
  if (residual) p = (fabsf(p) > dead ? p : 0.f) + -a;
  if (a <= hi) {
    const float lam = (hi - a) / (hi - lo);
    p = (1.f - lam) * p + lam * -a;
  }
  return ihk > 0.f ? fabsf(p) : -fabsf(p);
}

}  // namespace dev
}  // namespace casl
