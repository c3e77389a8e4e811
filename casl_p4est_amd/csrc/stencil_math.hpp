// Per-node stencil arithmetic shared by the reinitialization kernels.
//
// Every function restates one piece of the reference in its own operation
// order (compile with -ffp-contract=off so FP64 results are bit-identical to
// a strict IEEE evaluation of the reference expressions).  Paths are
// relative to the reference tree.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace casl {
namespace dev {

constexpr double EPS = 1e-13;  // src/casl_math.h:23

// Device view of the flattened stencils of casl_flat_stencil.  Distances are
// stored as DT: FP32 (exact for dyadic trees), or FP64 for the FP64 kernels
// of a grid whose distances FP32 does not hold exactly.
template <int D, class DT = float>
struct Tables {
  static constexpr bool REGULAR = false;  // see the REGULAR shortcuts below
  static constexpr bool FASTDIV = false;  // see second_difference
  static constexpr bool SLOTS = false;    // see the hanging-face functions below
  static constexpr bool UNIFORM = false;  // see second_difference
  static constexpr int F = 2 * D;
  static constexpr int C = 1 << (D - 1);
  static constexpr int T = 2 * (D - 1);
  const int32_t *__restrict__ nbr;  // [F*C][N]
  const DT *__restrict__ dist;      // [F][N] face distances, then [F*T][N] transverse
  const uint8_t *__restrict__ wall; // [N]
  int64_t N;

  __device__ __forceinline__ int32_t idx(int f, int c, int64_t n) const { return nbr[(int64_t(f) * C + c) * N + n]; }
  __device__ __forceinline__ DT d(int f, int64_t n) const { return dist[int64_t(f) * N + n]; }
  __device__ __forceinline__ DT dt(int f, int k, int64_t n) const {
    return dist[(int64_t(F) + int64_t(f) * T + k) * N + n];
  }
};

// MINMOD, src/casl_math.h:80-88
template <class R>
__device__ __forceinline__ R minmod(R a, R b) {
  if (a * b <= R(0)) return R(0);
  if (fabs(a) < fabs(b)) return a;
  return b;
}
// FP32 (the mixed path only): the median of a, b and 0, one instruction.  It
// is MINMOD's value except where a*b underflows to zero (|a*b| < 1e-38),
// there it keeps the smaller of two values below 1e-19 instead of 0.
__device__ __forceinline__ float minmod(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }
template <class R>
__host__ __device__ __forceinline__ R min2(R a, R b) { return a < b ? a : b; }
template <class R>
__host__ __device__ __forceinline__ R min3(R a, R b, R c) { return a < b ? min2(a, c) : min2(b, c); }
template <class R>
__host__ __device__ __forceinline__ R max2(R a, R b) { return b < a ? a : b; }

// interface_Location, src/casl_math.cpp:55-59
__device__ __forceinline__ double interface_location(double a, double b, double fa, double fb) {
  return 0.5 * (a + b + (a - b) * (fa + fb) / (fb - fa));
}

// interface_Location_With_Second_Order_Derivative, src/casl_math.cpp:67-97
__device__ __forceinline__ double interface_location_2nd(double a, double b, double fa, double fb, double fxxa,
                                                         double fxxb) {
  if (fabs(fa) < EPS) return a + EPS;
  if (fabs(fb) < EPS) return b - EPS;
  double fxx = minmod(fxxa, fxxb);
  double h = b - a;
  double c2 = 0.5 * fxx;
  double c1 = (fb - fa) / h;
  double c0 = 0.5 * (fb + fa) - h * h / 8 * fxx;
  double x;
  if (fabs(c2) < EPS)
    x = -c0 / c1;
  else if (fb < 0)
    x = (-2 * c0) / (c1 - sqrt(c1 * c1 - 4 * c2 * c0));
  else
    x = (-2 * c0) / (c1 + sqrt(c1 * c1 - 4 * c2 * c0));
  return x + 0.5 * (a + b);
}

// x / y.  FP64 divides (the path that reproduces the reference bit for bit);
// FP32 -- the mixed path, held to 1e-3 h of FP64 -- multiplies by the
// hardware reciprocal (about 1 ulp), two instructions instead of a range-safe
// division sequence.
template <class R>
__device__ __forceinline__ R qdiv(R x, R y) {
  return x / y;
}
template <>
__device__ __forceinline__ float qdiv<float>(float x, float y) {
  return x * __builtin_amdgcn_rcpf(y);
}

// sqrt: correctly rounded in FP64; the hardware square root in FP32 (the
// mixed path), without the denormal rescaling around it.
template <class R>
__device__ __forceinline__ R qsqrt(R x) {
  return sqrt(x);
}
template <>
__device__ __forceinline__ float qsqrt<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// a * b + c: one fused operation in FP32 (the mixed path); FP64 keeps the
// separately rounded product and sum of the reference expression.
template <class R>
__device__ __forceinline__ R qfma(R a, R b, R c) {
  return a * b + c;
}
template <>
__device__ __forceinline__ float qfma<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// Exact-inverse division.  A stencil accessor with TB::FASTDIV and t.fast
// has power-of-two face distances d_f = 1 / iq[f] (dyadic cells of a tree
// whose edges are powers of two, no wall arm).  Dividing by a power of two and
// multiplying by its exactly representable inverse give the same correctly
// rounded result, so these helpers return bit-for-bit what the plain
// division returns, without its instruction sequence.
// The helpers below serve the hanging-arm arithmetic, where the test on
// t.fast is a per-lane branch: it pays off against FP64 division sequences
// and costs more than FP32's short approximate division, so they apply the
// exact inverses in FP64 only (in FP32 the branches measured 1.84x slower
// for the whole stage kernel, from the added register pressure).
template <class R, class TB>
constexpr bool irregular_fastdiv() {
  return TB::FASTDIV && sizeof(R) == 8;
}

template <class R, class TB>
__device__ __forceinline__ R div_d(const TB &t, R x, int f, int64_t n) {
  if constexpr (irregular_fastdiv<R, TB>()) {
    if (t.fast) return x * t.iq[f];
  }
  return qdiv(x, R(t.d(f, n)));
}
// x / (d_{2a+1} + d_{2a}): a power of two when both arms are equal
template <class R, class TB>
__device__ __forceinline__ R div_dsum(const TB &t, R x, int a, int64_t n) {
  if constexpr (irregular_fastdiv<R, TB>()) {
    if (t.fast && t.dist[2 * a + 1] == t.dist[2 * a]) return x * (R(0.5) * t.iq[2 * a + 1]);
  }
  return qdiv(x, R(t.d(2 * a + 1, n)) + R(t.d(2 * a, n)));
}

// get_linear_interpolation_weights + interpolate_linearly,
// my_p4est_quad_neighbor_nodes_of_node.h:222-306
// TB is any stencil accessor with idx(f, c, n), d(f, n) and dt(f, k, n):
// Tables<D> (global SoA tables) or a decoded per-node stencil (LDS-local
// indices, see tiles_device.hpp).
//
// TB::REGULAR marks a node none of whose arms hangs on a coarser cell: every
// face then has exactly one corner c with a nonzero weight w, the reference's
// (sum_c f_c w_c)/sum-of-weights is (f_c w)/w = f_c exactly (w and the
// normalisation are equal powers of two), every m_ij of
// correct_naive_second_derivatives is 0 and every quadratic correction is
// 0.5*0*f'' = 0.  The REGULAR shortcuts therefore return the same values as
// the full expressions for finite inputs (at most the sign of a zero differs).
template <int D, class R, class G, class TB>
__device__ __forceinline__ R face_linear(const TB &t, int f, int64_t n, const G *__restrict__ g) {
  const R d_m0 = R(t.dt(f, 0, n)), d_p0 = R(t.dt(f, 1, n));
  if constexpr (TB::REGULAR) {
    return R(g[t.idx(f, 0, n)]);
  } else if constexpr (D == 2) {
    const R w_mm = d_p0;
    const R w_pm = d_m0;
    const R normalization = (d_m0 + d_p0);
    const R num = qfma(R(g[t.idx(f, 0, n)]), w_mm, R(g[t.idx(f, 1, n)]) * w_pm);
    // d_m0 + d_p0 is the arm cell's transverse size, a power of two on a
    // dyadic geometry: its exact inverse (inv_norm) replaces the division
    if constexpr (irregular_fastdiv<R, TB>()) {
      if (t.fast) return num * t.inv_norm(f);
    }
    return qdiv(num, normalization);
  } else {
    const R d_0m = R(t.dt(f, 2, n)), d_0p = R(t.dt(f, 3, n));
    const R w_mm = d_p0 * d_0p;
    const R w_pm = d_m0 * d_0p;
    const R w_mp = d_p0 * d_0m;
    const R w_pp = d_m0 * d_0m;
    const R normalization = (d_m0 + d_p0) * (d_0m + d_0p);
    const R num = qfma(R(g[t.idx(f, 0, n)]), w_mm,
                       qfma(R(g[t.idx(f, 2, n)]), w_mp, qfma(R(g[t.idx(f, 1, n)]), w_pm, R(g[t.idx(f, 3, n)]) * w_pp)));
    // the transverse sums are the arm cell's transverse sizes, powers of two
    // on a dyadic geometry (inv_norm: the inverse of their product)
    if constexpr (irregular_fastdiv<R, TB>()) {
      if (t.fast) return num * t.inv_norm(f);
    }
    return qdiv(num, normalization);
  }
}

// central_second_derivative, my_p4est_quad_neighbor_nodes_of_node.h:214-220
template <class R>
__device__ __forceinline__ R central_second_derivative(R fp, R f0, R fm, R dp, R dm) {
  return qdiv((qdiv(fp - f0, dp) + qdiv(fm - f0, dm)) * R(2.), dp + dm);
}

// central_second_derivative along axis a of a stencil.  When the accessor
// knows its face distances are powers of two (TB::FASTDIV and t.fast), the
// divisions by d_p and d_m become multiplications by their exact inverses and
// the division by d_p + d_m one by 1/(2 d_p) when d_p == d_m: the same
// correctly rounded results, without the division sequences.
//
// TB::UNIFORM: both arms of axis a have the same power-of-two length q (a
// uniform node of the tiled encoding on a dyadic geometry).  Then
// (fp - f0) * (1/q) and (fm - f0) * (1/q) are exact scalings, their correctly
// rounded sum is the rounded sum of the differences times 1/q, and the
// factors 2 and 1/(2q) are exact too: ((fp - f0) + (fm - f0)) * (1/q)^2 gives
// the same bits as the path below, in FP64 and (with its fused first product)
// in FP32.
template <class R, class TB>
__device__ __forceinline__ R second_difference(const TB &t, int a, R fp, R f0, R fm, int64_t n) {
  if constexpr (TB::UNIFORM) return ((fp - f0) + (fm - f0)) * t.iq2[a];
  const int fpp = 2 * a + 1, fmm = 2 * a;
  if constexpr (TB::FASTDIV) {
    if (t.fast) {
      const R x = qfma(fp - f0, t.iq[fpp], (fm - f0) * t.iq[fmm]) * R(2.);
      return t.dist[fpp] == t.dist[fmm] ? x * (R(0.5) * t.iq[fpp]) : qdiv(x, t.dist[fpp] + t.dist[fmm]);
    }
  }
  return central_second_derivative(fp, f0, fm, R(t.d(fpp, n)), R(t.d(fmm, n)));
}

template <class R>
__device__ __forceinline__ bool check_if_zero(R v) { return fabs(v) < R(EPS); }

// correct_naive_second_derivatives, my_p4est_quad_neighbor_nodes_of_node.cpp:44-298:
// the solve, given the m_ij (m13 .. m32 unused in 2D)
template <int D, class R>
__device__ __forceinline__ void correct_naive_solve(R m12, R m21, R m13, R m23, R m31, R m32, R &Dxx, R &Dyy,
                                                    R &Dzz) {
  if constexpr (D == 2) {
    if constexpr (sizeof(R) == 4) {  // FP32: the general inverse, as in 3D below
      const R idet = qdiv(R(1.0), R(1.0) - m12 * m21);
      const R nDxx = Dxx, nDyy = Dyy;
      Dxx = qfma(-nDyy, m12 * idet, nDxx * idet);
      Dyy = qfma(-nDxx, m21 * idet, nDyy * idet);
      return;
    }
    if (check_if_zero(m12)) {
      if (!check_if_zero(m21)) Dyy -= m21 * Dxx;
    } else if (check_if_zero(m21)) {
      Dxx -= m12 * Dyy;
    } else {
      const R idet = qdiv(R(1.0), R(1.0) - m12 * m21);
      const R nDxx = Dxx, nDyy = Dyy;
      Dxx = nDxx * idet - nDyy * (m12 * idet);
      Dyy = nDyy * idet - nDxx * (m21 * idet);
    }
  } else {
    if constexpr (sizeof(R) == 4) {
      // FP32 (the mixed path): the general inverse for every case.  The m_ij
      // are ratios of dyadic lengths, so each is exactly 0 or far above EPS,
      // and the reference's zero tests only pick which terms vanish; without
      // the branch tree the lanes of a wave no longer split over its cases.
      const R idet =
          qdiv(R(1.0), R(1.0) + m12 * m23 * m31 + m13 * m21 * m32 - m12 * m21 - m13 * m31 - m23 * m32);
      const R nDxx = Dxx, nDyy = Dyy, nDzz = Dzz;
      Dxx = qfma(nDzz, idet * (m12 * m23 - m13), qfma(nDyy, idet * (m32 * m13 - m12), nDxx * (idet * (R(1.0) - m23 * m32))));
      Dyy = qfma(nDzz, idet * (m21 * m13 - m23), qfma(nDyy, idet * (R(1.0) - m31 * m13), nDxx * (idet * (m31 * m23 - m21))));
      Dzz = qfma(nDzz, idet * (R(1.0) - m21 * m12), qfma(nDyy, idet * (m31 * m12 - m32), nDxx * (idet * (m21 * m32 - m31))));
      return;
    }
    if (check_if_zero(m12) && check_if_zero(m13)) {
      if (check_if_zero(m23)) {
        const bool z21 = !check_if_zero(m21), z31 = !check_if_zero(m31), z32 = !check_if_zero(m32);
        if (z21) Dyy -= m21 * Dxx;
        if (z31) Dzz -= m31 * Dxx;
        if (z32) Dzz -= m32 * Dyy;
      } else if (check_if_zero(m32)) {
        const bool z21 = !check_if_zero(m21), z31 = !check_if_zero(m31);
        if (z31) Dzz -= m31 * Dxx;
        Dyy -= m23 * Dzz;
        if (z21) Dyy -= m21 * Dxx;
      } else {
        const R idet = qdiv(R(1.0), R(1.0) - m23 * m32);
        const bool i21 = check_if_zero(m21), i31 = check_if_zero(m31);
        const R nDyy = Dyy, nDzz = Dzz;
        Dyy = ((i21 ? nDyy : (nDyy - m21 * Dxx)) - m23 * (i31 ? nDzz : (nDzz - m31 * Dxx))) * idet;
        Dzz = ((i31 ? nDzz : (nDzz - m31 * Dxx)) - m32 * (i21 ? nDyy : (nDyy - m21 * Dxx))) * idet;
      }
    } else if (check_if_zero(m21) && check_if_zero(m23)) {
      if (check_if_zero(m13)) {
        const bool z12 = !check_if_zero(m12), z31 = !check_if_zero(m31), z32 = !check_if_zero(m32);
        if (z12) Dxx -= m12 * Dyy;
        if (z31) Dzz -= m31 * Dxx;
        if (z32) Dzz -= m32 * Dyy;
      } else if (check_if_zero(m31)) {
        const bool z12 = !check_if_zero(m12), z32 = !check_if_zero(m32);
        if (z32) Dzz -= m32 * Dyy;
        Dxx -= m13 * Dzz;
        if (z12) Dxx -= m12 * Dyy;
      } else {
        const R idet = qdiv(R(1.0), R(1.0) - m13 * m31);
        const bool i12 = check_if_zero(m12), i32 = check_if_zero(m32);
        const R nDxx = Dxx, nDzz = Dzz;
        Dxx = ((i12 ? nDxx : (nDxx - m12 * Dyy)) - m13 * (i32 ? nDzz : (nDzz - m32 * Dyy))) * idet;
        Dzz = ((i32 ? nDzz : (nDzz - m32 * Dyy)) - m31 * (i12 ? nDxx : (nDxx - m12 * Dyy))) * idet;
      }
    } else if (check_if_zero(m31) && check_if_zero(m32)) {
      if (check_if_zero(m12)) {
        const bool z13 = !check_if_zero(m13), z21 = !check_if_zero(m21), z23 = !check_if_zero(m23);
        if (z13) Dxx -= m13 * Dzz;
        if (z21) Dyy -= m21 * Dxx;
        if (z23) Dyy -= m23 * Dzz;
      } else if (check_if_zero(m21)) {
        const bool z13 = !check_if_zero(m13), z23 = !check_if_zero(m23);
        if (z23) Dyy -= m23 * Dzz;
        Dxx -= m12 * Dyy;
        if (z13) Dxx -= m13 * Dzz;
      } else {
        const R idet = qdiv(R(1.0), R(1.0) - m12 * m21);
        const bool i13 = check_if_zero(m13), i23 = check_if_zero(m23);
        const R nDxx = Dxx, nDyy = Dyy;
        Dxx = ((i13 ? nDxx : (nDxx - m13 * Dzz)) - m12 * (i23 ? nDyy : (nDyy - m23 * Dzz))) * idet;
        Dyy = ((i23 ? nDyy : (nDyy - m23 * Dzz)) - m21 * (i13 ? nDxx : (nDxx - m13 * Dzz))) * idet;
      }
    } else {
      const R idet =
          qdiv(R(1.0), R(1.0) + m12 * m23 * m31 + m13 * m21 * m32 - m12 * m21 - m13 * m31 - m23 * m32);
      const R nDxx = Dxx, nDyy = Dyy, nDzz = Dzz;
      Dxx = (nDxx * (idet * (R(1.0) - m23 * m32)) + nDyy * (idet * (m32 * m13 - m12)) +
             nDzz * (idet * (m12 * m23 - m13)));
      Dyy = (nDxx * (idet * (m31 * m23 - m21)) + nDyy * (idet * (R(1.0) - m31 * m13)) +
             nDzz * (idet * (m21 * m13 - m23)));
      Dzz = (nDxx * (idet * (m21 * m32 - m31)) + nDyy * (idet * (m31 * m12 - m32)) +
             nDzz * (idet * (R(1.0) - m21 * m12)));
    }
  }
}

// correct_naive_second_derivatives with the m_ij of a stencil's faces
template <int D, class R, class TB>
__device__ __forceinline__ void correct_naive(const TB &t, int64_t n, R &Dxx, R &Dyy, R &Dzz) {
  if constexpr (TB::REGULAR) return;
  const R m12 = div_dsum<R>(t, div_d<R>(t, R(t.dt(0, 0, n)) * R(t.dt(0, 1, n)), 0, n), 0, n) +
                div_dsum<R>(t, div_d<R>(t, R(t.dt(1, 0, n)) * R(t.dt(1, 1, n)), 1, n), 0, n);
  const R m21 = div_dsum<R>(t, div_d<R>(t, R(t.dt(2, 0, n)) * R(t.dt(2, 1, n)), 2, n), 1, n) +
                div_dsum<R>(t, div_d<R>(t, R(t.dt(3, 0, n)) * R(t.dt(3, 1, n)), 3, n), 1, n);
  if constexpr (D == 2) {
    correct_naive_solve<D, R>(m12, m21, R(0), R(0), R(0), R(0), Dxx, Dyy, Dzz);
  } else {
    const R m13 = div_dsum<R>(t, div_d<R>(t, R(t.dt(0, 2, n)) * R(t.dt(0, 3, n)), 0, n), 0, n) +
                  div_dsum<R>(t, div_d<R>(t, R(t.dt(1, 2, n)) * R(t.dt(1, 3, n)), 1, n), 0, n);
    const R m23 = div_dsum<R>(t, div_d<R>(t, R(t.dt(2, 2, n)) * R(t.dt(2, 3, n)), 2, n), 1, n) +
                  div_dsum<R>(t, div_d<R>(t, R(t.dt(3, 2, n)) * R(t.dt(3, 3, n)), 3, n), 1, n);
    const R m31 = div_dsum<R>(t, div_d<R>(t, R(t.dt(4, 0, n)) * R(t.dt(4, 1, n)), 4, n), 2, n) +
                  div_dsum<R>(t, div_d<R>(t, R(t.dt(5, 0, n)) * R(t.dt(5, 1, n)), 5, n), 2, n);
    const R m32 = div_dsum<R>(t, div_d<R>(t, R(t.dt(4, 2, n)) * R(t.dt(4, 3, n)), 4, n), 2, n) +
                  div_dsum<R>(t, div_d<R>(t, R(t.dt(5, 2, n)) * R(t.dt(5, 3, n)), 5, n), 2, n);
    correct_naive_solve<D, R>(m12, m21, m13, m23, m31, m32, Dxx, Dyy, Dzz);
  }
}

// ------------------------------------------------- hanging faces by slot
// A stencil with TB::SLOTS (a node of the tiled encoding with hanging arms)
// names the faces that hang: t.hface(q) for q < TB::NH (HANG_NONE past the
// last one), with corners t.hidx(q, c), face geometry sg = t.slot_geom(f)
// and transverse offsets t.hdt(q, k, sg) (k: m0, p0, 0m, 0p).  Each other
// face f is regular: one corner t.id[f] carries its value and its transverse
// products d_m0 d_p0, d_0m d_0p are 0, so it adds nothing to the m_ij of
// correct_naive_second_derivatives nor to the quadratic corrections.  The
// functions below do the hanging-face work for the hanging faces only; the
// terms they leave out are exact zeros (finite inputs), so their results are
// those of the per-face functions above, in FP64 bit for bit.

// interpolate_linearly on hanging face q (my_p4est_quad_neighbor_nodes_of_node.h:222-306).
// The transverse sums are the arm cell's transverse extents (sg.qt1, qt2), so on a
// dyadic geometry the division is a multiplication by their exact inverses.
template <int D, class R, class G, class TB, class SG>
__device__ __forceinline__ R hang_face_linear(const TB &t, int q, const SG &sg, const G *__restrict__ g) {
  const R d_m0 = t.hdt(q, 0, sg), d_p0 = t.hdt(q, 1, sg);
  if constexpr (D == 2) {
    const R num = qfma(R(g[t.hidx(q, 0)]), d_p0, R(g[t.hidx(q, 1)]) * d_m0);
    if (t.pow2) return num * sg.iqt1;
    return qdiv(num, d_m0 + d_p0);
  } else {
    const R d_0m = t.hdt(q, 2, sg), d_0p = t.hdt(q, 3, sg);
    const R w_mm = d_p0 * d_0p, w_pm = d_m0 * d_0p, w_mp = d_p0 * d_0m, w_pp = d_m0 * d_0m;
    const R num = qfma(R(g[t.hidx(q, 0)]), w_mm,
                       qfma(R(g[t.hidx(q, 2)]), w_mp, qfma(R(g[t.hidx(q, 1)]), w_pm, R(g[t.hidx(q, 3)]) * w_pp)));
    if (t.pow2) return num * (sg.iqt1 * sg.iqt2);
    return qdiv(num, (d_m0 + d_p0) * (d_0m + d_0p));
  }
}

// the quadratic correction of hanging face q (axis a), as quadratic_corrections
template <int D, class R, class TB, class SG>
__device__ __forceinline__ R hang_correction(const TB &t, int q, int a, const SG &sg, const R *dd) {
  const R h = R(0.5);
  if constexpr (D == 2) {
    return h * t.hdt(q, 0, sg) * t.hdt(q, 1, sg) * (a == 0 ? dd[1] : dd[0]);
  } else {
    const R dd1 = (a == 0) ? dd[1] : dd[0], dd2 = (a == 2) ? dd[1] : dd[2];
    return qfma(h * t.hdt(q, 0, sg) * t.hdt(q, 1, sg), dd1, h * t.hdt(q, 2, sg) * t.hdt(q, 3, sg) * dd2);
  }
}

// every face's value of g: its corner's on regular faces, the interpolation
// (minus the quadratic correction with the node's dd, CORR) on hanging ones;
// vq[q] keeps hanging face q's value
template <int D, class R, bool CORR, class G, class TB>
__device__ __forceinline__ void hang_face_values(const TB &t, const G *__restrict__ g, R *fnb, const R *dd, R *vq) {
  constexpr int F = 2 * D;
#pragma unroll
  for (int f = 0; f < F; ++f) fnb[f] = R(g[t.id[f]]);
#pragma unroll
  for (int q = 0; q < TB::NH; ++q) {
    const int f = t.hface(q);
    if (q == 0 || f < F) {
      const auto sg = t.slot_geom(f);
      R v = hang_face_linear<D, R>(t, q, sg, g);
      if constexpr (CORR) v -= hang_correction<D, R>(t, q, f >> 1, sg, dd);
      vq[q] = v;
#pragma unroll
      for (int i = 0; i < F; ++i) fnb[i] = (i == f) ? v : fnb[i];
    }
  }
}

// per face f, the value of component f >> 1 of dd (entries stride apart)
template <int D, class R, class G, class TB>
__device__ __forceinline__ void hang_face_values_dd(const TB &t, const G *__restrict__ dd, int64_t stride, R *out) {
  constexpr int F = 2 * D;
#pragma unroll
  for (int f = 0; f < F; ++f) out[f] = R(dd[int64_t(f >> 1) * stride + t.id[f]]);
#pragma unroll
  for (int q = 0; q < TB::NH; ++q) {
    const int f = t.hface(q);
    if (q == 0 || f < F) {
      const auto sg = t.slot_geom(f);
      const R v = hang_face_linear<D, R>(t, q, sg, dd + int64_t(f >> 1) * stride);
#pragma unroll
      for (int i = 0; i < F; ++i) out[i] = (i == f) ? v : out[i];
    }
  }
}

// correct_naive_second_derivatives from the hanging faces' m terms: face f
// of axis a adds d_m0 d_p0 / d_f / (d_a+ + d_a-) to m_a,first transverse and
// (3D) d_0m d_0p / d_f / (d_a+ + d_a-) to m_a,second, as div_d / div_dsum
template <int D, class R, class TB>
__device__ __forceinline__ void hang_correct_naive(const TB &t, R &Dxx, R &Dyy, R &Dzz) {
  constexpr int F = 2 * D;
  R m[3][2] = {{R(0), R(0)}, {R(0), R(0)}, {R(0), R(0)}};
#pragma unroll
  for (int q = 0; q < TB::NH; ++q) {
    const int f = t.hface(q);
    if (q == 0 || f < F) {
      const int a = f >> 1;
      const auto sg = t.slot_geom(f);
      const R df = sg.df, iqf = sg.iqf, dp = sg.dp, dm = sg.dm, iqp = sg.iqp;
      auto term = [&](R x) {
        // div_d(x, f), then div_dsum(., a)
        if constexpr (irregular_fastdiv<R, TB>()) {
          if (t.fast) x = x * iqf;
          else x = qdiv(x, df);
          if (t.fast && dp == dm) return x * (R(0.5) * iqp);
          return qdiv(x, dp + dm);
        } else {
          return qdiv(qdiv(x, df), dp + dm);
        }
      };
      const R T1 = term(t.hdt(q, 0, sg) * t.hdt(q, 1, sg));
#pragma unroll
      for (int b = 0; b < D; ++b) m[b][0] = m[b][0] + ((a == b) ? T1 : R(0));
      if constexpr (D == 3) {
        const R T2 = term(t.hdt(q, 2, sg) * t.hdt(q, 3, sg));
#pragma unroll
        for (int b = 0; b < D; ++b) m[b][1] = m[b][1] + ((a == b) ? T2 : R(0));
      }
    }
  }
  correct_naive_solve<D, R>(m[0][0], m[1][0], m[0][1], m[1][1], m[2][0], m[2][1], Dxx, Dyy, Dzz);
}

// laplace of a node with hanging faces; vq[q] keeps hanging face q's value
template <int D, class R, class G, class TB>
__device__ __forceinline__ void hang_laplace(const TB &t, int64_t n, const G *__restrict__ g, R &f000, R *fnb, R *dd,
                                             R *vq) {
  f000 = R(g[n]);
  hang_face_values<D, R, false>(t, g, fnb, dd, vq);
  dd[0] = second_difference<R>(t, 0, fnb[1], f000, fnb[0], n);
  dd[1] = second_difference<R>(t, 1, fnb[3], f000, fnb[2], n);
  R dzz = R(0);
  if constexpr (D == 3) dzz = second_difference<R>(t, 2, fnb[5], f000, fnb[4], n);
  hang_correct_naive<D, R>(t, dd[0], dd[1], dzz);
  if constexpr (D == 3) dd[2] = dzz;
}

// laplace, my_p4est_quad_neighbor_nodes_of_node.h:615-635 (one field)
template <int D, class R, class G, class TB>
__device__ __forceinline__ void laplace(const TB &t, int64_t n, const G *__restrict__ g, R &f000, R *fnb, R *dd) {
  if constexpr (TB::SLOTS) {
    R vq[TB::NH];
    hang_laplace<D, R>(t, n, g, f000, fnb, dd, vq);
    return;
  }
  f000 = R(g[n]);
#pragma unroll
  for (int f = 0; f < 2 * D; ++f) fnb[f] = face_linear<D, R>(t, f, n, g);
  dd[0] = second_difference<R>(t, 0, fnb[1], f000, fnb[0], n);
  dd[1] = second_difference<R>(t, 1, fnb[3], f000, fnb[2], n);
  R dzz = R(0);
  if constexpr (D == 3) dzz = second_difference<R>(t, 2, fnb[5], f000, fnb[4], n);
  correct_naive<D, R>(t, n, dd[0], dd[1], dzz);
  if constexpr (D == 3) dd[2] = dzz;
}

// The third-order corrections of ngbd_with_quadratic_interpolation,
// my_p4est_quad_neighbor_nodes_of_node.h:1530-1538: fnb[f] -= 0.5 d_m0 d_p0 f_tt (+ ...)
template <int D, class R, class TB>
__device__ __forceinline__ void quadratic_corrections(const TB &t, int64_t n, R *fnb, const R *dd) {
  if constexpr (TB::REGULAR) return;
  const R h = R(0.5);
  if constexpr (D == 2) {
    fnb[0] -= (h * R(t.dt(0, 0, n)) * R(t.dt(0, 1, n)) * dd[1]);
    fnb[1] -= (h * R(t.dt(1, 0, n)) * R(t.dt(1, 1, n)) * dd[1]);
    fnb[2] -= (h * R(t.dt(2, 0, n)) * R(t.dt(2, 1, n)) * dd[0]);
    fnb[3] -= (h * R(t.dt(3, 0, n)) * R(t.dt(3, 1, n)) * dd[0]);
  } else {
    // fnb[f] -= (h d_m0 d_p0 f_tt1 + h d_0m d_0p f_tt2), the first product fused in FP32
    constexpr int t1[6] = {1, 1, 0, 0, 0, 0}, t2[6] = {2, 2, 2, 2, 1, 1};
#pragma unroll
    for (int f = 0; f < 6; ++f)
      fnb[f] -= qfma(h * R(t.dt(f, 0, n)) * R(t.dt(f, 1, n)), dd[t1[f]],
                     h * R(t.dt(f, 2, n)) * R(t.dt(f, 3, n)) * dd[t2[f]]);
  }
}

// ngbd_with_quadratic_interpolation, my_p4est_quad_neighbor_nodes_of_node.h:1514-1545
template <int D, class R, class G, class TB>
__device__ __forceinline__ void ngbd_quadratic(const TB &t, int64_t n, const G *__restrict__ g, R &f000, R *fnb,
                                               R *dd) {
  if constexpr (TB::SLOTS) {
    constexpr int F = 2 * D;
    R vq[TB::NH];
    hang_laplace<D, R>(t, n, g, f000, fnb, dd, vq);
#pragma unroll
    for (int q = 0; q < TB::NH; ++q) {
      const int f = t.hface(q);
      if (q == 0 || f < F) {
        const R v = vq[q] - hang_correction<D, R>(t, q, f >> 1, t.slot_geom(f), dd);
#pragma unroll
        for (int i = 0; i < F; ++i) fnb[i] = (i == f) ? v : fnb[i];
      }
    }
  } else {
    laplace<D, R>(t, n, g, f000, fnb, dd);
    quadratic_corrections<D, R>(t, n, fnb, dd);
  }
}

// The same with the node's own second derivatives already known (they are the
// laplace values of this node, computed by the same arithmetic elsewhere).
template <int D, class R, class G, class TB>
__device__ __forceinline__ void ngbd_quadratic_given_dd(const TB &t, int64_t n, const G *__restrict__ g,
                                                        const R *dd, R &f000, R *fnb) {
  f000 = R(g[n]);
  if constexpr (TB::SLOTS) {
    R vq[TB::NH];
    hang_face_values<D, R, true>(t, g, fnb, dd, vq);
    return;
  }
#pragma unroll
  for (int f = 0; f < 2 * D; ++f) fnb[f] = face_linear<D, R>(t, f, n, g);
  quadratic_corrections<D, R>(t, n, fnb, dd);
}

// per face f, the linear face value of component f >> 1 of dd (entries
// stride apart): f_{m00,p00,...}_linear of the stored second derivatives
template <int D, class R, class G, class TB>
__device__ __forceinline__ void face_values_dd(const TB &t, int64_t n, const G *__restrict__ dd, int64_t stride,
                                               R *out) {
  if constexpr (TB::SLOTS) {
    hang_face_values_dd<D, R>(t, dd, stride, out);
  } else {
#pragma unroll
    for (int f = 0; f < 2 * D; ++f) out[f] = face_linear<D, R>(t, f, n, dd + int64_t(f >> 1) * stride);
  }
}

// ---------------------------------------------------- first derivatives
// central_derivative, my_p4est_quad_neighbor_nodes_of_node.h:200-206
template <class R>
__host__ __device__ __forceinline__ R central_derivative(R fp, R f0, R fm, R dp, R dm) {
  return ((fp - f0) * dm / dp + (f0 - fm) * dp / dm) / (dp + dm);
}

// dd_correction_weight_to_naive_first_derivative, my_p4est_quad_neighbor_nodes_of_node.h:186-192
template <class R>
__host__ __device__ __forceinline__ R dd_correction_weight(R d_m, R d_p, R d_m_m, R d_m_p, R d_p_m, R d_p_p) {
  return R(0.5) * (d_p_m * d_p_p * d_m / d_p - d_m_m * d_m_p * d_p / d_m) / (d_m + d_p);
}

// 1 / (the largest arm or transverse distance), the list construct_neighbors
// uses (src/my_p4est_node_neighbors.cpp:1168-1186; in 3D its list pairs
// d_00m_0m with d_00p_0p and leaves d_00m_0p out).
template <int D, class TB>
__device__ __forceinline__ double inverse_d_max(const TB &t, int64_t n) {
  double m = max2(double(t.d(0, n)), double(t.d(1, n)));
  m = max2(max2(m, double(t.d(2, n))), double(t.d(3, n)));
  if constexpr (D == 3) m = max2(max2(m, double(t.d(4, n))), double(t.d(5, n)));
  // face f, transverse k (m0, p0[, 0m, 0p])
#pragma unroll
  for (int f = 0; f < 4; ++f) m = max2(max2(m, double(t.dt(f, 0, n))), double(t.dt(f, 1, n)));
  if constexpr (D == 3) {
#pragma unroll
    for (int f = 0; f < 4; ++f) m = max2(max2(m, double(t.dt(f, 2, n))), double(t.dt(f, 3, n)));
    m = max2(max2(m, double(t.dt(4, 0, n))), double(t.dt(4, 1, n)));  // d_00m_m0, d_00m_p0
    m = max2(max2(m, double(t.dt(4, 2, n))), double(t.dt(5, 3, n)));  // d_00m_0m, d_00p_0p
    m = max2(max2(m, double(t.dt(5, 0, n))), double(t.dt(5, 1, n)));  // d_00p_m0, d_00p_p0
    m = max2(max2(m, double(t.dt(5, 2, n))), double(t.dt(5, 3, n)));  // d_00p_0m, d_00p_0p
  }
  return 1.0 / m;
}

// Central first derivatives of f at one node, dx_central / dy_central /
// dz_central (my_p4est_quad_neighbor_nodes_of_node.cpp:638-700 and the y, z
// counterparts): the naive central derivative on the linearly interpolated
// face values, minus f_tt * weight for each transverse axis t whose weight
// (scaled by inverse_d_max) is not zero -- f_tt from laplace, computed only
// when some correction is needed.  FP64.
template <int D, class G, class TB>
__device__ __forceinline__ void gradient_central(const TB &t, int64_t n, const G *__restrict__ f, double *grad) {
  const double f000 = double(f[n]);
  double fnb[2 * D];
#pragma unroll
  for (int q = 0; q < 2 * D; ++q) fnb[q] = face_linear<D, double>(t, q, n, f);
#pragma unroll
  for (int a = 0; a < D; ++a)
    grad[a] = central_derivative(fnb[2 * a + 1], f000, fnb[2 * a], double(t.d(2 * a + 1, n)), double(t.d(2 * a, n)));
  const double idm = inverse_d_max<D>(t, n);
  double w[D][2];
  bool need[D][2];
  bool any = false;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int k = 0; k < D - 1; ++k) {
      const int fm = 2 * a, fp = 2 * a + 1;
      w[a][k] = dd_correction_weight(double(t.d(fm, n)), double(t.d(fp, n)), double(t.dt(fm, 2 * k, n)),
                                     double(t.dt(fm, 2 * k + 1, n)), double(t.dt(fp, 2 * k, n)),
                                     double(t.dt(fp, 2 * k + 1, n)));
      need[a][k] = !check_if_zero(w[a][k] * idm);
      any = any || need[a][k];
    }
  if (any) {
    double f0, fn2[2 * D], dd[D];
    laplace<D, double>(t, n, f, f0, fn2, dd);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int k = 0; k < D - 1; ++k) {
        // transverse axes of axis a, in order: x -> (y, z), y -> (x, z), z -> (x, y)
        const int ta = (k == 0) ? (a == 0 ? 1 : 0) : (a == 2 ? 1 : 2);
        if (need[a][k]) grad[a] -= dd[ta] * w[a][k];
      }
  }
}

}  // namespace dev
}  // namespace casl
