// Learned-curvature inputs on the adaptive grid, for 2D and 3D trees.
//
// The reference's learned curvature (kml::Curvature, src/my_p4est_curvature_ml.cpp)
// is 2D and reads 3x3 phi stencils of uniform-grid nodes next to the
// interface.  Here the same inputs come from any grid, through the
// construct_neighbors tables: for every node with |phi| < band, the block of
// 3^dim nodes at spacing h around it is found by walking face neighbours
// (x, then y, then z), every arm on the way being exactly h long and ending
// on a node (a zero transverse distance along every transverse axis).  A
// node without such a block keeps the numerical value; it is counted, not
// sampled.
//
//   features  phi/h over the block (x slowest, z fastest: the 2D order of
//             src/my_p4est_curvature_ml.h:49-60 extended by one axis), then
//             h*kappa interpolated (bi/trilinearly, inside the block) at the
//             projection x - phi n of the node onto the interface ("ihk",
//             src/my_p4est_curvature_ml.cpp:414-480), kappa being
//             compute_mean_curvature's div(n) and n compute_normals' output
//   3D normalisation  negate when ihk > 0; reflect every axis along which the
//             centred difference of phi is negative; order the axes by
//             decreasing difference.  Every one of the 48 symmetries of the
//             cube maps a stencil to the same input: the 3D counterpart of the
//             2D rotation to the first quadrant plus y = x reflection
//             (:327-398), with one row per sample instead of two.  The 27
//             values then go to a basis that BF16 inputs resolve: the centre
//             value, the half centred differences, and the deviations from
//             the plane through the centre with those slopes.
//   3D hybrid  as the 2D one (:483-556): numerical value where |ihk| < lo,
//             blend towards it up to hi, sign of ihk; a residual network's
//             corrections at most `deadband` in size are dropped.
#include <hip/hip_runtime.h>

#include "curv_device.hpp"
#include "reinit_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

constexpr int BLOCK = 256;

inline unsigned grid_for(int64_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }

template <int D>
struct Block {
  static constexpr int Q = (D == 2 ? 9 : 27);
  static constexpr int NF = Q + 1;
  // position of offsets (i, j[, k]) in {-1, 0, 1}^D, x slowest
  __host__ __device__ static constexpr int q(int i, int j, int k = 0) {
    return D == 2 ? 3 * (i + 1) + (j + 1) : 9 * (i + 1) + 3 * (j + 1) + (k + 1);
  }
};

// the arm of node n through face f if it is exactly h long and ends on a
// node: along every transverse axis one of the two transverse distances is
// zero, and the face corner on that side is the node
template <int D>
__device__ __forceinline__ bool arm(const Tables<D> &t, int f, int64_t n, float h, int32_t &out) {
  if (t.d(f, n) != h) return false;
  int c = 0;
#pragma unroll
  for (int k = 0; k < D - 1; ++k) {
    if (t.dt(f, 2 * k, n) == 0.f) {
      // corner bit 0 along this axis
    } else if (t.dt(f, 2 * k + 1, n) == 0.f) {
      c |= 1 << k;
    } else {
      return false;
    }
  }
  out = t.idx(f, c, n);
  return true;
}

template <int D>
__device__ bool gather_block(const Tables<D> &t, int64_t n, float h, int32_t *blk) {
  using B = Block<D>;
  blk[B::q(0, 0, 0)] = int32_t(n);
  if (!arm<D>(t, 0, n, h, blk[B::q(-1, 0, 0)]) || !arm<D>(t, 1, n, h, blk[B::q(1, 0, 0)])) return false;
  for (int i = -1; i <= 1; ++i) {
    const int32_t b = blk[B::q(i, 0, 0)];
    if (!arm<D>(t, 2, b, h, blk[B::q(i, -1, 0)]) || !arm<D>(t, 3, b, h, blk[B::q(i, 1, 0)])) return false;
  }
  if constexpr (D == 3) {
    for (int i = -1; i <= 1; ++i)
      for (int j = -1; j <= 1; ++j) {
        const int32_t b = blk[B::q(i, j, 0)];
        if (!arm<D>(t, 4, b, h, blk[B::q(i, j, -1)]) || !arm<D>(t, 5, b, h, blk[B::q(i, j, 1)])) return false;
      }
  }
  return true;
}

template <int D>
__global__ void __launch_bounds__(BLOCK)
    k_grid_samples(Tables<D> t, const double *__restrict__ phi, const double *__restrict__ nrm,
                   const double *__restrict__ kap, double h, double band, int64_t cap, float *__restrict__ feat,
                   double *__restrict__ offset, int64_t *__restrict__ node, unsigned long long *__restrict__ count) {
  using B = Block<D>;
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  const double f0 = phi[n];
  if (!(fabs(f0) < band)) return;
  int32_t blk[B::Q];
  if (!gather_block<D>(t, n, float(h), blk)) {
    atomicAdd(count + 1, 1ull);
    return;
  }
  // the projection's offset, in units of h clamped to the block, and the
  // cell of the block that holds it
  double off[D], w[D];
  int lo[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    off[a] = -f0 * nrm[int64_t(a) * t.N + n];
    double u = off[a] / h;
    u = u < -1.0 ? -1.0 : (u > 1.0 ? 1.0 : u);
    lo[a] = u < 0.0 ? -1 : 0;
    w[a] = u - double(lo[a]);
  }
  double k = 0.0;
#pragma unroll
  for (int c = 0; c < (1 << D); ++c) {
    double wt = 1.0;
    int o[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int bit = (c >> a) & 1;
      wt = wt * (bit ? w[a] : 1.0 - w[a]);
      o[a] = lo[a] + bit;
    }
    k = k + wt * kap[blk[B::q(o[0], o[1], o[2])]];
  }
  const unsigned long long s = atomicAdd(count, 1ull);
  if (int64_t(s) >= cap) return;
  float *fo = feat + int64_t(s) * B::NF;
#pragma unroll
  for (int q = 0; q < B::Q; ++q) fo[q] = float(phi[blk[q]] / h);
  fo[B::Q] = float(h * k);
#pragma unroll
  for (int a = 0; a < D; ++a) offset[int64_t(s) * D + a] = off[a];
  node[s] = n;
}

__global__ void k_zero2(unsigned long long *c) {
  c[0] = 0ull;
  c[1] = 0ull;
}

// [n][28] samples -> [n][28] network inputs (one canonical orientation)
__global__ void __launch_bounds__(BLOCK) k_normalize3(const float *__restrict__ feat, int64_t n,
                                                      float *__restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  curv_normalize3(feat + 28 * s, out + 28 * s);
}

// Hybrid h*kappa of 3D samples: one network output per sample.  With a
// residual network, corrections no larger than `dead` are dropped: on a
// well-resolved surface the numerical value is more accurate than the
// network's noise floor.
__global__ void __launch_bounds__(BLOCK) k_hybrid3(const float *__restrict__ feat, const float *__restrict__ pred,
                                                   int64_t n, float lo, float hi, int residual, float dead,
                                                   float *__restrict__ hk) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  hk[s] = curv_hybrid3(feat[28 * s + 27], pred[s], lo, hi, residual, dead);
}

}  // namespace

extern "C" {

int casl_curvature_grid_samples(casl_stencil_t *st, const double *phi, const double *normals, const double *kappa,
                                double h, double band, float *feat, double *offset, int64_t *node,
                                int64_t max_samples, int64_t *n_samples, int64_t *n_skipped, void *stream) {
  return guarded([&]() {
    if (!st || !phi || !normals || !kappa || !n_samples || !(h > 0) || !(band > 0) || max_samples < 0)
      throw std::invalid_argument("bad arguments");
    if (max_samples > 0 && (!feat || !offset || !node)) throw std::invalid_argument("null output arrays");
    if (double(float(h)) != h) throw std::invalid_argument("h must be exact in FP32 (a grid spacing)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    st->red.ensure(2);
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, s, st->red.p);
    if (st->dim == 2)
      hipLaunchKernelGGL(k_grid_samples<2>, dim3(grid_for(st->N)), dim3(BLOCK), 0, s, st->tables<2>(), phi, normals,
                         kappa, h, band, max_samples, feat, offset, node, st->red.p);
    else
      hipLaunchKernelGGL(k_grid_samples<3>, dim3(grid_for(st->N)), dim3(BLOCK), 0, s, st->tables<3>(), phi, normals,
                         kappa, h, band, max_samples, feat, offset, node, st->red.p);
    HIP_CHECK(hipGetLastError());
    unsigned long long c[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(c, st->red.p, sizeof(c), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    *n_samples = int64_t(c[0]);
    if (n_skipped) *n_skipped = int64_t(c[1]);
  });
}

int casl_curvature_normalize_3d(const float *feat, int64_t n, float *out, void *stream) {
  return guarded([&]() {
    if (n < 0 || (n > 0 && (!feat || !out))) throw std::invalid_argument("bad arguments");
    if (n == 0) return;
    hipLaunchKernelGGL(k_normalize3, dim3(grid_for(n)), dim3(BLOCK), 0, static_cast<hipStream_t>(stream), feat, n,
                       out);
    HIP_CHECK(hipGetLastError());
  });
}

int casl_curvature_hybrid_3d(const float *feat, const float *pred, int64_t n, float lo, float hi, int residual,
                             float deadband, float *hk, void *stream) {
  return guarded([&]() {
    if (n < 0 || (n > 0 && (!feat || !pred || !hk)) || !(hi > lo) || !(lo >= 0.f) || !(deadband >= 0.f))
      throw std::invalid_argument("bad arguments");
    if (n == 0) return;
    hipLaunchKernelGGL(k_hybrid3, dim3(grid_for(n)), dim3(BLOCK), 0, static_cast<hipStream_t>(stream), feat, pred, n,
                       lo, hi, residual, deadband, hk);
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
