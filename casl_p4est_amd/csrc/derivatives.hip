// First derivatives, normals and mean curvature on the flattened stencils
// (the numerical curvature chain of the reference), FP64:
//   casl_first_derivatives  my_p4est_node_neighbors_t::first_derivatives_central
//                           (src/my_p4est_node_neighbors.h:568-600), per node
//                           dx/dy/dz_central (src/my_p4est_quad_neighbor_nodes_of_node.cpp:638-760)
//   casl_normals            compute_normals (src/my_p4est_utils.cpp:1720-1740):
//                           gradient / |gradient|, 0 where |gradient| < EPS
//   casl_mean_curvature     compute_mean_curvature(neighbors, normals, kappa)
//                           (src/my_p4est_utils.cpp:1648-1690) with
//                           qnnn::get_curvature(normals) (src/my_p4est_quad_neighbor_nodes_of_node.h:2167-2171):
//                           dx_central(n_x) + dy_central(n_y) [+ dz_central(n_z)]
// One thread per node on the grid-wide tables; every call is one pass.
#include <hip/hip_runtime.h>

#include "reinit_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

constexpr int BLOCK = 256;

template <int D, class TB>
__global__ void __launch_bounds__(BLOCK) k_gradient(TB t, const double *__restrict__ f, double *__restrict__ fd) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  double g[D];
  gradient_central<D>(t, n, f, g);
#pragma unroll
  for (int a = 0; a < D; ++a) fd[int64_t(a) * t.N + n] = g[a];
}

template <int D, class TB>
__global__ void __launch_bounds__(BLOCK) k_normals(TB t, const double *__restrict__ phi, double *__restrict__ nrm) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  double g[D];
  gradient_central<D>(t, n, phi, g);
  double s = g[0] * g[0] + g[1] * g[1];
  if constexpr (D == 3) s = s + g[2] * g[2];
  const double a = sqrt(s);
#pragma unroll
  for (int k = 0; k < D; ++k) nrm[int64_t(k) * t.N + n] = (a < EPS) ? 0.0 : g[k] / a;
}

// kappa = sum_a d(n_a)/dx_a, each with its own transverse corrections
template <int D, class TB>
__global__ void __launch_bounds__(BLOCK) k_curvature(TB t, const double *__restrict__ nrm, double *__restrict__ kap) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  double k = 0.0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double g[D];
    gradient_central<D>(t, n, nrm + int64_t(a) * t.N, g);
    k = (a == 0) ? g[0] : k + g[a];
  }
  kap[n] = k;
}

inline unsigned grid_for(int64_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }

void check_stencil(const casl_stencil *st, const void *a, const void *b) {
  if (!st || !a || !b) throw std::invalid_argument("null argument");
}

}  // namespace

extern "C" {

int casl_first_derivatives(casl_stencil_t *st, const double *f, double *fd, void *stream) {
  return guarded([&]() {
    check_stencil(st, f, fd);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (st->dim == 2)
      st->with_tables<2>(true, [&](auto t) {
        hipLaunchKernelGGL((k_gradient<2, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, f, fd);
      });
    else
      st->with_tables<3>(true, [&](auto t) {
        hipLaunchKernelGGL((k_gradient<3, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, f, fd);
      });
    HIP_CHECK(hipGetLastError());
  });
}

int casl_normals(casl_stencil_t *st, const double *phi, double *normals, void *stream) {
  return guarded([&]() {
    check_stencil(st, phi, normals);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (st->dim == 2)
      st->with_tables<2>(true, [&](auto t) {
        hipLaunchKernelGGL((k_normals<2, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, phi, normals);
      });
    else
      st->with_tables<3>(true, [&](auto t) {
        hipLaunchKernelGGL((k_normals<3, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, phi, normals);
      });
    HIP_CHECK(hipGetLastError());
  });
}

int casl_mean_curvature(casl_stencil_t *st, const double *normals, double *kappa, void *stream) {
  return guarded([&]() {
    check_stencil(st, normals, kappa);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (st->dim == 2)
      st->with_tables<2>(true, [&](auto t) {
        hipLaunchKernelGGL((k_curvature<2, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, normals, kappa);
      });
    else
      st->with_tables<3>(true, [&](auto t) {
        hipLaunchKernelGGL((k_curvature<3, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, normals, kappa);
      });
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
