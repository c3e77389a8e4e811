// Numerical curvature and learned-curvature input samples on uniform patches.
//
// Numerical curvature follows the reference: normals n = grad(phi)/|grad(phi)|
// from central differences, zero where |grad(phi)| < EPS (compute_normals,
// src/my_p4est_utils.cpp:1698-1740), then kappa = d(nx)/dx + d(ny)/dy by
// central differences (qnnn::get_curvature(normals),
// src/my_p4est_quad_neighbor_nodes_of_node.h:2167-2171).  A sample is a node
// next to the interface (a face neighbour of opposite sign); its features are
// the 3x3 phi stencil over h (x slowest, y fastest: mm, m0, mp, 0m, 00, 0p,
// pm, p0, pp, src/my_p4est_curvature_ml.h:49-60) and h*kappa bilinearly
// interpolated at the projection x - phi n onto the interface (the
// reference's "ihk", src/my_p4est_curvature_ml.cpp:414-480).
#include <hip/hip_runtime.h>

#include "curv_device.hpp"
#include "reinit_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

struct Patch {
  const double *phi;     // [n_patch][n_nodes]
  const int32_t *grid;   // [(M+1)^2], node index of (i, j) (x index i, y index j)
  int M;
  int64_t n_nodes;
  double h;
};

// central_derivative on a uniform grid, my_p4est_quad_neighbor_nodes_of_node.h:200-206
__device__ __forceinline__ double cdiff(double fp, double f0, double fm, double h) {
  return ((fp - f0) * h / h + (f0 - fm) * h / h) / (h + h);
}

__global__ void k_normals(Patch P, int n_patch, double *__restrict__ nrm) {
  const int W = P.M + 1;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(n_patch) * W * W) return;
  const int p = int(t / (W * W)), ij = int(t % (W * W)), i = ij / W, j = ij % W;
  double nx = 0.0, ny = 0.0;
  if (i >= 1 && i < P.M && j >= 1 && j < P.M) {
    const double *f = P.phi + int64_t(p) * P.n_nodes;
    const double f0 = f[P.grid[i * W + j]];
    const double fx = cdiff(f[P.grid[(i + 1) * W + j]], f0, f[P.grid[(i - 1) * W + j]], P.h);
    const double fy = cdiff(f[P.grid[i * W + j + 1]], f0, f[P.grid[i * W + j - 1]], P.h);
    const double a = sqrt(fx * fx + fy * fy);
    if (a >= dev::EPS) {
      nx = fx / a;
      ny = fy / a;
    }
  }
  nrm[2 * t] = nx;
  nrm[2 * t + 1] = ny;
}

__global__ void k_kappa(Patch P, int n_patch, const double *__restrict__ nrm, double *__restrict__ kap) {
  const int W = P.M + 1;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(n_patch) * W * W) return;
  const int ij = int(t % (W * W)), i = ij / W, j = ij % W;
  double k = 0.0;
  if (i >= 2 && i < P.M - 1 && j >= 2 && j < P.M - 1) {
    const int64_t b = t - ij;
    k = cdiff(nrm[2 * (b + (i + 1) * W + j)], nrm[2 * t], nrm[2 * (b + (i - 1) * W + j)], P.h) +
        cdiff(nrm[2 * (b + i * W + j + 1) + 1], nrm[2 * t + 1], nrm[2 * (b + i * W + j - 1) + 1], P.h);
  }
  kap[t] = k;
}

__global__ void k_samples(Patch P, int n_patch, const double *__restrict__ nrm, const double *__restrict__ kap,
                          double x0, double y0, int64_t cap, float *__restrict__ feat, double *__restrict__ proj,
                          int64_t *__restrict__ node, unsigned long long *__restrict__ count) {
  const int W = P.M + 1;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(n_patch) * W * W) return;
  const int p = int(t / (W * W)), ij = int(t % (W * W)), i = ij / W, j = ij % W;
  if (i < 4 || i > P.M - 4 || j < 4 || j > P.M - 4) return;
  const double *f = P.phi + int64_t(p) * P.n_nodes;
  const double f0 = f[P.grid[i * W + j]];
  const double fe = f[P.grid[(i + 1) * W + j]], fw = f[P.grid[(i - 1) * W + j]];
  const double fn = f[P.grid[i * W + j + 1]], fs = f[P.grid[i * W + j - 1]];
  if (!(f0 * fe <= 0.0 || f0 * fw <= 0.0 || f0 * fn <= 0.0 || f0 * fs <= 0.0)) return;
  // projection onto the interface and bilinear h*kappa there
  const double x = x0 + i * P.h, y = y0 + j * P.h;
  const double xg = x - f0 * nrm[2 * t], yg = y - f0 * nrm[2 * t + 1];
  double u = (xg - x0) / P.h, v = (yg - y0) / P.h;
  int ci = int(floor(u)), cj = int(floor(v));
  ci = ci < 2 ? 2 : (ci > P.M - 3 ? P.M - 3 : ci);
  cj = cj < 2 ? 2 : (cj > P.M - 3 ? P.M - 3 : cj);
  u -= ci;
  v -= cj;
  const int64_t b = t - ij;
  const double k00 = kap[b + ci * W + cj], k10 = kap[b + (ci + 1) * W + cj];
  const double k01 = kap[b + ci * W + cj + 1], k11 = kap[b + (ci + 1) * W + cj + 1];
  const double kg = (1 - u) * (1 - v) * k00 + u * (1 - v) * k10 + (1 - u) * v * k01 + u * v * k11;
  const unsigned long long s = atomicAdd(count, 1ull);
  if (int64_t(s) >= cap) return;
  float *o = feat + s * 10;
  int q = 0;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj) o[q++] = float(f[P.grid[(i + di) * W + j + dj]] / P.h);
  o[9] = float(P.h * kg);
  proj[2 * s] = xg;
  proj[2 * s + 1] = yg;
  node[s] = int64_t(p) * P.n_nodes + P.grid[i * W + j];
}

__global__ void k_zero_count(unsigned long long *c) { *c = 0ull; }

// Network inputs of a sample (kml::Curvature::_computeHybridHK,
// src/my_p4est_curvature_ml.cpp:483-520): negate every feature when the
// numerical h*kappa is positive, rotate the stencil by multiples of 90 degrees
// so that the gradient at the centre points into the first quadrant
// (rotateStencilToFirstQuadrant, :373-398; rotateStencil90, :327-353; a zero
// gradient component counts as +EPS), and emit it twice: as is (row 2s) and
// reflected about y = x (row 2s+1, reflectStencil_yEqx :356-370).  The
// gradient is the central difference of the phi stencil (the reference
// carries normal stencils for it; only their direction matters here).
__global__ void k_normalize(const float *__restrict__ feat, int64_t n, float *__restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  curv_normalize2(feat + 10 * s, out + 20 * s);
}

// Hybrid h*kappa (src/my_p4est_curvature_ml.cpp:483-556): the numerical value
// where |ihk| < lo; elsewhere the mean of the two network outputs (plus the
// normalised numerical value -|ihk| when the network predicts a correction),
// blended linearly towards -|ihk| for |ihk| <= hi, and given the sign of ihk.
__global__ void k_hybrid(const float *__restrict__ feat, const float *__restrict__ pred, int64_t n, float lo,
                         float hi, int residual, float *__restrict__ hk) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  hk[s] = curv_hybrid2(feat[10 * s + 9], pred[2 * s], pred[2 * s + 1], lo, hi, residual);
}

}  // namespace

extern "C" {

int casl_curvature_samples(const double *phi, int n_patch, int64_t n_nodes, int M, const int32_t *grid, double h,
                           double x0, double y0, float *feat, double *proj, int64_t *node, int64_t max_samples,
                           int64_t *n_samples, void *stream) {
  return guarded([&]() {
    if (!phi || !grid || !feat || !proj || !node || !n_samples || n_patch <= 0 || M < 9 || !(h > 0))
      throw std::invalid_argument("bad arguments");
    if (int64_t(M + 1) * (M + 1) != n_nodes) throw std::invalid_argument("grid size does not match n_nodes");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t tot = int64_t(n_patch) * (M + 1) * (M + 1);
    DevBuf<double> nrm, kap;
    DevBuf<unsigned long long> cnt;
    nrm.ensure(size_t(2 * tot));
    kap.ensure(size_t(tot));
    cnt.ensure(1);
    Patch P{phi, grid, M, n_nodes, h};
    const unsigned g = unsigned((tot + 255) / 256);
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, cnt.p);
    hipLaunchKernelGGL(k_normals, dim3(g), dim3(256), 0, s, P, n_patch, nrm.p);
    hipLaunchKernelGGL(k_kappa, dim3(g), dim3(256), 0, s, P, n_patch, nrm.p, kap.p);
    hipLaunchKernelGGL(k_samples, dim3(g), dim3(256), 0, s, P, n_patch, nrm.p, kap.p, x0, y0, max_samples, feat, proj,
                       node, cnt.p);
    HIP_CHECK(hipGetLastError());
    unsigned long long c = 0;
    HIP_CHECK(hipMemcpyAsync(&c, cnt.p, sizeof(c), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    *n_samples = int64_t(c);
  });
}

int casl_curvature_normalize(const float *feat, int64_t n, float *out, void *stream) {
  return guarded([&]() {
    if (n < 0 || (n > 0 && (!feat || !out))) throw std::invalid_argument("bad arguments");
    if (n == 0) return;
    hipLaunchKernelGGL(k_normalize, dim3(unsigned((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       feat, n, out);
    HIP_CHECK(hipGetLastError());
  });
}

int casl_curvature_hybrid(const float *feat, const float *pred, int64_t n, float lo, float hi, int residual,
                          float *hk, void *stream) {
  return guarded([&]() {
    if (n < 0 || (n > 0 && (!feat || !pred || !hk)) || !(hi > lo) || !(lo >= 0.f))
      throw std::invalid_argument("bad arguments");
    if (n == 0) return;
    hipLaunchKernelGGL(k_hybrid, dim3(unsigned((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       feat, pred, n, lo, hi, residual, hk);
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
