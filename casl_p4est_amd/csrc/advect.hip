// advect_in_normal_direction (src/my_p4est_level_set.cpp:559-726, 785-815)
// and the time step of its velocity-function form (:744-783), FP64:
//   phi_n = phi;  dd = second_derivatives_central(phi_n)
//   phi   = stage(vn,    dd, phi_n)            (advect_in_normal_direction_one_iteration)
//   dd    = second_derivatives_central(phi)
//   phi   = (phi_n + stage(vn_np1 or vn, dd, phi)) / 2
// Tiled path (default when the stencil has a tiled encoding): one fused
// kernel per stage over the Morton tiles, pass A of the stage's input for
// T u H1 in LDS and the stage for T, phi read once per stage; the first
// stage writes a scratch array, the second averages into phi in place.
// Grid-wide path: pass A and the stage as separate kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "reinit_device.hpp"
#include "tiles_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

constexpr int BLOCK = 256;

inline unsigned grid_for(int64_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }

template <int D, class TB>
__global__ void __launch_bounds__(BLOCK) k_adv_dd(TB t, const double *__restrict__ f, double *__restrict__ fdd) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n < t.N) node_dd<D, double>(t, n, f, fdd);
}

// out = stage, or (RK_AVG) out = (out + stage) / 2 with out holding phi_n
template <int D, bool RK_AVG, class TB>
__global__ void __launch_bounds__(BLOCK) k_advect(TB t, const double *__restrict__ vn, double dt,
                                                  const double *__restrict__ phin, const double *__restrict__ dd,
                                                  double *out) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  const double r = advect_value<D>(t, n, t.wall[n], vn[n], dt, phin, dd, t.N);
  if constexpr (RK_AVG)
    out[n] = 0.5 * (out[n] + r);
  else
    out[n] = r;
}

// One advection stage of a tile: phi of T u H1 u H2 into LDS, pass A for
// T u H1, then advect_value for T.  out = result, or (RK_AVG) out = (out +
// result) / 2.  The loads a node needs (its record, slot and vn, and the
// value being averaged) are issued before the first barrier.
template <int D, bool RK_AVG>
__global__ void __launch_bounds__(TILE_BLOCK_F64) __attribute__((amdgpu_waves_per_eu(CASL_TILE_WAVES_F64)))
    k_tile_advect(TileArgs<double> A, const double *__restrict__ vn, double dt) {
  constexpr int DMAX = TileLimits<D>::DMAX;
  constexpr int KT = TileLimits<D>::TMAX / TILE_BLOCK_F64;
  __shared__ double phis[TileLimits<D>::LMAX];
  __shared__ double dds[D * DMAX];
  const int tile = xcd_tile(int(blockIdx.x), int(gridDim.x));
  TileCtx<D> c;
  c.load(A, tile);
  const int s = c.s, nT = c.nT, hb = c.hb, nH = c.nH, nH1 = c.nH1;
  const bool ufast = A.geom.pow2 != 0 && A.geom.cubic != 0;
  RecRegs trec[KT];
  int tslot[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int i = int(threadIdx.x) + k * TILE_BLOCK_F64;
    if (i < nT) tslot[k] = fetch_record<D>(A, c, i, trec[k]);
  }
  for (int j = threadIdx.x; j < nT; j += TILE_BLOCK_F64) phis[j] = A.phin[s + j];
  for (int j = threadIdx.x; j < nH; j += TILE_BLOCK_F64) phis[nT + j] = A.phin[A.halo[hb + j]];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int i = int(threadIdx.x) + k * TILE_BLOCK_F64;
    if (i < nT) tile_pass_a<D, double>(A, c, ufast, trec[k], tslot[k], i, phis, dds);
  }
  for (int i = nT + int(threadIdx.x); i < nT + nH1; i += TILE_BLOCK_F64) {
    RecRegs r;
    const int js = fetch_record<D>(A, c, i, r);
    tile_pass_a<D, double>(A, c, ufast, r, js, i, phis, dds);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int i = int(threadIdx.x) + k * TILE_BLOCK_F64;
    if (i >= nT) continue;
    int j = tslot[k];
    double r;
    if (ufast && c.uniform(i)) {
      UniStencil<D, double> st;
      j = st.load(trec[k].w, A.geom);
      r = advect_value<D>(st, j, 0, vn[s + j], dt, phis, dds, DMAX);
    } else {
      const TileRec rec = record_at<D>(c, i, trec[k], j);
      const uint8_t w = uint8_t((rec.geo >> 24) & 63u);
      if (!(rec.geo & GEO_IRREGULAR)) {
        RegStencil<D, double> st;
        st.load(rec, A.geom);
        r = advect_value<D>(st, j, w, vn[s + j], dt, phis, dds, DMAX);
      } else {
        HangStencil<D, double> st;
        st.load(rec, c.ext_of(A.ext, i), A.geom);
        r = advect_value<D>(st, j, w, vn[s + j], dt, phis, dds, DMAX);
      }
    }
    // vn and the phi_n being averaged are read here rather than held in
    // registers through the passes (the 128-VGPR cap would spill them)
    if constexpr (RK_AVG)
      A.out[s + j] = 0.5 * (A.out[s + j] + r);
    else
      A.out[s + j] = r;
  }
}

// min over nodes of 0.8 |s_min / vn|, s_min the shortest |arm|; non-negative
// doubles order like their bit patterns, so an unsigned atomicMin reduces them
template <int D, class TB>
__global__ void __launch_bounds__(BLOCK) k_cfl(TB t, const double *__restrict__ vn,
                                               unsigned long long *__restrict__ result) {
  __shared__ double red[BLOCK];
  double m = DBL_MAX;
  for (int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; n < t.N; n += int64_t(gridDim.x) * blockDim.x) {
    double s_min = min2(fabs(double(t.d(0, n))), fabs(double(t.d(1, n))));
    s_min = min2(s_min, min2(fabs(double(t.d(2, n))), fabs(double(t.d(3, n)))));
    if constexpr (D == 3) s_min = min2(s_min, min2(fabs(double(t.d(4, n))), fabs(double(t.d(5, n)))));
    m = min2(m, 0.8 * fabs(s_min / vn[n]));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = BLOCK / 2; k > 0; k >>= 1) {
    if (int(threadIdx.x) < k) red[threadIdx.x] = min2(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMin(result, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

template <int D>
void run_advect(casl_stencil *st, double *phi, const double *vn, const double *vnp1, double dt, hipStream_t s) {
  const int64_t N = st->N;
  if (st->use_tiles(true)) {
    st->phi0.ensure(N);
    TileArgs<double> ta = tile_args<double>(st);
    const dim3 g(unsigned(st->n_tiles)), b(TILE_BLOCK_F64);
    ta.phin = phi;
    ta.out = st->phi0.p;
    prof_begin("k_tile_advect", s);
    hipLaunchKernelGGL((k_tile_advect<D, false>), g, b, 0, s, ta, vn, dt);
    prof_end("k_tile_advect", s);
    ta.phin = st->phi0.p;
    ta.out = phi;
    prof_begin("k_tile_advect", s);
    hipLaunchKernelGGL((k_tile_advect<D, true>), g, b, 0, s, ta, vnp1 ? vnp1 : vn, dt);
    prof_end("k_tile_advect", s);
    HIP_CHECK(hipGetLastError());
    return;
  }
  st->phi0.ensure(N);
  st->ddn64.ensure(size_t(D) * N);
  double *phin = st->phi0.p, *dd = st->ddn64.p;
  const unsigned g = grid_for(N);
  HIP_CHECK(hipMemcpyAsync(phin, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  st->with_tables<D>(true, [&](auto t) {
    using TB = decltype(t);
    prof_begin("k_adv_dd", s);
    hipLaunchKernelGGL((k_adv_dd<D, TB>), dim3(g), dim3(BLOCK), 0, s, t, phin, dd);
    prof_end("k_adv_dd", s);
    prof_begin("k_advect", s);
    hipLaunchKernelGGL((k_advect<D, false, TB>), dim3(g), dim3(BLOCK), 0, s, t, vn, dt, phin, dd, phi);
    prof_end("k_advect", s);
    prof_begin("k_adv_dd", s);
    hipLaunchKernelGGL((k_adv_dd<D, TB>), dim3(g), dim3(BLOCK), 0, s, t, phi, dd);
    prof_end("k_adv_dd", s);
    // phin <- (phi_n + stage(phi)) / 2, then back into phi
    prof_begin("k_advect", s);
    hipLaunchKernelGGL((k_advect<D, true, TB>), dim3(g), dim3(BLOCK), 0, s, t, vnp1 ? vnp1 : vn, dt, phi, dd,
                       phin);
    prof_end("k_advect", s);
  });
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(phi, phin, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
}

}  // namespace

extern "C" {

int casl_advect_in_normal_direction(casl_stencil_t *st, double *phi, const double *vn, const double *vn_np1,
                                    double dt, void *stream) {
  return guarded([&]() {
    if (!st || !phi || !vn) throw std::invalid_argument("null argument");
    if (!(dt >= 0.0)) throw std::invalid_argument("time step must be >= 0");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (st->dim == 2)
      run_advect<2>(st, phi, vn, vn_np1, dt, s);
    else
      run_advect<3>(st, phi, vn, vn_np1, dt, s);
  });
}

int casl_advect_cfl_dt(casl_stencil_t *st, const double *vn, double *dt, void *stream) {
  return guarded([&]() {
    if (!st || !vn || !dt) throw std::invalid_argument("null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    st->red.ensure(1);
    const double big = DBL_MAX;
    unsigned long long init;
    std::memcpy(&init, &big, sizeof(init));
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, st->red.p, init);
    const unsigned g = std::min<unsigned>(grid_for(st->N), 2048u);
    if (st->dim == 2)
      st->with_tables<2>(true, [&](auto t) {
        hipLaunchKernelGGL((k_cfl<2, decltype(t)>), dim3(g), dim3(BLOCK), 0, s, t, vn, st->red.p);
      });
    else
      st->with_tables<3>(true, [&](auto t) {
        hipLaunchKernelGGL((k_cfl<3, decltype(t)>), dim3(g), dim3(BLOCK), 0, s, t, vn, st->red.p);
      });
    HIP_CHECK(hipGetLastError());
    unsigned long long bits = 0;
    HIP_CHECK(hipMemcpyAsync(&bits, st->red.p, sizeof(bits), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(dt, &bits, sizeof(bits));
  });
}

}  // extern "C"
