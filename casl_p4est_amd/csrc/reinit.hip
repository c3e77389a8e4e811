// Level-set reinitialization on MI355X (gfx950): grid-wide kernels and the
// stencil part of the C ABI of include/casl_reinit.h.
//
// Per reinitialize_within_range_of_phi_0 call (src/my_p4est_level_set.cpp:290-557):
//   1. phi_0 = phi (device copy);  dd0 = second_derivatives_central(phi_0)
//   2. k_prep: everything that depends on phi_0 and the grid only -- the
//      anchoring / band / mask mode, the sign of phi_0 and, for nodes next to
//      the interface, the subcell arm lengths s (interface_Location*), which
//      the reference recomputes identically at every stage (:86-211).  FP64.
//   3. per pseudo-time stage: k_dd (pass A, second_derivatives_central of the
//      stage input) then k_stage (pass B, reinitialize_one_iteration :59-288);
//      the RK2 average phi_n = (phi_n + phi_np2)/2 (:497-498) is fused into
//      the second stage's store.
// The mixed path runs step 3 in FP32 (values, second derivatives, Godunov
// update) and keeps step 2 -- the subcell root finding where phi values
// cancel -- and the band error reduction in FP64.
#include <hip/hip_runtime.h>

#include <cmath>

#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"
#include "reinit_device.hpp"
#include "tiles_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

constexpr int BLOCK = 256;

inline unsigned grid_for(int64_t n, int block = BLOCK) { return unsigned((n + block - 1) / block); }

template <int D, class R, class TB>
__global__ void __launch_bounds__(BLOCK) k_dd(TB t, const R *__restrict__ f, R *__restrict__ fdd) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  node_dd<D, R>(t, n, f, fdd);
}

template <int D, class S, class TB>
__global__ void __launch_bounds__(BLOCK)
    k_prep(TB t, const double *__restrict__ phi0, const double *__restrict__ dd0,
           const uint8_t *__restrict__ mask, double lim_low, double lim_high, double zdt, int second_order,
           uint16_t *__restrict__ flags, S *__restrict__ s_fix) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  node_prep<D, S>(t, n, phi0, dd0, mask, lim_low, lim_high, zdt, second_order != 0, flags, s_fix);
}

// Pass B of one pseudo-time stage.  out[n] = result, or for the second RK2
// stage out[n] = 0.5*(out[n] + result): out then holds phi_n, each node reads
// and writes only its own entry, and phin is a different buffer.
template <int D, class R, bool SECOND_ORDER, bool RK_AVG, class TB>
__global__ void __launch_bounds__(BLOCK)
    k_stage(TB t, const uint16_t *__restrict__ flags, const R *__restrict__ s_fix,
            const double *__restrict__ phi0, const R *__restrict__ phin, const R *__restrict__ dd, R *out) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  const R r = node_stage<D, R, SECOND_ORDER>(t, n, flags[n], s_fix, phi0, phin, dd);
  if constexpr (RK_AVG)
    out[n] = R(0.5) * (out[n] + r);
  else
    out[n] = r;
}

__global__ void k_d2f(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = float(a[i]);
}
// The mixed path's way back to FP64.  Nodes the stages never update keep
// their FP64 input bit for bit, as in the reference: outside the phi_0 limits
// phi_np1 = phi_0 (src/my_p4est_level_set.cpp:277-278), masked nodes keep
// their current value, which is the input (phi_0) at every stage.
__device__ __forceinline__ double mixed_result(float v, uint16_t fl, double p0) {
  const int mode = fl & 3;
  return (mode == MODE_COPY0 || mode == MODE_MASKED) ? p0 : double(v);
}
// flags in node order (grid-wide path)
__global__ void k_f2d_keep(const float *__restrict__ a, const uint16_t *__restrict__ flags,
                           const double *__restrict__ phi0, double *__restrict__ b, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = mixed_result(a[i], flags[i], phi0[i]);
}
// flags in tile-record order (tiled path): one block per tile
template <int D>
__global__ void __launch_bounds__(BLOCK)
    k_tile_f2d_keep(TileArgs<float> T, const float *__restrict__ a, const uint16_t *__restrict__ flags,
                    const double *__restrict__ phi0, double *__restrict__ b) {
  TileCtx<D> c;
  c.load(T, int(blockIdx.x));
  for (int i = int(threadIdx.x); i < c.nT; i += BLOCK) {
    RecRegs r;
    const int js = fetch_record<D>(T, c, i, r);
    const int64_t n = int64_t(c.s) + slot_at<D>(c, i, r, js);
    b[n] = mixed_result(a[n], flags[c.s + i], phi0[n]);
  }
}
__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// max |a - b| over nodes with |ref| < band; non-negative doubles order like
// their bit patterns, so an unsigned 64-bit atomicMax reduces them.  A NaN
// difference counts as +inf (fmax would drop it), so a NaN in the band cannot
// hide behind a small error.
__global__ void __launch_bounds__(BLOCK) k_band_linf(const double *__restrict__ a, const double *__restrict__ b,
                                                     const double *__restrict__ ref, double band, int64_t n,
                                                     unsigned long long *__restrict__ result) {
  __shared__ double red[BLOCK];
  double m = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    if (fabs(ref[i]) < band) {
      const double d = fabs(a[i] - b[i]);
      m = fmax(m, d == d ? d : double(INFINITY));
    }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(result, (unsigned long long)__double_as_longlong(red[0]));
}

void require_device() {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) throw hip_error("no HIP device available");
}

}  // namespace

namespace casl {
namespace {
// The word of a float, or the high word of a double, that holds its exponent.
uint32_t exp_word(float x) {
  uint32_t b;
  std::memcpy(&b, &x, sizeof(b));
  return b;
}
uint32_t exp_word(double x) {
  uint64_t b;
  std::memcpy(&b, &x, sizeof(b));
  return uint32_t(b >> 32);
}
}  // namespace

// The tile arrays and geometry of a stencil's tiled encoding, in arithmetic R
// (per-call fields -- flags, s_fix, phi_0, phi in/out -- are the caller's).
template <class R>
dev::TileArgs<R> tile_args(const casl_stencil *st) {
  dev::TileArgs<R> ta{};
  ta.tile_start = st->t_tile_start.p;
  ta.halo_start = st->t_halo_start.p;
  ta.h1count = st->t_h1count.p;
  ta.halo = st->t_halo.p;
  ta.rec_start = st->t_rec_start.p;
  ta.rec = reinterpret_cast<const TileRec *>(st->t_rec.p);
  ta.rec_slot = st->t_rec_slot.p;
  ta.ext = st->t_ext.p;
  ta.ext_start = st->t_ext_start.p;
  ta.nreg = st->t_nreg.p;
  ta.urec = st->t_urec.p;
  ta.urec_start = st->t_urec_start.p;
  ta.nuni = st->t_nuni.p;
  ta.geom.pow2 = st->dyadic_geometry() ? 1 : 0;
  for (int a = 0; a < 3; ++a) {
    ta.geom.base[a] = R(st->geom_base[a]);
    ta.geom.unit[a] = R(st->geom_unit[a]);
    ta.geom.inv_base[a] = R(1.0 / st->geom_base[a]);
    ta.geom.base_bits[a] = exp_word(ta.geom.base[a]);
    ta.geom.inv_bits[a] = exp_word(ta.geom.inv_base[a]);
  }
  ta.geom.cubic = 1;
  for (int a = 1; a < st->dim; ++a) ta.geom.cubic &= int(st->geom_base[a] == st->geom_base[0]);
  ta.geom.d_pack = 0;
  {
    int e = 0;
    std::frexp(st->geom_base[0] / st->geom_unit[0], &e);
    ta.geom.lmax = e - 1;  // base / unit = 2^lmax, exact for any base
    int ex[3];
    for (int a = 0; a < 3; ++a) std::frexp(st->geom_base[a], &ex[a]);
    for (int a = 0; a < 3; ++a) {
      const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
      // cubic cells: zero; non-cubic ones come on a pow2 geometry only
      const int d1 = ta.geom.cubic ? 0 : ex[t1] - ex[a], d2 = ta.geom.cubic ? 0 : ex[t2] - ex[a];
      ta.geom.d_pack |= uint64_t(uint32_t((d1 + 128) & 255) | (uint32_t((d2 + 128) & 255) << 8)) << (16 * a);
    }
  }
  ta.N = st->N;
  return ta;
}
template dev::TileArgs<float> tile_args<float>(const casl_stencil *);
template dev::TileArgs<double> tile_args<double>(const casl_stencil *);
}  // namespace casl

namespace {

// reinitialize_within_range_of_phi_0 driver, templated on dimension and the
// stage arithmetic type R (double: FP64, float: mixed).

// The device buffers of one call: the FP64 copy of phi_0, the per-node flags
// and subcell distances (tile-record order on the tiled path), the stage
// buffers (mixed: the FP32 phi and its partner; FP64: phi itself and tmp)
// and, on the two-pass path, the pass-A outputs.
template <class R>
struct ReinitBufs {
  double *phi0 = nullptr;
  uint16_t *flags = nullptr;
  double *dd0 = nullptr;  // two-pass, second order: pass A of phi_0
  R *sfix = nullptr;      // [F][N]
  R *work = nullptr;      // mixed only: FP32 phi
  R *tmp = nullptr;
  R *ddn = nullptr;       // two-pass, second order: pass A of the stage input
};

// Sizes (elements) of the buffers a call of this kind uses; zero when unused.
struct ReinitSizes {
  size_t phi0, flags, dd0, sfix, work, tmp, ddn;
};
template <int D, class R>
ReinitSizes reinit_sizes(size_t N, bool so, bool tiled) {
  constexpr int F = 2 * D;
  ReinitSizes z{};
  z.phi0 = N;
  z.flags = N;
  z.dd0 = (so && !tiled) ? size_t(D) * N : 0;
  z.sfix = size_t(F) * N;
  z.work = sizeof(R) == 4 ? N : 0;
  z.tmp = N;
  z.ddn = (so && !tiled) ? size_t(D) * N : 0;
  return z;
}

// Every scratch buffer a reinitialization call of this kind uses, allocated
// up front in the stencil (a recorded graph must not allocate).
template <int D, class R>
ReinitBufs<R> reinit_scratch(casl_stencil *st, bool so, bool tiled) {
  const ReinitSizes z = reinit_sizes<D, R>(size_t(st->N), so, tiled);
  ReinitBufs<R> b;
  st->phi0.ensure(z.phi0);
  st->flags.ensure(z.flags);
  if (z.dd0) st->dd0.ensure(z.dd0);
  b.phi0 = st->phi0.p;
  b.flags = st->flags.p;
  b.dd0 = st->dd0.p;
  if constexpr (sizeof(R) == 8) {
    st->sfix64.ensure(z.sfix);
    st->tmp64.ensure(z.tmp);
    if (z.ddn) st->ddn64.ensure(z.ddn);
    b.sfix = st->sfix64.p;
    b.tmp = st->tmp64.p;
    b.ddn = st->ddn64.p;
  } else {
    st->sfixf.ensure(z.sfix);
    st->phif.ensure(z.work);
    st->tmpf.ensure(z.tmp);
    if (z.ddn) st->ddnf.ensure(z.ddn);
    b.sfix = st->sfixf.p;
    b.work = st->phif.p;
    b.tmp = st->tmpf.p;
    b.ddn = st->ddnf.p;
  }
  return b;
}

template <int D, class R>
void run_reinit(casl_stencil *st, double *phi, const uint8_t *mask, int order_space, int order_time, int n_iter,
                double limit_high, double limit_low, hipStream_t s, const ReinitBufs<R> &B) {
  const int64_t N = st->N;
  double lim_low, lim_high;
  widened_limits(limit_high, limit_low, st->zdt, lim_low, lim_high);
  const bool so = order_space > 1;
  const unsigned grid = grid_for(N);

  const bool tiled = st->use_tiles(sizeof(R) == 8);
  HIP_CHECK(hipMemcpyAsync(B.phi0, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  if (so && !tiled) {
    prof_begin("k_dd", s);
    st->with_tables<D>(true, [&](auto tt) {
      hipLaunchKernelGGL((k_dd<D, double, decltype(tt)>), dim3(grid), dim3(BLOCK), 0, s, tt, B.phi0, B.dd0);
    });
    prof_end("k_dd", s);
    HIP_CHECK(hipGetLastError());
  }
  R *sfix = B.sfix, *cur, *nxt = B.tmp, *ddn = B.ddn;
  if constexpr (sizeof(R) == 8) {
    cur = phi;
  } else {
    cur = B.work;
    hipLaunchKernelGGL(k_d2f, dim3(grid), dim3(BLOCK), 0, s, phi, B.work, N);
    HIP_CHECK(hipGetLastError());
  }
  if (tiled) {
    // flags and s_fix in tile-record order, the order the tiled stages read them
    TileArgs<double> tp = tile_args<double>(st);
    tp.phin = B.phi0;
    const dim3 g(unsigned(st->n_tiles)), b(TILE_BLOCK_F64);
    prof_begin("k_tile_prep", s);
    if (so)
      hipLaunchKernelGGL((k_tile_prep<D, R, true>), g, b, 0, s, tp, mask, lim_low, lim_high, st->zdt, B.flags,
                         sfix);
    else
      hipLaunchKernelGGL((k_tile_prep<D, R, false>), g, b, 0, s, tp, mask, lim_low, lim_high, st->zdt, B.flags,
                         sfix);
    prof_end("k_tile_prep", s);
  } else {
    prof_begin("k_prep", s);
    st->with_tables<D>(true, [&](auto tt) {
      hipLaunchKernelGGL((k_prep<D, R, decltype(tt)>), dim3(grid), dim3(BLOCK), 0, s, tt, B.phi0, B.dd0, mask,
                         lim_low, lim_high, st->zdt, int(so), B.flags, sfix);
    });
    prof_end("k_prep", s);
  }
  HIP_CHECK(hipGetLastError());

  // the FP64 stages read the FP64 distances when the stencil has them; the
  // mixed path's FP32 stages read the FP32 tables
  auto dd_pass = [&](const R *in) {
    prof_begin("k_dd", s);
    st->with_tables<D>(sizeof(R) == 8, [&](auto tt) {
      hipLaunchKernelGGL((k_dd<D, R, decltype(tt)>), dim3(grid), dim3(BLOCK), 0, s, tt, in, ddn);
    });
    prof_end("k_dd", s);
  };
  auto stage_launch = [&](const R *in, R *out, bool avg) {
    st->with_tables<D>(sizeof(R) == 8, [&](auto tt) {
      using TB = decltype(tt);
      if (so) {
        if (avg)
          hipLaunchKernelGGL((k_stage<D, R, true, true, TB>), dim3(grid), dim3(BLOCK), 0, s, tt, B.flags, sfix,
                             B.phi0, in, ddn, out);
        else
          hipLaunchKernelGGL((k_stage<D, R, true, false, TB>), dim3(grid), dim3(BLOCK), 0, s, tt, B.flags, sfix,
                             B.phi0, in, ddn, out);
      } else {
        if (avg)
          hipLaunchKernelGGL((k_stage<D, R, false, true, TB>), dim3(grid), dim3(BLOCK), 0, s, tt, B.flags, sfix,
                             B.phi0, in, ddn, out);
        else
          hipLaunchKernelGGL((k_stage<D, R, false, false, TB>), dim3(grid), dim3(BLOCK), 0, s, tt, B.flags, sfix,
                             B.phi0, in, ddn, out);
      }
    });
  };
  auto stage = [&](const R *in, R *out, bool avg) {
    prof_begin("k_stage", s);
    stage_launch(in, out, avg);
    prof_end("k_stage", s);
  };
  TileArgs<R> ta{};
  if (tiled) {
    ta = tile_args<R>(st);
    ta.flags = B.flags;
    ta.sfix = sfix;
    ta.phi0 = B.phi0;
  }
  auto tile_stage = [&](const R *in, R *out, bool avg) {
    ta.phin = in;
    ta.out = out;
    const dim3 g(unsigned(st->n_tiles)), b(stage_block<R>());
    prof_begin("k_tile_stage", s);
    if (so) {
      if (avg)
        hipLaunchKernelGGL((k_tile_stage<D, R, true, true>), g, b, 0, s, ta);
      else
        hipLaunchKernelGGL((k_tile_stage<D, R, true, false>), g, b, 0, s, ta);
    } else {
      if (avg)
        hipLaunchKernelGGL((k_tile_stage<D, R, false, true>), g, b, 0, s, ta);
      else
        hipLaunchKernelGGL((k_tile_stage<D, R, false, false>), g, b, 0, s, ta);
    }
    prof_end("k_tile_stage", s);
  };
  for (int it = 0; it < n_iter; ++it) {
    if (tiled) {
      tile_stage(cur, nxt, false);
      if (order_time > 1)
        tile_stage(nxt, cur, true);
      else
        std::swap(cur, nxt);
      HIP_CHECK(hipGetLastError());
      continue;
    }
    if (so) dd_pass(cur);
    stage(cur, nxt, false);  // phi_n -> phi_np1
    if (order_time > 1) {
      if (so) dd_pass(nxt);
      stage(nxt, cur, true);  // phi_n <- (phi_n + stage(phi_np1)) / 2, in place
    } else {
      std::swap(cur, nxt);  // memcpy(phi_n, phi_np1) (:521)
    }
    HIP_CHECK(hipGetLastError());
  }
  if constexpr (sizeof(R) == 8) {
    if (cur != phi) HIP_CHECK(hipMemcpyAsync(phi, cur, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  } else {
    if (tiled)
      hipLaunchKernelGGL(k_tile_f2d_keep<D>, dim3(unsigned(st->n_tiles)), dim3(BLOCK), 0, s, tile_args<float>(st),
                         cur, B.flags, B.phi0, phi);
    else
      hipLaunchKernelGGL(k_f2d_keep, dim3(grid), dim3(BLOCK), 0, s, cur, B.flags, B.phi0, phi, N);
    HIP_CHECK(hipGetLastError());
  }
}

void validate_tables(const casl_flat_stencil *tab) {
  if (!tab) throw std::invalid_argument("null stencil tables");
  if (tab->dim != 2 && tab->dim != 3) throw std::invalid_argument("dim must be 2 or 3");
  if (tab->n_nodes <= 0 || tab->n_nodes >= int64_t(INT32_MAX)) throw std::invalid_argument("bad n_nodes");
  if (!tab->nbr || !tab->d || !tab->d_tr || !tab->wall) throw std::invalid_argument("null table pointer");
  if (!(tab->zero_distance_threshold > 0)) throw std::invalid_argument("zero_distance_threshold must be > 0");
}

}  // namespace

namespace casl {

// Independent bounds check of every record of a tiled encoding before
// anything reaches the GPU: tiles within their LDS limits, halo nodes in
// range, record and extension ranges matching the per-tile counts, every slot
// used once, every corner slot readable, hanging faces valid.
void validate_tile_host(const TileHost &H, int D, int64_t n_nodes) {
  const int F = 2 * D, C = 1 << (D - 1);
  const int WORDS = D == 2 ? HangExt<2>::WORDS : HangExt<3>::WORDS, SW = D == 2 ? HangExt<2>::SW : HangExt<3>::SW;
  const int UW = D == 2 ? UniRec<2>::WORDS : UniRec<3>::WORDS;
  const int LMAX = D == 2 ? TileLimits<2>::LMAX : TileLimits<3>::LMAX;
  const int DMAX = D == 2 ? TileLimits<2>::DMAX : TileLimits<3>::DMAX;
  const int64_t nt = int64_t(H.tile_start.size()) - 1;
  if (H.tile_start.front() != 0 || H.tile_start.back() != n_nodes)
    throw std::runtime_error("tile ranges do not cover the nodes");
  if (int64_t(H.ext_start.size()) != nt + 1 || int64_t(H.nreg.size()) != nt || H.ext_start.front() != 0 ||
      size_t(H.ext_start.back()) * size_t(WORDS) != H.ext.size())
    throw std::runtime_error("extension ranges do not match the tiles");
  if (int64_t(H.urec_start.size()) != nt + 1 || int64_t(H.nuni.size()) != nt || H.urec_start.front() != 0 ||
      size_t(H.urec_start.back()) * size_t(UW) != H.urec.size() || H.rec_slot.size() != H.rec.size())
    throw std::runtime_error("record ranges do not match the tiles");
  for (int64_t t = 0; t < nt; ++t) {
    const int nT = H.tile_start[t + 1] - H.tile_start[t];
    const int nH = H.halo_start[t + 1] - H.halo_start[t];
    const int nH1 = H.h1count[t];
    const int nR = H.rec_start[t + 1] - H.rec_start[t];
    const int nU = H.urec_start[t + 1] - H.urec_start[t];
    const int nTu = int(H.nuni[t] & 0xffffu), nH1u = int(H.nuni[t] >> 16);
    if (nT <= 0 || nH1 < 0 || nH1 > nH || nT + nH > LMAX || nT + nH1 > DMAX || nR + nU != nT + nH1 ||
        nTu > nT || nH1u > nH1 || nTu + nH1u != nU)
      throw std::runtime_error("tile exceeds its LDS limits or its record counts disagree");
    for (int k = 0; k < nH; ++k) {
      const int32_t g = H.halo_idx[size_t(H.halo_start[t]) + k];
      if (g < 0 || g >= n_nodes) throw std::runtime_error("halo node out of range");
    }
    // non-uniform records: regular first within T and within H1; the extension
    // records of the irregular ones, in record order, fill [ext_start[t], ext_start[t+1])
    const int nTn = nT - nTu, nH1n = nH1 - nH1u;  // non-uniform records of T and of H1
    const int nTreg = int(H.nreg[t] & 0xffffu), nH1reg = int(H.nreg[t] >> 16);
    if (nTreg > nTn || nH1reg > nH1n || H.ext_start[t + 1] - H.ext_start[t] != (nTn - nTreg) + (nH1n - nH1reg))
      throw std::runtime_error("regular record counts do not match the extension range");
    std::vector<char> seen(size_t(nT + nH1), 0);
    // the tile's own records hold slots < nT, its H1 records slots nT .. nT + nH1 - 1, each exactly once
    auto take_slot = [&](int slot, bool own) {
      if (own ? (slot < 0 || slot >= nT) : (slot < nT || slot >= nT + nH1))
        throw std::runtime_error("record slot out of range");
      if (seen[size_t(slot)]++) throw std::runtime_error("record slot repeated");
    };
    for (int j = 0; j < nU; ++j) {
      const uint32_t *w = &H.urec[(size_t(H.urec_start[t]) + size_t(j)) * size_t(UW)];
      uint16_t idx[6];
      int slot = 0, level = 0;
      if (D == 2)
        uni_unpack<2>(w, idx, slot, level);
      else
        uni_unpack<3>(w, idx, slot, level);
      const bool own = j < nTu;
      take_slot(slot, own);
      // pass A of any record reads phi slots < nT + nH; pass B of a tile node
      // also reads second-derivative slots < nT + nH1 at the same corners
      const int bound = own ? nT + nH1 : nT + nH;
      for (int f = 0; f < F; ++f)
        if (idx[f] >= bound) throw std::runtime_error("corner slot out of range");
    }
    for (int j = 0; j < nR; ++j) {
      const TileRec &rr = H.rec[size_t(H.rec_start[t]) + j];
      const int slot = H.rec_slot[size_t(H.rec_start[t]) + j];
      const bool own = j < nTn;
      take_slot(slot, own);
      const bool irregular = (rr.geo & GEO_IRREGULAR) != 0;
      if (irregular != (own ? (j >= nTreg) : (j >= nTn + nH1reg)))
        throw std::runtime_error("record out of its regular/irregular group");
      const int bound = own ? nT + nH1 : nT + nH;
      for (int f = 0; f < F; ++f)
        if (rr.idx[f] >= bound) throw std::runtime_error("corner slot out of range");
      if (irregular) {
        const int rank = own ? j - nTreg : (nTn - nTreg) + (j - nTn - nH1reg);
        const uint16_t *e = &H.ext[(size_t(H.ext_start[t]) + size_t(rank)) * size_t(WORDS)];
        const int h0 = D == 2 ? e[0] : (e[0] & 15), h1 = D == 2 ? int(HANG_NONE) : ((e[0] >> 4) & 15);
        if (h0 >= F || (h1 != HANG_NONE && (h1 >= F || h1 <= h0))) throw std::runtime_error("bad hanging face");
        for (int q = 0; q < (h1 == HANG_NONE ? 1 : 2); ++q)
          for (int c = 0; c < C; ++c)
            if (e[1 + q * SW + c] >= bound) throw std::runtime_error("corner slot out of range");
      }
    }
  }
}

// Uploads a validated tiled encoding of n_nodes nodes into st (single geometry).
void upload_tile_host(const TileHost &H, int64_t n_nodes, casl_stencil *st) {
  if (H.tile_start.empty() || H.tile_start.back() != n_nodes) throw std::runtime_error("tile ranges do not cover the nodes");
  auto up = [](auto &buf, const auto &vec) {
    using T = typename std::decay_t<decltype(vec)>::value_type;
    buf.ensure(vec.size());
    HIP_CHECK(hipMemcpy(buf.p, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice));
  };
  up(st->t_tile_start, H.tile_start);
  up(st->t_halo_start, H.halo_start);
  up(st->t_h1count, H.h1count);
  up(st->t_halo, H.halo_idx.empty() ? std::vector<int32_t>(1, 0) : H.halo_idx);
  up(st->t_rec_start, H.rec_start);
  up(st->t_rec_slot, H.rec_slot.empty() ? std::vector<uint16_t>(1, 0) : H.rec_slot);
  st->t_rec.ensure(std::max<size_t>(H.rec.size(), 1));
  if (!H.rec.empty())
    HIP_CHECK(hipMemcpy(st->t_rec.p, H.rec.data(), H.rec.size() * sizeof(TileRec), hipMemcpyHostToDevice));
  up(st->t_ext, H.ext.empty() ? std::vector<uint16_t>(16, 0) : H.ext);
  up(st->t_ext_start, H.ext_start);
  up(st->t_nreg, H.nreg);
  up(st->t_urec, H.urec.empty() ? std::vector<uint32_t>(4, 0) : H.urec);
  up(st->t_urec_start, H.urec_start);
  up(st->t_nuni, H.nuni);
  st->n_tiles = int64_t(H.tile_start.size()) - 1;
  st->n_halo = int64_t(H.halo_idx.size());
  st->n_records = int64_t(H.rec.size()) + int64_t(H.urec_start.back());
  st->n_uniform = int64_t(H.urec_start.back());
  st->n_irregular = H.n_irregular;
  for (int a = 0; a < 3; ++a) {
    st->geom_base[a] = H.size[a];
    st->geom_unit[a] = H.size[a] * std::ldexp(1.0, -H.lmax);
  }
  st->tiled = true;
}

// Builds and uploads the tiled encoding; on failure the stencil keeps only
// the grid-wide tables (two-pass kernels) and records why.
void upload_tile_tables(const casl_flat_stencil *tab, casl_stencil *st) {
  st->tiled = false;
  TileHost H;
  try {
    build_tiles(*tab, H);
  } catch (const std::exception &e) {
    st->tile_error = e.what();
    return;
  }
  validate_tile_host(H, tab->dim, tab->n_nodes);
  upload_tile_host(H, tab->n_nodes, st);
  // One block per tile: a grid of fewer tiles than compute units leaves most
  // of the chip idle, and there the grid-wide kernels (more, smaller blocks)
  // measured faster (the 15-tile 2D star: 9.5e8 against 7.7e8 FP64
  // node-iterations/s).  casl_stencil_set_path overrides the choice.
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, st->device));
  st->path = st->n_tiles >= int64_t(prop.multiProcessorCount) ? CASL_PATH_TILED : CASL_PATH_TWO_PASS;
}

void upload_tables(const casl_flat_stencil *tab, casl_stencil *st, bool with_tiles, bool allow_fp64_copy) {
  validate_tables(tab);
  require_device();
  const int D = tab->dim;
  const int64_t N = tab->n_nodes;
  const int F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
  const size_t nidx = size_t(F) * C * N;
  for (size_t i = 0; i < nidx; ++i)
    if (tab->nbr[i] < 0 || tab->nbr[i] >= N)
      throw std::runtime_error("stencil node index out of range at entry " + std::to_string(i));
  const size_t nd = size_t(F) * N, nt = size_t(F) * T * N;
  // FP32 distances for every kernel; when FP32 does not hold them exactly (a
  // box whose edges are not dyadic) the FP64 kernels read an FP64 copy
  std::vector<float> dist(nd + nt);
  bool exact = true;
  for (size_t i = 0; i < nd; ++i) {
    dist[i] = float(tab->d[i]);
    exact = exact && double(dist[i]) == tab->d[i];
    if (!(tab->d[i] != 0.0)) throw std::runtime_error("zero face distance");
  }
  for (size_t i = 0; i < nt; ++i) {
    dist[nd + i] = float(tab->d_tr[i]);
    exact = exact && double(dist[nd + i]) == tab->d_tr[i];
  }
  if (!exact && !allow_fp64_copy) throw std::domain_error("stencil distance not exact in FP32");
  st->dim = D;
  st->N = N;
  st->zdt = tab->zero_distance_threshold;
  HIP_CHECK(hipGetDevice(&st->device));
  st->nbr.ensure(nidx);
  st->dist.ensure(dist.size());
  st->wall.ensure(N);
  HIP_CHECK(hipMemcpy(st->nbr.p, tab->nbr, nidx * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(st->dist.p, dist.data(), dist.size() * sizeof(float), hipMemcpyHostToDevice));
  st->wide = !exact;
  if (!exact) {
    st->dist64.ensure(nd + nt);
    HIP_CHECK(hipMemcpy(st->dist64.p, tab->d, nd * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(st->dist64.p + nd, tab->d_tr, nt * sizeof(double), hipMemcpyHostToDevice));
  }
  HIP_CHECK(hipMemcpy(st->wall.p, tab->wall, N * sizeof(uint8_t), hipMemcpyHostToDevice));
  if (with_tiles) upload_tile_tables(tab, st);
}

// ------------------------------------------------------------ kernel timing
namespace {
struct Profiler {
  std::mutex mu;
  std::atomic<bool> on{false};
  struct Pending {
    hipEvent_t a = nullptr, b = nullptr;
  };
  std::map<std::string, std::vector<Pending>> events;
  std::map<std::string, Pending> open;
  std::map<std::string, std::pair<double, int64_t>> done;

  void flush(const std::string &k) {
    auto &v = events[k];
    for (auto &p : v) {
      float ms = 0.f;
      if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        done[k].first += ms;
        done[k].second += 1;
      }
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    v.clear();
  }
};
Profiler &profiler() {
  static Profiler p;
  return p;
}
}  // namespace

// Set while a call is being recorded into a graph: no timing events then.
thread_local bool g_recording = false;

bool prof_on() { return profiler().on.load(std::memory_order_relaxed); }

void prof_begin(const char *kernel, hipStream_t s) {
  Profiler &P = profiler();
  if (!P.on || g_recording) return;
  std::lock_guard<std::mutex> g(P.mu);
  Profiler::Pending p;
  HIP_CHECK(hipEventCreate(&p.a));
  HIP_CHECK(hipEventCreate(&p.b));
  HIP_CHECK(hipEventRecord(p.a, s));
  P.open[kernel] = p;
}

void prof_end(const char *kernel, hipStream_t s) {
  Profiler &P = profiler();
  if (!P.on || g_recording) return;
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.open.find(kernel);
  if (it == P.open.end()) return;
  HIP_CHECK(hipEventRecord(it->second.b, s));
  auto &v = P.events[kernel];
  v.push_back(it->second);
  P.open.erase(it);
  if (v.size() >= 4096) P.flush(kernel);
}

}  // namespace casl

extern "C" {

int casl_profile_enable(int enable) {
  casl::profiler().on = (enable != 0);
  return CASL_OK;
}

int casl_profile_reset(void) {
  return guarded([&]() {
    Profiler &P = casl::profiler();
    std::lock_guard<std::mutex> g(P.mu);
    for (auto &kv : P.events) P.flush(kv.first);
    P.done.clear();
  });
}

int casl_profile_query(const char *kernel, double *total_ms, int64_t *launches) {
  return guarded([&]() {
    if (!kernel) throw std::invalid_argument("null kernel name");
    Profiler &P = casl::profiler();
    std::lock_guard<std::mutex> g(P.mu);
    P.flush(kernel);
    auto it = P.done.find(kernel);
    if (total_ms) *total_ms = it == P.done.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == P.done.end() ? 0 : it->second.second;
  });
}

int casl_stencil_create(const casl_flat_stencil *tab, casl_stencil_t **out) {
  return guarded([&]() {
    if (!out) throw std::invalid_argument("null out");
    *out = nullptr;
    auto st = std::make_unique<casl_stencil>();
    upload_tables(tab, st.get(), true, true);
    *out = st.release();
  });
}

int casl_stencil_destroy(casl_stencil_t *st) {
  delete st;
  return CASL_OK;
}

int casl_stencil_set_path(casl_stencil_t *st, int path) {
  return guarded([&]() {
    if (!st) throw std::invalid_argument("null stencil");
    if (path != CASL_PATH_TILED && path != CASL_PATH_TWO_PASS) throw std::invalid_argument("unknown path");
    st->path = path;
  });
}

int casl_stencil_get_path(const casl_stencil_t *st, int32_t *path) {
  return guarded([&]() {
    if (!st || !path) throw std::invalid_argument("null argument");
    *path = st->tiled ? st->path : CASL_PATH_TWO_PASS;
  });
}

int casl_stencil_tile_info(const casl_stencil_t *st, int32_t *tiled, int64_t *n_tiles, int64_t *n_halo,
                           int64_t *n_records, int64_t *n_irregular) {
  return guarded([&]() {
    if (!st) throw std::invalid_argument("null stencil");
    if (tiled) *tiled = st->tiled ? 1 : 0;
    if (n_tiles) *n_tiles = st->n_tiles;
    if (n_halo) *n_halo = st->n_halo;
    if (n_records) *n_records = st->n_records;
    if (n_irregular) *n_irregular = st->n_irregular;
    if (!st->tiled && !st->tile_error.empty()) set_error(st->tile_error);
  });
}

int casl_tile_plan(const casl_flat_stencil *tables, int64_t *n_tiles, int64_t *n_records, int64_t *n_irregular,
                   int64_t *n_irregular_nodes, int64_t *n_two_hanging) {
  return guarded([&]() {
    if (!tables) throw std::invalid_argument("null tables");
    validate_tables(tables);
    TileHost H;
    build_tiles(*tables, H);
    validate_tile_host(H, tables->dim, tables->n_nodes);
    const int64_t nt = int64_t(H.tile_start.size()) - 1;
    const int W = tables->dim == 2 ? HangExt<2>::WORDS : HangExt<3>::WORDS;
    int64_t irr_nodes = 0, two = 0;
    for (int64_t t = 0; t < nt; ++t)
      irr_nodes += (H.tile_start[t + 1] - H.tile_start[t]) - int64_t(H.nuni[t] & 0xffffu) - int64_t(H.nreg[t] & 0xffffu);
    if (tables->dim == 3)
      for (int64_t i = 0; i < H.n_irregular; ++i) two += ((H.ext[size_t(i) * W] >> 4) & 15) != HANG_NONE;
    if (n_tiles) *n_tiles = nt;
    if (n_records) *n_records = int64_t(H.rec.size()) + int64_t(H.urec_start.back());
    if (n_irregular) *n_irregular = H.n_irregular;
    if (n_irregular_nodes) *n_irregular_nodes = irr_nodes;
    if (n_two_hanging) *n_two_hanging = two;
  });
}

int casl_stencil_tile_check(const casl_stencil_t *st, const casl_flat_stencil *tables, int32_t *equal) {
  return guarded([&]() {
    if (!st || !tables || !equal) throw std::invalid_argument("null argument");
    validate_tables(tables);
    if (tables->n_nodes != st->N || tables->dim != st->dim) throw std::invalid_argument("tables of another grid");
    *equal = 0;
    TileHost H;
    try {
      build_tiles(*tables, H);
    } catch (const std::exception &e) {
      // neither encoding exists: equal
      if (!st->tiled) *equal = 1;
      else set_error(std::string("the host builder refuses these tables: ") + e.what());
      return;
    }
    if (!st->tiled) {
      set_error("the stencil has no tiled encoding: " + st->tile_error);
      return;
    }
    validate_tile_host(H, tables->dim, tables->n_nodes);
    auto same = [&](const char *name, const void *dev, const void *host, size_t bytes) {
      std::vector<unsigned char> d(bytes);
      if (bytes) HIP_CHECK(hipMemcpy(d.data(), dev, bytes, hipMemcpyDeviceToHost));
      if (bytes && std::memcmp(d.data(), host, bytes) != 0) {
        set_error(std::string("tiled encodings differ in ") + name);
        return false;
      }
      return true;
    };
    const int64_t nt = int64_t(H.tile_start.size()) - 1;
    if (st->n_tiles != nt || st->n_halo != int64_t(H.halo_idx.size()) ||
        st->n_records != int64_t(H.rec.size()) + int64_t(H.urec_start.back()) || st->n_irregular != H.n_irregular) {
      set_error("tiled encodings differ in their counts");
      return;
    }
    for (int a = 0; a < 3; ++a)
      if (st->geom_base[a] != H.size[a] || st->geom_unit[a] != H.size[a] * std::ldexp(1.0, -H.lmax)) {
        set_error("tiled encodings differ in their scales");
        return;
      }
    const bool ok =
        same("tile_start", st->t_tile_start.p, H.tile_start.data(), H.tile_start.size() * sizeof(int32_t)) &&
        same("halo_start", st->t_halo_start.p, H.halo_start.data(), H.halo_start.size() * sizeof(int32_t)) &&
        same("h1count", st->t_h1count.p, H.h1count.data(), H.h1count.size() * sizeof(int32_t)) &&
        same("halo", st->t_halo.p, H.halo_idx.data(), H.halo_idx.size() * sizeof(int32_t)) &&
        same("rec_start", st->t_rec_start.p, H.rec_start.data(), H.rec_start.size() * sizeof(int32_t)) &&
        same("rec", st->t_rec.p, H.rec.data(), H.rec.size() * sizeof(TileRec)) &&
        same("rec_slot", st->t_rec_slot.p, H.rec_slot.data(), H.rec_slot.size() * sizeof(uint16_t)) &&
        same("ext", st->t_ext.p, H.ext.data(), H.ext.size() * sizeof(uint16_t)) &&
        same("ext_start", st->t_ext_start.p, H.ext_start.data(), H.ext_start.size() * sizeof(int32_t)) &&
        same("nreg", st->t_nreg.p, H.nreg.data(), H.nreg.size() * sizeof(uint32_t)) &&
        same("urec", st->t_urec.p, H.urec.data(), H.urec.size() * sizeof(uint32_t)) &&
        same("urec_start", st->t_urec_start.p, H.urec_start.data(), H.urec_start.size() * sizeof(int32_t)) &&
        same("nuni", st->t_nuni.p, H.nuni.data(), H.nuni.size() * sizeof(uint32_t));
    *equal = ok ? 1 : 0;
  });
}

int casl_stencil_info(const casl_stencil_t *st, int32_t *dim, int64_t *n_nodes) {
  return guarded([&]() {
    if (!st) throw std::invalid_argument("null stencil");
    if (dim) *dim = st->dim;
    if (n_nodes) *n_nodes = st->N;
  });
}

int casl_second_derivatives(casl_stencil_t *st, const double *f, double *fdd, void *stream) {
  return guarded([&]() {
    if (!st || !f || !fdd) throw std::invalid_argument("null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (st->dim == 2)
      st->with_tables<2>(true, [&](auto t) {
        hipLaunchKernelGGL((k_dd<2, double, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, f, fdd);
      });
    else
      st->with_tables<3>(true, [&](auto t) {
        hipLaunchKernelGGL((k_dd<3, double, decltype(t)>), dim3(grid_for(st->N)), dim3(BLOCK), 0, s, t, f, fdd);
      });
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"

static void check_reinit_args(casl_stencil_t *st, double *phi, int order_space, int order_time, int n_iter,
                              double limit_high, double limit_low, int precision) {
  if (!st || !phi) throw std::invalid_argument("null argument");
  if (order_space == 0 || order_space > 2 || order_time == 0 || order_time > 2)
    throw std::invalid_argument(
        "reinitialize_within_range_of_phi_0: valid orders of accuracy for space and pseudotime are 1 and 2 only!");
  if (n_iter < 0) throw std::invalid_argument("negative number of iterations");
  if (!(limit_high > limit_low)) throw std::invalid_argument("limit_high must exceed limit_low");
  if (precision != CASL_PRECISION_FP64 && precision != CASL_PRECISION_MIXED)
    throw std::invalid_argument("unknown precision mode");
}

// One call on the stencil's own scratch buffers (allocated on first use);
// prepare_only: allocate them, launch nothing.
template <int D, class R>
static void reinit_on_stencil(casl_stencil_t *st, double *phi, const uint8_t *mask, int order_space,
                              int order_time, int n_iter, double limit_high, double limit_low, hipStream_t s,
                              bool prepare_only) {
  const bool so = order_space > 1, tiled = st->use_tiles(sizeof(R) == 8);
  const ReinitBufs<R> B = reinit_scratch<D, R>(st, so, tiled);
  if (!prepare_only) run_reinit<D, R>(st, phi, mask, order_space, order_time, n_iter, limit_high, limit_low, s, B);
}

static void dispatch_reinit(casl_stencil_t *st, double *phi, const uint8_t *mask, int order_space, int order_time,
                            int n_iter, double limit_high, double limit_low, int precision, hipStream_t s,
                            bool prepare_only = false) {
  const bool fp64 = precision == CASL_PRECISION_FP64;
  if (st->dim == 2) {
    if (fp64)
      reinit_on_stencil<2, double>(st, phi, mask, order_space, order_time, n_iter, limit_high, limit_low, s,
                                   prepare_only);
    else
      reinit_on_stencil<2, float>(st, phi, mask, order_space, order_time, n_iter, limit_high, limit_low, s,
                                  prepare_only);
  } else {
    if (fp64)
      reinit_on_stencil<3, double>(st, phi, mask, order_space, order_time, n_iter, limit_high, limit_low, s,
                                   prepare_only);
    else
      reinit_on_stencil<3, float>(st, phi, mask, order_space, order_time, n_iter, limit_high, limit_low, s,
                                  prepare_only);
  }
}

// One call with its buffers carved from [base, base + cap) instead of the
// stencil's own scratch: nothing is allocated.  The buffers are 256-byte
// aligned pieces in the order of ReinitBufs.
template <int D, class R>
static void reinit_in_buffer_t(casl_stencil_t *st, double *phi, int order_space, int order_time, int n_iter,
                               double limit_high, double limit_low, hipStream_t s, unsigned char *base, size_t cap) {
  const bool so = order_space > 1, tiled = st->use_tiles(sizeof(R) == 8);
  const ReinitSizes z = reinit_sizes<D, R>(size_t(st->N), so, tiled);
  const size_t bytes[7] = {z.phi0 * sizeof(double), z.flags * sizeof(uint16_t), z.dd0 * sizeof(double),
                           z.sfix * sizeof(R), z.work * sizeof(R), z.tmp * sizeof(R), z.ddn * sizeof(R)};
  size_t total = 0;
  for (size_t b : bytes) total += (b + 255) & ~size_t(255);
  if (total > cap)
    throw std::invalid_argument("reinitialization of " + std::to_string(st->N) + " nodes needs " +
                                std::to_string(total) + " bytes of scratch, more than the fixed " +
                                std::to_string(cap) + " bytes");
  unsigned char *p[7];
  size_t off = 0;
  for (int i = 0; i < 7; ++i) {
    p[i] = bytes[i] ? base + off : nullptr;
    off += (bytes[i] + 255) & ~size_t(255);
  }
  ReinitBufs<R> B;
  B.phi0 = reinterpret_cast<double *>(p[0]);
  B.flags = reinterpret_cast<uint16_t *>(p[1]);
  B.dd0 = reinterpret_cast<double *>(p[2]);
  B.sfix = reinterpret_cast<R *>(p[3]);
  B.work = reinterpret_cast<R *>(p[4]);
  B.tmp = reinterpret_cast<R *>(p[5]);
  B.ddn = reinterpret_cast<R *>(p[6]);
  run_reinit<D, R>(st, phi, nullptr, order_space, order_time, n_iter, limit_high, limit_low, s, B);
}

namespace casl {
void reinit_in_buffer(casl_stencil_t *st, double *phi, int order_space, int order_time, int n_iter,
                      double limit_high, double limit_low, int precision, hipStream_t s, unsigned char *base,
                      size_t cap) {
  check_reinit_args(st, phi, order_space, order_time, n_iter, limit_high, limit_low, precision);
  if (!base) throw std::invalid_argument("null scratch");
  // no timing events: creating them is the kind of call this path avoids
  const bool rec = g_recording;
  g_recording = true;
  try {
    const bool fp64 = precision == CASL_PRECISION_FP64;
    if (st->dim == 2)
      fp64 ? reinit_in_buffer_t<2, double>(st, phi, order_space, order_time, n_iter, limit_high, limit_low, s, base,
                                           cap)
           : reinit_in_buffer_t<2, float>(st, phi, order_space, order_time, n_iter, limit_high, limit_low, s, base,
                                          cap);
    else
      fp64 ? reinit_in_buffer_t<3, double>(st, phi, order_space, order_time, n_iter, limit_high, limit_low, s, base,
                                           cap)
           : reinit_in_buffer_t<3, float>(st, phi, order_space, order_time, n_iter, limit_high, limit_low, s, base,
                                          cap);
  } catch (...) {
    g_recording = rec;
    throw;
  }
  g_recording = rec;
}
}  // namespace casl

// ------------------------------------------------------ host-buffer calls
// phi in host memory (a PETSc Vec's array) reaches the device through the
// stencil's pinned staging buffer in chunks: the host threads copy chunk k
// into pinned memory while the DMA engine moves chunk k - 1 to the device,
// and on the way back each chunk is copied out as soon as its D2H transfer
// has finished.  Copies from pageable memory would go through the driver's
// own bounce buffer one piece at a time.
namespace {

int stage_chunks(int64_t N) { return int(std::max<int64_t>(1, std::min<int64_t>(16, N >> 20))); }

void chunk_range(int64_t N, int K, int k, int64_t &b, int64_t &e) {
  b = N * k / K;
  e = N * (k + 1) / K;
}

// pageable src -> pinned stage -> device dst, chunk by chunk on stream s
void stage_in(casl_stencil *st, const double *src, double *dst, int64_t N, hipStream_t s) {
  const int K = stage_chunks(N), T = std::max(1, std::min(num_threads(), 16));
  double *stage = st->host_stage.p;
  std::vector<std::atomic<int>> done(static_cast<size_t>(K));
  for (auto &d : done) d.store(0);
  std::vector<std::thread> pool;
  pool.reserve(size_t(T));
  for (int t = 0; t < T; ++t)
    pool.emplace_back([&, t]() {
      for (int k = 0; k < K; ++k) {
        int64_t b, e;
        chunk_range(N, K, k, b, e);
        const int64_t pb = b + (e - b) * t / T, pe = b + (e - b) * (t + 1) / T;
        if (pe > pb) std::memcpy(stage + pb, src + pb, size_t(pe - pb) * sizeof(double));
        done[size_t(k)].fetch_add(1, std::memory_order_release);
      }
    });
  hipError_t err = hipSuccess;
  for (int k = 0; k < K; ++k) {
    while (done[size_t(k)].load(std::memory_order_acquire) < T) std::this_thread::yield();
    int64_t b, e;
    chunk_range(N, K, k, b, e);
    if (err == hipSuccess && e > b)
      err = hipMemcpyAsync(dst + b, stage + b, size_t(e - b) * sizeof(double), hipMemcpyHostToDevice, s);
  }
  for (auto &th : pool) th.join();
  if (err != hipSuccess) throw hip_error(std::string("staging copy to the device failed: ") + hipGetErrorString(err));
}

// device src -> pinned stage -> pageable dst; returns when dst is written
void stage_out(casl_stencil *st, const double *src, double *dst, int64_t N, hipStream_t s) {
  const int K = stage_chunks(N), T = std::max(1, std::min(num_threads(), 16));
  double *stage = st->host_stage.p;
  for (int k = 0; k < K; ++k) {
    int64_t b, e;
    chunk_range(N, K, k, b, e);
    HIP_CHECK(hipMemcpyAsync(stage + b, src + b, size_t(e - b) * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipEventRecord(st->stage_events[size_t(k)], s));
  }
  std::atomic<int> bad{0};
  std::vector<std::thread> pool;
  pool.reserve(size_t(T));
  for (int t = 0; t < T; ++t)
    pool.emplace_back([&, t]() {
      for (int k = 0; k < K; ++k) {
        if (hipEventSynchronize(st->stage_events[size_t(k)]) != hipSuccess) {
          bad.store(1);
          return;
        }
        int64_t b, e;
        chunk_range(N, K, k, b, e);
        const int64_t pb = b + (e - b) * t / T, pe = b + (e - b) * (t + 1) / T;
        if (pe > pb) std::memcpy(dst + pb, stage + pb, size_t(pe - pb) * sizeof(double));
      }
    });
  for (auto &th : pool) th.join();
  if (bad.load()) throw hip_error("staging copy from the device failed");
}

void ensure_staging(casl_stencil *st) {
  const size_t N = size_t(st->N);
  st->host_phi.ensure(N);
  st->host_stage.ensure(N);
  const size_t K = size_t(stage_chunks(st->N));
  while (st->stage_events.size() < K) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    st->stage_events.push_back(e);
  }
}

}  // namespace

extern "C" {

int casl_reinit_host(casl_stencil_t *st, double *phi, const uint8_t *mask, int order_space, int order_pseudotime,
                     int n_iter, double limit_high, double limit_low, int precision, void *stream) {
  return guarded([&]() {
    check_reinit_args(st, phi, order_space, order_pseudotime, n_iter, limit_high, limit_low, precision);
    if (mask && (order_space != 2 || order_pseudotime != 2))
      throw std::invalid_argument("reinitialize_2nd_order_with_mask: a mask comes with orders 2 and 2 only");
    hipStream_t s = static_cast<hipStream_t>(stream);
    ensure_staging(st);
    const uint8_t *dmask = nullptr;
    if (mask) {
      st->host_mask.ensure(size_t(st->N));
      st->host_mask_stage.ensure(size_t(st->N));
      std::memcpy(st->host_mask_stage.p, mask, size_t(st->N));
      HIP_CHECK(hipMemcpyAsync(st->host_mask.p, st->host_mask_stage.p, size_t(st->N), hipMemcpyHostToDevice, s));
      dmask = st->host_mask.p;
    }
    stage_in(st, phi, st->host_phi.p, st->N, s);
    dispatch_reinit(st, st->host_phi.p, dmask, order_space, order_pseudotime, n_iter, limit_high, limit_low,
                    precision, s);
    HIP_CHECK(hipGetLastError());
    stage_out(st, st->host_phi.p, phi, st->N, s);
  });
}

static int reinit_common(casl_stencil_t *st, double *phi, const uint8_t *mask, int order_space, int order_time,
                         int n_iter, double limit_high, double limit_low, int precision, void *stream) {
  return guarded([&]() {
    check_reinit_args(st, phi, order_space, order_time, n_iter, limit_high, limit_low, precision);
    dispatch_reinit(st, phi, mask, order_space, order_time, n_iter, limit_high, limit_low, precision,
                    static_cast<hipStream_t>(stream));
  });
}

int casl_reinit(casl_stencil_t *st, double *phi, int order_space, int order_pseudotime, int n_iter, double limit_high,
                double limit_low, int precision, void *stream) {
  return reinit_common(st, phi, nullptr, order_space, order_pseudotime, n_iter, limit_high, limit_low, precision,
                       stream);
}

int casl_reinit_masked(casl_stencil_t *st, double *phi, const uint8_t *mask, int n_iter, double limit_high,
                       double limit_low, int precision, void *stream) {
  if (!mask) {
    set_error("null mask");
    return CASL_ERR_INVALID_ARGUMENT;
  }
  return reinit_common(st, phi, mask, 2, 2, n_iter, limit_high, limit_low, precision, stream);
}

int casl_reinit_graph(casl_stencil_t *st, double *phi, int order_space, int order_pseudotime, int n_iter,
                      double limit_high, double limit_low, int precision, void *stream) {
  return guarded([&]() {
    check_reinit_args(st, phi, order_space, order_pseudotime, n_iter, limit_high, limit_low, precision);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // what the recording depends on: the call's arguments, the stage path and
    // every scratch address the recorded kernels read or write
    dispatch_reinit(st, phi, nullptr, order_space, order_pseudotime, n_iter, limit_high, limit_low, precision, s,
                    true);
    uint64_t lh, ll;
    std::memcpy(&lh, &limit_high, sizeof(lh));
    std::memcpy(&ll, &limit_low, sizeof(ll));
    const std::vector<uintptr_t> key = {
        uintptr_t(phi), uintptr_t(order_space), uintptr_t(order_pseudotime), uintptr_t(n_iter), uintptr_t(lh),
        uintptr_t(ll), uintptr_t(precision), uintptr_t(st->path), uintptr_t(st->phi0.p), uintptr_t(st->flags.p),
        uintptr_t(st->dd0.p), uintptr_t(st->sfix64.p), uintptr_t(st->tmp64.p), uintptr_t(st->ddn64.p),
        uintptr_t(st->sfixf.p), uintptr_t(st->phif.p), uintptr_t(st->tmpf.p), uintptr_t(st->ddnf.p)};
    if (!st->graph_exec || st->graph_key != key) {
      st->drop_graph();
      hipStream_t cap;
      HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      g_recording = true;
      try {
        HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
          dispatch_reinit(st, phi, nullptr, order_space, order_pseudotime, n_iter, limit_high, limit_low, precision,
                          cap);
        } catch (...) {
          hipGraph_t broken = nullptr;
          (void)hipStreamEndCapture(cap, &broken);
          if (broken) (void)hipGraphDestroy(broken);
          throw;
        }
        HIP_CHECK(hipStreamEndCapture(cap, &st->graph));
        HIP_CHECK(hipGraphInstantiate(&st->graph_exec, st->graph, nullptr, nullptr, 0));
      } catch (...) {
        g_recording = false;
        (void)hipStreamDestroy(cap);
        st->drop_graph();
        throw;
      }
      g_recording = false;
      HIP_CHECK(hipStreamDestroy(cap));
      st->graph_key = key;
    }
    HIP_CHECK(hipGraphLaunch(st->graph_exec, s));
  });
}

int casl_band_linf(casl_stencil_t *st, const double *a, const double *b, const double *ref, double band,
                   double *result, void *stream) {
  return guarded([&]() {
    if (!st || !a || !b || !ref || !result) throw std::invalid_argument("null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    st->red.ensure(1);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, st->red.p);
    const unsigned grid = std::min<unsigned>(grid_for(st->N), 2048);
    hipLaunchKernelGGL(k_band_linf, dim3(grid), dim3(BLOCK), 0, s, a, b, ref, band, st->N, st->red.p);
    HIP_CHECK(hipGetLastError());
    unsigned long long bits = 0;
    HIP_CHECK(hipMemcpyAsync(&bits, st->red.p, sizeof(bits), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *result = v;
  });
}

}  // extern "C"
