// Per-node bodies of the reinitialization passes, shared by the grid-wide
// kernels (reinit.hip) and the persistent batched kernel (batch.hip), plus
// the host-side device-buffer helpers.  Paths are relative to the reference
// tree.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/casl_reinit.h"
#include "stencil_math.hpp"
#include "tiles.hpp"

namespace casl {
void set_error(const std::string &msg);

struct hip_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_CHECK(expr)                                                                         \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess)                                                                       \
      throw ::casl::hip_error(std::string(#expr) + " failed: " + hipGetErrorString(_e));        \
  } while (0)

// The library's device allocations (devmem.hip): hipMalloc, or a block of
// the opt-in cache of released blocks.  device_alloc returns the block's
// size (at least bytes) and device; device_free takes both back.
void *device_alloc(size_t bytes, size_t *granted, int *device);
void device_free(void *p, size_t bytes, int device);

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  size_t bytes = 0;  // the block's size (at least n * sizeof(T))
  int device = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) device_free(p, bytes, device);
    p = nullptr;
    n = 0;
    bytes = 0;
  }
  void ensure(size_t count) {
    if (n >= count && p) return;
    release();
    if (count == 0) return;
    p = static_cast<T *>(device_alloc(count * sizeof(T), &bytes, &device));
    n = count;
  }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(bytes, o.bytes);
    std::swap(device, o.device);
  }
};

// Page-locked host memory (hipHostMalloc), grown on demand.
template <class T>
struct PinnedBuf {
  T *p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (n >= count && p) return;
    release();
    if (count == 0) return;
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      if (e == hipErrorOutOfMemory) throw std::bad_alloc();
      throw hip_error(std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    n = count;
  }
};

// Maps exceptions to the C ABI status codes.
template <class Fn>
int guarded(Fn fn) {
  try {
    fn();
    return CASL_OK;
  } catch (const std::invalid_argument &e) {
    set_error(e.what());
    return CASL_ERR_INVALID_ARGUMENT;
  } catch (const std::bad_alloc &) {
    set_error("device or host allocation failed");
    return CASL_ERR_OUT_OF_MEMORY;
  } catch (const hip_error &e) {
    set_error(e.what());
    return CASL_ERR_HIP;
  } catch (const std::domain_error &e) {
    set_error(e.what());
    return CASL_ERR_INEXACT_DISTANCES;
  } catch (const std::exception &e) {
    set_error(e.what());
    return CASL_ERR_INVALID_STENCIL;
  }
}

inline double largest_dbl_smaller_than_dbl_max() {  // src/my_p4est_utils.h:322-330
  volatile double p = 1.0;
  while (DBL_MAX - p == DBL_MAX) p *= 2.0;
  return DBL_MAX - p;
}

// Widened phi_0 limits of reinitialize_one_iteration (src/my_p4est_level_set.cpp:75-76).
inline void widened_limits(double limit_high, double limit_low, double zdt, double &lo, double &hi) {
  const double ldbl = largest_dbl_smaller_than_dbl_max();
  lo = (limit_low < -ldbl ? -DBL_MAX : limit_low - dev::max2(zdt, dev::EPS * std::fabs(limit_low)));
  hi = (limit_high > ldbl ? DBL_MAX : limit_high + dev::max2(zdt, dev::EPS * std::fabs(limit_high)));
}

namespace dev {

// flags[n] (uint16):
//   bits 0-1  mode: 0 update, 1 anchor (phi = 0), 2 outside limits (phi = phi_0), 3 masked (phi = phi_n)
//   bit  2    phi_0 > 0
//   bits 3-8  subcell crossing on face f (bit 3 + f)
enum : uint16_t { MODE_UPDATE = 0, MODE_ANCHOR = 1, MODE_COPY0 = 2, MODE_MASKED = 3 };

// Pass A at one node: second_derivatives_central (laplace) of f.
template <int D, class R, class TB>
__device__ __forceinline__ void node_dd(const TB &t, int64_t n, const R *__restrict__ f, R *__restrict__ fdd) {
  R f000, fnb[2 * D], dd[D];
  laplace<D, R>(t, n, f, f000, fnb, dd);
#pragma unroll
  for (int a = 0; a < D; ++a) fdd[int64_t(a) * t.N + n] = dd[a];
}

// The phi_0-only part of reinitialize_one_iteration (src/my_p4est_level_set.cpp:82-211):
// mode, sign of phi_0 and the subcell arm lengths of crossing arms,
// s_fix[f] = MAX(interface_Location*(..), zero_distance_threshold).
// The reference recomputes these identically at every stage.  Returns the
// node's flags; writes s_fix of its crossing arms at sfix_m[f * sfix_stride].
//   t, j      stencil accessor and the node's index in it and in phi0 / dd0
//   masked    the node is outside the mask
//   dd0       pass A of phi_0, component a at dd0[a * dd_stride + index]
//   OWN_DD    the node's pass-A values are dd0[a * dd_stride + j]
template <int D, class S, class TB, bool OWN_DD = false>
__device__ __forceinline__ uint16_t prep_value(const TB &t, int64_t j, bool masked, const double *__restrict__ phi0,
                                               const double *__restrict__ dd0, int64_t dd_stride, double lim_low,
                                               double lim_high, double zdt, bool second_order,
                                               S *__restrict__ sfix_m, int64_t sfix_stride) {
  const double p = phi0[j];
  if (masked) return MODE_MASKED;
  if (fabs(p) < zdt) return MODE_ANCHOR;
  if (!(lim_low < p && p < lim_high)) return MODE_COPY0;
  constexpr int F = 2 * D;
  double p0_000, p0_f[F], dloc[D];
  if constexpr (OWN_DD) {
#pragma unroll
    for (int a = 0; a < D; ++a) dloc[a] = dd0[int64_t(a) * dd_stride + j];
    ngbd_quadratic_given_dd<D, double>(t, j, phi0, dloc, p0_000, p0_f);
  } else {
    ngbd_quadratic<D, double>(t, j, phi0, p0_000, p0_f, dloc);
  }
  uint16_t fl = MODE_UPDATE | (p0_000 > 0.0 ? 4 : 0);
  unsigned cross = 0;
#pragma unroll
  for (int f = 0; f < F; ++f)
    if (p0_000 * p0_f[f] < 0.0) cross |= 1u << f;
  if (cross) {
#pragma unroll
    for (int f = 0; f < F; ++f) {
      if (!(cross & (1u << f))) continue;
      const int a = f >> 1;
      double s = double(t.d(f, j));
      if (second_order) {
        const double dd_000 = dd0[int64_t(a) * dd_stride + j];
        const double dd_f = face_linear<D, double>(t, f, j, dd0 + int64_t(a) * dd_stride);
        if ((f & 1) == 0)
          s = -interface_location_2nd(-s, 0.0, p0_f[f], p0_000, dd_f, dd_000);
        else
          s = interface_location_2nd(0.0, s, p0_000, p0_f[f], dd_000, dd_f);
      } else {
        if ((f & 1) == 0)
          s = -interface_location(-s, 0.0, p0_f[f], p0_000);
        else
          s = interface_location(0.0, s, p0_000, p0_f[f]);
      }
      sfix_m[int64_t(f) * sfix_stride] = S(max2(s, zdt));
    }
    fl |= uint16_t(cross << 3);
  }
  return fl;
}

// prep_value on the grid-wide tables, flags and s_fix in node order
template <int D, class S, class TB>
__device__ __forceinline__ void node_prep(const TB &t, int64_t n, const double *__restrict__ phi0,
                                          const double *__restrict__ dd0, const uint8_t *__restrict__ mask,
                                          double lim_low, double lim_high, double zdt, bool second_order,
                                          uint16_t *__restrict__ flags, S *__restrict__ s_fix) {
  flags[n] = prep_value<D, S>(t, n, mask && mask[n] == 0, phi0, dd0, t.N, lim_low, lim_high, zdt, second_order,
                              s_fix + n, t.N);
}

// Pass B at one node: the stage result of reinitialize_one_iteration
// (src/my_p4est_level_set.cpp:82-281) given the prepared flags / arm lengths.
//   t, j      stencil accessor and the node's index in it and in phin / dd
//   fl, w     flags and wall bits of the node
//   sfix_n    s_fix + n (entries a stride sfix_stride apart, one per face)
//   phi0_n    phi_0 + n (read only for nodes outside the limits)
//   dd        pass-A output, component a at dd[a * dd_stride + index]
//   OWN_DD    the node's pass-A values are taken from dd[a * dd_stride + j]
//             (they are the laplace of this node by the same arithmetic)
template <int D, class R, bool SECOND_ORDER, class TB, bool OWN_DD = false>
__device__ __forceinline__ R stage_value(const TB &t, int64_t j, uint16_t fl, uint8_t w,
                                         const R *__restrict__ sfix_n, int64_t sfix_stride,
                                         const double *__restrict__ phi0_n, const R *__restrict__ phin,
                                         const R *__restrict__ dd, int64_t dd_stride) {
  constexpr int F = 2 * D;
  const int mode = fl & 3;
  if (mode == MODE_ANCHOR) return R(0.0);
  if (mode == MODE_COPY0) return R(*phi0_n);
  if (mode == MODE_MASKED) return phin[j];
  R pn_000, pn_f[F], pn_dd_000[D];
  // ngbd_with_quadratic_interpolation of phi_n; its laplace is the node's own
  // pass-A value (same arithmetic), so phi_n_xx_000 comes from there.
  if constexpr (OWN_DD && SECOND_ORDER) {
#pragma unroll
    for (int a = 0; a < D; ++a) pn_dd_000[a] = dd[int64_t(a) * dd_stride + j];
    ngbd_quadratic_given_dd<D, R>(t, j, phin, pn_dd_000, pn_000, pn_f);
  } else {
    ngbd_quadratic<D, R>(t, j, phin, pn_000, pn_f, pn_dd_000);
  }
  R pn_dd_f[F];
  if constexpr (SECOND_ORDER) face_values_dd<D, R>(t, j, dd, dd_stride, pn_dd_f);
  R s[F];
#pragma unroll
  for (int f = 0; f < F; ++f) s[f] = R(t.d(f, j));
  const unsigned cross = (fl >> 3) & ((1u << F) - 1);
  if (cross) {
#pragma unroll
    for (int f = 0; f < F; ++f)
      if (cross & (1u << f)) {
        s[f] = sfix_n[int64_t(f) * sfix_stride];
        pn_f[f] = R(0.0);
      }
    // MAX(s, zero_distance_threshold) of the non-crossing arms only changes
    // negative wall arms, which the Neumann mirror below overwrites.
  }
  // Neumann boundary condition on the walls (:216-225); nodes off the walls
  // (all but a thin layer) skip the selects
  if (w) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int fm = 2 * a, fp = 2 * a + 1;
      if (w & (1u << fm)) {
        s[fm] = s[fp];
        pn_f[fm] = pn_f[fp];
        if constexpr (SECOND_ORDER) pn_dd_000[a] = pn_dd_f[fm] = pn_dd_f[fp] = R(0.0);
      }
      if (w & (1u << fp)) {
        s[fp] = s[fm];
        pn_f[fp] = pn_f[fm];
        if constexpr (SECOND_ORDER) pn_dd_000[a] = pn_dd_f[fm] = pn_dd_f[fp] = R(0.0);
      }
    }
  }
  // one-sided differences with ENO/minmod correction (:230-248)
  R Dp[D], Dm[D];
  bool divided = false;
  if constexpr (TB::FASTDIV) {
    if (t.fast) {
      // no wall arm: every non-crossing s[f] is the power-of-two arm length,
      // so dividing by it is multiplying by its exact inverse
#pragma unroll
      for (int a = 0; a < D; ++a) {
        Dp[a] = (pn_f[2 * a + 1] - pn_000) * t.iq[2 * a + 1];
        Dm[a] = (pn_000 - pn_f[2 * a]) * t.iq[2 * a];
      }
      if (cross) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          if (cross & (2u << (2 * a))) Dp[a] = qdiv(pn_f[2 * a + 1] - pn_000, s[2 * a + 1]);
          if (cross & (1u << (2 * a))) Dm[a] = qdiv(pn_000 - pn_f[2 * a], s[2 * a]);
        }
      }
      divided = true;
    }
  }
  if (!divided) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      Dp[a] = qdiv(pn_f[2 * a + 1] - pn_000, s[2 * a + 1]);
      Dm[a] = qdiv(pn_000 - pn_f[2 * a], s[2 * a]);
    }
  }
  if constexpr (SECOND_ORDER) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const R ddm = minmod(pn_dd_f[2 * a], pn_dd_000[a]);
      Dm[a] = qfma(R(0.5) * s[2 * a], ddm, Dm[a]);
      const R ddp = minmod(pn_dd_f[2 * a + 1], pn_dd_000[a]);
      Dp[a] = qfma(R(-0.5) * s[2 * a + 1], ddp, Dp[a]);
    }
  }
  // local time step and Godunov upwinding (:250-271)
  const R sgn = (fl & 4) ? R(1.0) : R(-1.0);
  // arm lengths are positive: fmin is MIN's value, in one instruction
  R dt = fmin(fmin(s[0], s[1]), fmin(s[2], s[3]));
  if constexpr (D == 3) dt = fmin(dt, fmin(s[4], s[5]));
  // FP64 keeps the reference's division; in FP32 (the mixed path, held
  // to 1e-3 h of FP64) a multiplication by 1/D differs by an ulp at most
  if constexpr (sizeof(R) == 4)
    dt *= R(1.0 / D);
  else
    dt /= R(D);
  // upwinding: D+ is dropped where sgn*D+ > 0 and D- where sgn*D- < 0; only
  // the squares enter, and (sgn*x)^2 == x^2 exactly, so clamp sgn*x instead
#pragma unroll
  for (int a = 0; a < D; ++a) {
    Dp[a] = fmin(sgn * Dp[a], R(0.0));
    Dm[a] = fmax(sgn * Dm[a], R(0.0));
  }
  R sum = fmax(Dp[0] * Dp[0], Dm[0] * Dm[0]) + fmax(Dp[1] * Dp[1], Dm[1] * Dm[1]);
  if constexpr (D == 3) sum = sum + fmax(Dp[2] * Dp[2], Dm[2] * Dm[2]);
  R r = qfma(-(dt * sgn), qsqrt(sum) - R(1.), pn_000);
  // sign guard phi_0*phi_np1 < 0 -> flip (:273-274); phi_0 != 0 for updated nodes
  if ((fl & 4) ? (r < R(0.0)) : (r > R(0.0))) r *= R(-1.0);
  return r;
}

// One stage of advect_in_normal_direction_one_iteration at one node
// (src/my_p4est_level_set.cpp:559-652): quadratic neighbor values of phi_n,
// Neumann mirror on walls, one-sided differences with the ENO/minmod
// correction, upwinding by the sign of vn, and
// phi_n - dt vn sqrt(sum of both one-sided squares).  FP64.
template <int D, class TB>
__device__ __forceinline__ double advect_value(const TB &t, int64_t j, uint8_t w, double vn, double dt,
                                               const double *__restrict__ phin, const double *__restrict__ dd,
                                               int64_t dd_stride) {
  constexpr int F = 2 * D;
  double pn_000, pn_f[F], dloc[D];
  ngbd_quadratic<D, double>(t, j, phin, pn_000, pn_f, dloc);
  double s[F], dd_000[D], dd_f[F];
#pragma unroll
  for (int f = 0; f < F; ++f) s[f] = double(t.d(f, j));
#pragma unroll
  for (int a = 0; a < D; ++a) dd_000[a] = dd[int64_t(a) * dd_stride + j];
  face_values_dd<D, double>(t, j, dd, dd_stride, dd_f);
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const int fm = 2 * a, fp = 2 * a + 1;
    if (w & (1u << fm)) {
      s[fm] = s[fp];
      pn_f[fm] = pn_f[fp];
      dd_000[a] = dd_f[fm] = dd_f[fp] = 0.0;
    }
    if (w & (1u << fp)) {
      s[fp] = s[fm];
      pn_f[fp] = pn_f[fm];
      dd_000[a] = dd_f[fm] = dd_f[fp] = 0.0;
    }
  }
  double Dp[D], Dm[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    Dp[a] = (pn_f[2 * a + 1] - pn_000) / s[2 * a + 1];
    Dm[a] = (pn_000 - pn_f[2 * a]) / s[2 * a];
  }
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double ddm = minmod(dd_f[2 * a], dd_000[a]);
    Dm[a] += 0.5 * s[2 * a] * ddm;
    const double ddp = minmod(dd_f[2 * a + 1], dd_000[a]);
    Dp[a] -= 0.5 * s[2 * a + 1] * ddp;
  }
  const double sgn = vn > 0.0 ? 1.0 : -1.0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    if (sgn * Dp[a] > 0.0) Dp[a] = 0.0;
    if (sgn * Dm[a] < 0.0) Dm[a] = 0.0;
  }
  double sum = (Dp[0] * Dp[0] + Dm[0] * Dm[0]) + (Dp[1] * Dp[1] + Dm[1] * Dm[1]);
  if constexpr (D == 3) sum = sum + (Dp[2] * Dp[2] + Dm[2] * Dm[2]);
  return pn_000 - dt * vn * sqrt(sum);
}

// stage_value on the grid-wide tables
template <int D, class R, bool SECOND_ORDER, class TB>
__device__ __forceinline__ R node_stage(const TB &t, int64_t n, uint16_t fl, const R *__restrict__ s_fix,
                                        const double *__restrict__ phi0, const R *__restrict__ phin,
                                        const R *__restrict__ dd) {
  return stage_value<D, R, SECOND_ORDER>(t, n, fl, t.wall[n], s_fix + n, t.N, phi0 + n, phin, dd, t.N);
}

}  // namespace dev
}  // namespace casl

// Device tables and scratch of one stencil (opaque casl_stencil_t).
struct casl_stencil {
  int dim = 0;
  int64_t N = 0;
  int device = 0;
  double zdt = 0;
  casl::DevBuf<int32_t> nbr;
  casl::DevBuf<float> dist;
  // FP64 copy of the distances when FP32 does not hold them exactly (a box
  // whose edges are not dyadic): the FP64 grid-wide kernels read it
  casl::DevBuf<double> dist64;
  bool wide = false;
  casl::DevBuf<uint8_t> wall;
  // scratch, grown on demand
  casl::DevBuf<double> phi0, dd0, tmp64, ddn64, sfix64;
  casl::DevBuf<float> phif, tmpf, ddnf, sfixf;
  casl::DevBuf<uint16_t> flags;
  casl::DevBuf<unsigned long long> red;
  // tiled encoding (tiles.hpp)
  bool tiled = false;
  int path = CASL_PATH_TILED;
  int64_t n_tiles = 0, n_halo = 0, n_records = 0, n_irregular = 0, n_uniform = 0;
  double geom_base[3] = {0, 0, 0}, geom_unit[3] = {0, 0, 0};
  std::string tile_error;
  casl::DevBuf<int32_t> t_tile_start, t_halo_start, t_h1count, t_halo, t_rec_start;
  casl::DevBuf<uint16_t> t_rec_slot;
  casl::DevBuf<uint4> t_rec;
  casl::DevBuf<uint16_t> t_ext;
  casl::DevBuf<int32_t> t_ext_start;
  casl::DevBuf<uint32_t> t_nreg;
  casl::DevBuf<uint32_t> t_urec;
  casl::DevBuf<int32_t> t_urec_start;
  casl::DevBuf<uint32_t> t_nuni;
  // casl_reinit_host: phi on the device and its pinned staging copy, and one
  // event per staging chunk
  casl::DevBuf<double> host_phi;
  casl::PinnedBuf<double> host_stage;
  casl::DevBuf<uint8_t> host_mask;
  casl::PinnedBuf<uint8_t> host_mask_stage;
  std::vector<hipEvent_t> stage_events;
  // casl_reinit_graph: the recorded call and what it was recorded for
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uintptr_t> graph_key;

  casl_stencil() = default;
  casl_stencil(const casl_stencil &) = delete;
  casl_stencil &operator=(const casl_stencil &) = delete;
  ~casl_stencil() {
    drop_graph();
    for (hipEvent_t e : stage_events) (void)hipEventDestroy(e);
  }
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_key.clear();
  }

  template <int D>
  casl::dev::Tables<D> tables() const {
    casl::dev::Tables<D> t;
    t.nbr = nbr.p;
    t.dist = dist.p;
    t.wall = wall.p;
    t.N = N;
    return t;
  }
  template <int D>
  casl::dev::Tables<D, double> tables64() const {
    casl::dev::Tables<D, double> t;
    t.nbr = nbr.p;
    t.dist = dist64.p;
    t.wall = wall.p;
    t.N = N;
    return t;
  }
  // fn(tables) with the tables an FP64 kernel reads: the FP64 copy when
  // there is one, else the FP32 tables (exact then)
  template <int D, class Fn>
  void with_tables(bool fp64, Fn &&fn) const {
    if (fp64 && wide)
      fn(tables64<D>());
    else
      fn(tables<D>());
  }

  // Every tile scale is a normal power of two whose arm cells (down to 2^-15
  // of it) and their inverses stay normal: true for a tree of any
  // power-of-two edge.  The tiled kernels' shortcuts are exact then.
  bool dyadic_geometry() const { return casl::dyadic_scales(geom_base, 3); }
  // The tiled kernels run this call.  They take a regular face's value as
  // its one corner's value, which equals the reference's weighted form
  // (g w) / w only when w, a product of transverse extents, is a power of
  // two: FP64 on any other geometry runs the grid-wide kernels, which keep
  // the weighted form.
  bool use_tiles(bool fp64) const {
    return tiled && path == CASL_PATH_TILED && (!fp64 || dyadic_geometry());
  }
};

namespace casl {
// Validates and uploads host tables into st (shared by stencils and batches).
// with_tiles also builds the tiled encoding for the fused stage kernel;
// allow_fp64_copy keeps an FP64 copy of distances FP32 does not hold exactly
// (else such tables are refused).
void upload_tables(const casl_flat_stencil *tab, casl_stencil *st, bool with_tiles, bool allow_fp64_copy);
// Bounds check of a tiled encoding of n_nodes nodes (throws), and its upload
// into st (tile arrays, one geometry).
void validate_tile_host(const TileHost &H, int D, int64_t n_nodes);
void upload_tile_host(const TileHost &H, int64_t n_nodes, casl_stencil *st);

// Grid-wide tables of a stencil from tables already on the device (checked
// there, distances converted to FP32 exactly), and the tiled encoding built
// on the device (tiles_gpu.hip); on failure the latter leaves the stencil on
// the two-pass path with the reason in tile_error.
void upload_device_tables(const casl_flat_stencil &dtab, casl_stencil *st, hipStream_t s);
void build_tiles_device(const casl_flat_stencil &dtab, casl_stencil *st, hipStream_t s);

// A reinitialization call (casl_reinit's arguments, no mask) whose buffers are
// carved from [base, base + cap) instead of the stencil's own scratch:
// nothing is allocated.  Throws std::invalid_argument when they do not fit.
void reinit_in_buffer(casl_stencil *st, double *phi, int order_space, int order_time, int n_iter,
                      double limit_high, double limit_low, int precision, hipStream_t s, unsigned char *base,
                      size_t cap);

// Kernel timing (casl_profile_*): prof_begin/prof_end bracket one launch.
bool prof_on();
void prof_begin(const char *kernel, hipStream_t s);
void prof_end(const char *kernel, hipStream_t s);
}
