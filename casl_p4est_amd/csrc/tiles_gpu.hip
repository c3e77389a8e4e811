// The tiled stencil encoding (tiles.hpp) built on the device from tables that
// are already there (a grid built by casl_device_grid_build), so a grid that
// changes every step does not travel to the host and back.  The result is the
// host builder's (tiles.cpp) byte for byte:
//
//   scales     the coarsest arm cell and the finest level: two reductions
//   tiles      candidates of TMAX consecutive nodes; every tile whose halo
//              exceeds the LDS limits is halved, round after round, which
//              gives the tiles of the host's depth-first halving in order
//   halos      H1 and H2 of every tile at once: (tile, node) keys of the
//              needed nodes outside the tile, radix-sorted and made unique
//   records    every record position classified (uniform, regular, hanging
//              on 1 or 2 faces; own node or H1), a stable sort by (tile,
//              group) gives the host's emission order, and each record is
//              encoded, checked against the tables and written in place
//   check      the independent bounds check of validate_tile_host, per tile
// The stencil it fills keeps only device arrays; the few scalars the host
// needs (the scales, the tile and record counts, the error flags) are the
// only values copied back.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "reinit_device.hpp"
#include "tiles_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

constexpr int BLOCK = 256;
inline unsigned grid_for(int64_t n) {
  if (n >= (int64_t(1) << 31)) throw std::runtime_error("grid too large for one launch");
  return unsigned(std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
}

// error bits raised by the kernels
enum : int {
  BAD_INDEX = 1,       // a stencil node index out of range
  BAD_ZERO = 2,        // a zero face distance
  BAD_INEXACT = 4,     // a distance not exact in FP32
  BAD_SCALE = 8,       // an arm cell that is not base * 2^-l, 0 <= l <= 15
  BAD_ENCODE = 16,     // a record the encoding cannot hold (corner outside T u H1 u H2, too many hanging faces)
  BAD_UNITS = 32,      // a transverse offset that is not a multiple of the finest cell
  BAD_VERIFY = 64,     // a decoded record differs from the tables
  BAD_BOUNDS = 128,    // the independent bounds check failed
};

// The flattened tables on the device (FP64 distances), as casl_flat_stencil.
template <int D>
struct DTab {
  static constexpr int F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
  const int32_t *__restrict__ nbr;
  const double *__restrict__ d;
  const double *__restrict__ dtr;
  const uint8_t *__restrict__ wall;
  int64_t N;
  __device__ int32_t nb(int f, int c, int64_t n) const { return nbr[(int64_t(f) * C + c) * N + n]; }
  __device__ double df(int f, int64_t n) const { return d[int64_t(f) * N + n]; }
  __device__ double dt(int f, int k, int64_t n) const { return dtr[(int64_t(f) * T + k) * N + n]; }
  __device__ bool w(int f, int64_t n) const { return (wall[n] >> f) & 1; }
  // size of the cell that carries face f's neighbours
  __device__ double arm_cell(int f, int64_t n) const { return w(f, n) ? (-df(f, n) - df(f ^ 1, n)) : df(f, n); }
  // get_linear_interpolation_weights of face f's corners
  __device__ void weights(int f, int64_t n, double *wt) const {
    const double m0 = dt(f, 0, n), p0 = dt(f, 1, n);
    if constexpr (D == 2) {
      wt[0] = p0;
      wt[1] = m0;
    } else {
      const double zm = dt(f, 2, n), zp = dt(f, 3, n);
      wt[0] = p0 * zp;
      wt[1] = m0 * zp;
      wt[2] = p0 * zm;
      wt[3] = m0 * zm;
    }
  }
};

// The scales of the encoding: per axis, the coarsest arm cell and the finest
// cell (a cell of level l spans base[a] * 2^-l along axis a).
struct Scales {
  double base[3], unit[3];
  int lmax;
};

__device__ __forceinline__ void flag(int *bad, int bit) { atomicOr(bad, bit); }

// ---------------------------------------------------------------- tables
// The grid-wide tables of a stencil from device tables: indices checked,
// distances converted to FP32 (exact, or the stencil cannot be made).
template <int D>
__global__ void k_check_convert(DTab<D> t, float *__restrict__ dist, int *__restrict__ bad) {
  constexpr int F = DTab<D>::F, C = DTab<D>::C, T = DTab<D>::T;
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  for (int f = 0; f < F; ++f) {
    for (int c = 0; c < C; ++c) {
      const int32_t g = t.nb(f, c, n);
      if (g < 0 || g >= t.N) flag(bad, BAD_INDEX);
    }
    const double v = t.df(f, n);
    if (!(v != 0.0)) flag(bad, BAD_ZERO);
    const float fv = float(v);
    if (double(fv) != v) flag(bad, BAD_INEXACT);
    dist[int64_t(f) * t.N + n] = fv;
    for (int k = 0; k < T; ++k) {
      const double u = t.dt(f, k, n);
      const float fu = float(u);
      if (double(fu) != u) flag(bad, BAD_INEXACT);
      dist[(int64_t(F) + int64_t(f) * T + k) * t.N + n] = fu;
    }
  }
}

// ---------------------------------------------------------------- scales
// Blocks of the two scale reductions: each thread strides over the nodes, the
// block reduces through wave shuffles and LDS, and one atomic per block
// reaches memory (one per node serialized on the few target words).
constexpr unsigned REDUCE_BLOCKS = 2048;
inline unsigned reduce_grid(int64_t n) { return std::max(1u, std::min(REDUCE_BLOCKS, grid_for(n))); }

template <class T>
__device__ __forceinline__ T block_max(T v) {
  __shared__ T part[BLOCK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  __syncthreads();  // part[] may still be read by a previous call
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < BLOCK / 64; ++k) v = part[k] > v ? part[k] : v;
  return v;  // the block's maximum in thread 0
}

// max arm cell per axis, out[D] (non-negative doubles order like their bit
// patterns)
template <int D>
__global__ void __launch_bounds__(BLOCK) k_arm_max(DTab<D> t, unsigned long long *__restrict__ out,
                                                   int *__restrict__ bad) {
  double m[D];
#pragma unroll
  for (int a = 0; a < D; ++a) m[a] = 0.0;
  for (int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; n < t.N; n += int64_t(gridDim.x) * blockDim.x) {
#pragma unroll
    for (int a = 0; a < D; ++a)
      for (int f = 2 * a; f < 2 * a + 2; ++f) {
        const double q = t.arm_cell(f, n);
        if (!(q > 0)) flag(bad, BAD_SCALE);
        m[a] = fmax(m[a], q);
      }
  }
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const unsigned long long b = block_max((unsigned long long)__double_as_longlong(m[a]));
    if (threadIdx.x == 0) atomicMax(out + a, b);
  }
}

// level_of of every arm cell.  The host builder takes l = the nearest
// integer of log2(base / q) and requires base * 2^-l == q with 0 <= l <= 15;
// that accepts exactly the q with base's significand and a binary exponent l
// below base's, and returns that l, so the exponents are compared here
// without the logarithm.
__device__ __forceinline__ int level_of(double base, double q, int *bad) {
  if (!(q > 0)) {
    flag(bad, BAD_SCALE);
    return 0;
  }
  int eb = 0, eq = 0;
  const double mb = frexp(base, &eb), mq = frexp(q, &eq);
  const int l = eb - eq;
  if (mb != mq || l < 0 || l > 15) {
    flag(bad, BAD_SCALE);
    return 0;
  }
  return l;
}

template <int D>
__global__ void __launch_bounds__(BLOCK) k_level_max(DTab<D> t, Scales sc, int *__restrict__ out,
                                                     int *__restrict__ bad) {
  int m = 0;
  for (int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; n < t.N; n += int64_t(gridDim.x) * blockDim.x)
    for (int f = 0; f < DTab<D>::F; ++f) m = max(m, level_of(sc.base[f >> 1], t.arm_cell(f, n), bad));
  m = block_max(m);
  if (threadIdx.x == 0) atomicMax(out, m);
}

// ---------------------------------------------------------------- halos
__device__ __forceinline__ int tile_of(const int32_t *__restrict__ ts, int nt, int64_t n) {
  int lo = 0, hi = nt;  // the last tile whose start is <= n
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ts[mid] <= n) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t hkey(int t, int32_t g) { return (uint64_t(uint32_t(t)) << 32) | uint32_t(g); }

// binary search of g in the sorted keys [b, e) of one tile
__device__ __forceinline__ int64_t find_key(const uint64_t *__restrict__ keys, int64_t b, int64_t e, uint64_t k) {
  while (b < e) {
    const int64_t m = (b + e) >> 1;
    if (keys[m] < k) b = m + 1;
    else e = m;
  }
  return b;
}

// the nodes with a nonzero weight in node n's stencil; returns how many
template <int D>
__device__ __forceinline__ int needed(const DTab<D> &t, int64_t n, int32_t *out) {
  int k = 0;
  for (int f = 0; f < DTab<D>::F; ++f) {
    double w[4];
    t.weights(f, n, w);
    for (int c = 0; c < DTab<D>::C; ++c)
      if (w[c] != 0.0) out[k++] = t.nb(f, c, n);
  }
  return k;
}

// H1 candidates: needed nodes outside the node's tile.  EMIT: write the
// keys at off[n], else count them into cnt[n].
template <int D, bool EMIT>
__global__ void k_h1(DTab<D> t, const int32_t *__restrict__ ts, int nt, int64_t *__restrict__ cnt,
                     const int64_t *__restrict__ off, uint64_t *__restrict__ keys) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= t.N) return;
  const int tl = tile_of(ts, nt, n);
  const int32_t s = ts[tl], e = ts[tl + 1];
  int32_t buf[DTab<D>::F * DTab<D>::C];
  const int k = needed<D>(t, n, buf);
  int64_t o = EMIT ? off[n] : 0;
  int c = 0;
  for (int i = 0; i < k; ++i)
    if (buf[i] < s || buf[i] >= e) {
      if (EMIT) keys[o++] = hkey(tl, buf[i]);
      ++c;
    }
  if (!EMIT) cnt[n] = c;
}

// first key of every tile in a sorted unique key array: seg[t] (t <= nt)
__global__ void k_segments(const uint64_t *__restrict__ keys, int64_t nk, int nt, int64_t *__restrict__ seg) {
  const int t = int(blockIdx.x) * blockDim.x + int(threadIdx.x);
  if (t > nt) return;
  seg[t] = find_key(keys, 0, nk, uint64_t(uint32_t(t)) << 32);
}

// H2 candidates: for every H1 entry, its needed nodes outside the tile and
// outside the tile's H1.
template <int D, bool EMIT>
__global__ void k_h2(DTab<D> t, const int32_t *__restrict__ ts, const uint64_t *__restrict__ h1, int64_t nh1,
                     const int64_t *__restrict__ h1seg, int64_t *__restrict__ cnt, const int64_t *__restrict__ off,
                     uint64_t *__restrict__ keys) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= nh1) return;
  const int tl = int(h1[q] >> 32);
  const int32_t h = int32_t(uint32_t(h1[q]));
  const int32_t s = ts[tl], e = ts[tl + 1];
  const int64_t b = h1seg[tl], en = h1seg[tl + 1];
  int32_t buf[DTab<D>::F * DTab<D>::C];
  const int k = needed<D>(t, h, buf);
  int64_t o = EMIT ? off[q] : 0;
  int c = 0;
  for (int i = 0; i < k; ++i) {
    const int32_t g = buf[i];
    if (g >= s && g < e) continue;
    const uint64_t key = hkey(tl, g);
    const int64_t p = find_key(h1, b, en, key);
    if (p < en && h1[p] == key) continue;
    if (EMIT) keys[o++] = key;
    ++c;
  }
  if (!EMIT) cnt[q] = c;
}

// 1 where a tile exceeds its LDS limits (the host's halo() test)
template <int D>
__global__ void k_tile_fits(const int32_t *__restrict__ ts, int nt, const int64_t *__restrict__ h1seg,
                            const int64_t *__restrict__ h2seg, uint8_t *__restrict__ fail) {
  const int t = int(blockIdx.x) * blockDim.x + int(threadIdx.x);
  if (t >= nt) return;
  const int64_t nT = ts[t + 1] - ts[t], n1 = h1seg[t + 1] - h1seg[t], n2 = h2seg[t + 1] - h2seg[t];
  fail[t] = (nT + n1 > TileLimits<D>::DMAX || nT + n1 + n2 > TileLimits<D>::LMAX) ? 1 : 0;
}

// ---------------------------------------------------------------- records
// The tile arrays a record needs.
struct TileView {
  const int32_t *ts;      // [nt + 1] node ranges
  const int64_t *rpos;    // [nt + 1] first record position of each tile (prefix of nT + nH1)
  const uint64_t *h1, *h2;
  const int64_t *h1seg, *h2seg;
  int nt;
  Scales sc;
};

// A tile handled by one block: its halo node ids, H1 then H2 (each in
// ascending order), staged in LDS, and the LDS slot of a node of the tile.
template <int D>
struct TileLocal {
  int32_t s = 0, e = 0;  // the tile's node range
  int n1 = 0, n2 = 0;    // H1 and H2 entries
  const int32_t *h = nullptr;
  // LDS slot of node g in T u H1 u H2, -1 if not there
  __device__ int32_t operator()(int32_t g) const {
    if (g >= s && g < e) return g - s;
    int p = lower(0, n1, g);
    if (p < n1 && h[p] == g) return (e - s) + p;
    p = lower(n1, n1 + n2, g);
    if (p < n1 + n2 && h[p] == g) return (e - s) + p;
    return -1;
  }
  // node of record position i (the tile's nodes, then its H1 nodes)
  __device__ int64_t node(int i) const { return i < e - s ? int64_t(s) + i : int64_t(h[i - (e - s)]); }
  __device__ int lower(int b, int end, int32_t g) const {
    while (b < end) {
      const int m = (b + end) >> 1;
      if (h[m] < g) b = m + 1;
      else end = m;
    }
    return b;
  }
};

// Stages tile tl's halo ids into hl (every thread of the block calls it);
// false (and an error bit) if they exceed the LDS array, which the tile
// limits rule out.
template <int D>
__device__ __forceinline__ bool stage_tile(const TileView &v, int tl, int32_t *hl, TileLocal<D> &loc, int *bad) {
  loc.s = v.ts[tl];
  loc.e = v.ts[tl + 1];
  const int64_t b1 = v.h1seg[tl], b2 = v.h2seg[tl];
  loc.n1 = int(v.h1seg[tl + 1] - b1);
  loc.n2 = int(v.h2seg[tl + 1] - b2);
  loc.h = hl;
  if (loc.n1 + loc.n2 > TileLimits<D>::LMAX) {
    if (threadIdx.x == 0) flag(bad, BAD_BOUNDS);
    return false;
  }
  for (int j = int(threadIdx.x); j < loc.n1 + loc.n2; j += int(blockDim.x))
    hl[j] = int32_t(uint32_t(j < loc.n1 ? v.h1[b1 + j] : v.h2[b2 + (j - loc.n1)]));
  __syncthreads();
  return true;
}

__device__ __forceinline__ void transverse(int D, int a, int &t1, int &t2) {
  if (D == 2) {
    t1 = 1 - a;
    t2 = -1;
  } else {
    t1 = (a == 0 ? 1 : 0);
    t2 = (a == 2 ? 1 : 2);
  }
}

// The record of node n in tile tl, as the host builder's encode(): corners of
// nonzero weight, levels, walls, hanging faces and their extension words.
// Returns the number of hanging faces (-1 when the record cannot be encoded).
// SLOTS = false: the levels, wall bits and hanging faces only (what a
// record's group depends on), without the slot lookups and the extension
// words; the emission encodes every record in full and flags what fails.
template <int D, bool SLOTS = true>
__device__ int encode(const DTab<D> &t, const TileView &v, const TileLocal<D> &loc, int64_t n, TileRec &r,
                      uint16_t *e, int *bad) {
  constexpr int F = DTab<D>::F, C = DTab<D>::C;
  constexpr int NH = HangExt<D>::NH, WORDS = HangExt<D>::WORDS, SW = HangExt<D>::SW;
  uint32_t geo = 0;
  int hf[2] = {HANG_NONE, HANG_NONE}, k = 0;
  for (int f = 0; f < F; ++f) {
    geo |= uint32_t(level_of(v.sc.base[f >> 1], t.arm_cell(f, n), bad)) << (4 * f);
    if (t.w(f, n)) geo |= 1u << (24 + f);
    double w[4];
    t.weights(f, n, w);
    int nz = 0, first = -1;
    for (int c = 0; c < C; ++c)
      if (w[c] != 0.0) {
        ++nz;
        if (first < 0) first = c;
      }
    if (first < 0) {
      flag(bad, BAD_ENCODE);
      return -1;
    }
    if constexpr (SLOTS) {
      const int32_t l = loc(t.nb(f, first, n));
      if (l < 0 || l > 65535) {
        flag(bad, BAD_ENCODE);
        return -1;
      }
      r.idx[f] = uint16_t(l);
    }
    if (nz != 1) {
      if (k == NH) {
        flag(bad, BAD_ENCODE);
        return -1;
      }
      hf[k++] = f;
    }
  }
  if constexpr (D == 2) r.idx[4] = r.idx[5] = 0;
  if (k > 0) geo |= GEO_IRREGULAR;
  if (SLOTS && k > 0) {
    for (int q = 0; q < WORDS; ++q) e[q] = 0;
    e[0] = uint16_t(D == 2 ? hf[0] : (hf[0] | (hf[1] << 4)));
    for (int q = 0; q < k; ++q) {
      const int f = hf[q];
      int t1, t2;
      transverse(D, f >> 1, t1, t2);
      double w[4];
      t.weights(f, n, w);
      int first = -1;
      for (int c = 0; c < C; ++c)
        if (w[c] != 0.0 && first < 0) first = c;
      uint16_t *fe = e + 1 + q * SW;
      for (int c = 0; c < C; ++c) {
        const int32_t l = loc(t.nb(f, w[c] != 0.0 ? c : first, n));
        if (l < 0 || l > 65535) {
          flag(bad, BAD_ENCODE);
          return -1;
        }
        fe[c] = uint16_t(l);
      }
      // units(): the offset in finest cells of its transverse axis
      for (int m = 0; m < (D == 3 ? 2 : 1); ++m) {
        const double val = t.dt(f, 2 * m, n);
        const double u = v.sc.unit[m == 0 ? t1 : t2];
        const double kk = val / u;
        if (!(kk >= 0) || kk > 65535 || floor(kk) != kk || kk * u != val) {
          flag(bad, BAD_UNITS);
          return -1;
        }
        fe[C + m] = uint16_t(kk);
      }
    }
  }
  r.geo = geo;
  return k;
}

// The host builder's verify(): the decoded record reproduces the tables
// (distances bit-exact, one nonzero-weight corner on a face without a
// hanging arm).  Its slots are what encode() looked up for the same nodes
// just before, so they are not looked up again.
template <int D>
__device__ void verify(const DTab<D> &t, const TileView &v, int64_t n, const TileRec &r, const uint16_t *e,
                       int *bad) {
  constexpr int F = DTab<D>::F, C = DTab<D>::C, SW = HangExt<D>::SW;
  const double *base = v.sc.base;
  double dd[6];
  for (int f = 0; f < F; ++f) dd[f] = base[f >> 1] * ldexp(1.0, -int((r.geo >> (4 * f)) & 15u));
  for (int f = 0; f < F; ++f) {
    if ((r.geo >> (24 + f)) & 1u) dd[f] = -dd[f ^ 1] - base[f >> 1] * ldexp(1.0, -int((r.geo >> (4 * f)) & 15u));
    if (dd[f] != t.df(f, n)) flag(bad, BAD_VERIFY);
  }
  int hf[2] = {HANG_NONE, HANG_NONE};
  if (r.geo & GEO_IRREGULAR) {
    hf[0] = D == 2 ? e[0] : (e[0] & 15);
    if (D == 3) hf[1] = (e[0] >> 4) & 15;
    if (hf[0] >= F || (hf[1] != HANG_NONE && (hf[1] >= F || hf[1] <= hf[0]))) flag(bad, BAD_VERIFY);
  }
  for (int f = 0; f < F; ++f) {
    double w[4];
    t.weights(f, n, w);
    const int q = (f == hf[0]) ? 0 : (f == hf[1]) ? 1 : -1;
    if (q < 0) {
      int nz = 0;
      for (int c = 0; c < C; ++c) nz += (w[c] != 0.0) ? 1 : 0;
      if (nz != 1) flag(bad, BAD_VERIFY);
      continue;
    }
    const uint16_t *fe = e + 1 + q * SW;
    // the arm cell's transverse extents, at the face's level
    int t1, t2;
    transverse(D, f >> 1, t1, t2);
    const double lv = ldexp(1.0, -int((r.geo >> (4 * f)) & 15u));
    const double qh1 = base[t1] * lv;
    const double m0 = fe[C] * v.sc.unit[t1];
    if (m0 != t.dt(f, 0, n) || qh1 - m0 != t.dt(f, 1, n)) flag(bad, BAD_VERIFY);
    if (D == 3) {
      const double qh2 = base[t2] * lv;
      const double zm = fe[C + 1] * v.sc.unit[t2];
      if (zm != t.dt(f, 2, n) || qh2 - zm != t.dt(f, 3, n)) flag(bad, BAD_VERIFY);
    }
  }
}

// a regular record off the walls whose arms all sit at one level
template <int D>
__device__ __forceinline__ bool uniform_rec(const TileRec &r) {
  if (r.geo & (GEO_IRREGULAR | (63u << 24))) return false;
  for (int f = 1; f < 2 * D; ++f)
    if (((r.geo >> (4 * f)) & 15u) != (r.geo & 15u)) return false;
  return true;
}

// group of a record: 0 uniform, 1 regular, 2 / 3 hanging on 1 / 2 faces,
// +4 for an H1 record
template <int D>
__device__ __forceinline__ int group_of(const TileRec &r, int nh, bool h1) {
  const int g = nh > 0 ? 1 + nh : (uniform_rec<D>(r) ? 0 : 1);
  return g + (h1 ? 4 : 0);
}

// sort key (tile, group) of every record position; the value is the
// position.  One block per tile.
template <int D>
__global__ void __launch_bounds__(BLOCK) k_classify(DTab<D> t, TileView v, uint64_t *__restrict__ keys,
                                                    int64_t *__restrict__ vals, int *__restrict__ bad) {
  __shared__ int32_t hl[TileLimits<D>::LMAX];
  const int tl = int(blockIdx.x);
  TileLocal<D> loc;
  if (!stage_tile<D>(v, tl, hl, loc, bad)) return;
  const int nT = loc.e - loc.s;
  const int64_t p0 = v.rpos[tl];
  for (int i = int(threadIdx.x); i < nT + loc.n1; i += BLOCK) {
    TileRec r;
    uint16_t e[HangExt<D>::WORDS];
    const int nh = encode<D, false>(t, v, loc, loc.node(i), r, e, bad);
    keys[p0 + i] = (uint64_t(uint32_t(tl)) << 3) | uint32_t(group_of<D>(r, nh < 0 ? 0 : nh, i >= nT));
    vals[p0 + i] = p0 + i;
  }
}

// per tile: where each group starts in the sorted order (gs[t * 9 + g])
__global__ void k_group_starts(const uint64_t *__restrict__ keys, int64_t R, int nt, int64_t *__restrict__ gs) {
  const int t = int(blockIdx.x) * blockDim.x + int(threadIdx.x);
  if (t >= nt) return;
  for (int g = 0; g <= 8; ++g) {
    const uint64_t k = (uint64_t(uint32_t(t)) << 3) + uint64_t(g);
    int64_t b = 0, e = R;
    while (b < e) {
      const int64_t m = (b + e) >> 1;
      if (keys[m] < k) b = m + 1;
      else e = m;
    }
    gs[int64_t(t) * 9 + g] = b;
  }
}

// per tile counts of uniform, non-uniform and irregular records, H1 + H2
// halo entries, and the packed nreg / nuni / h1count
__global__ void k_tile_counts(const int64_t *__restrict__ gs, const int32_t *__restrict__ ts,
                              const int64_t *__restrict__ h1seg, const int64_t *__restrict__ h2seg, int nt,
                              int64_t *__restrict__ cu, int64_t *__restrict__ cn, int64_t *__restrict__ cx,
                              int64_t *__restrict__ ch, uint32_t *__restrict__ nreg, uint32_t *__restrict__ nuni,
                              int32_t *__restrict__ h1count) {
  const int t = int(blockIdx.x) * blockDim.x + int(threadIdx.x);
  if (t >= nt) return;
  const int64_t *g = gs + int64_t(t) * 9;
  const int64_t c[8] = {g[1] - g[0], g[2] - g[1], g[3] - g[2], g[4] - g[3],
                        g[5] - g[4], g[6] - g[5], g[7] - g[6], g[8] - g[7]};
  cu[t] = c[0] + c[4];
  cn[t] = c[1] + c[2] + c[3] + c[5] + c[6] + c[7];
  cx[t] = c[2] + c[3] + c[6] + c[7];
  const int64_t n1 = h1seg[t + 1] - h1seg[t], n2 = h2seg[t + 1] - h2seg[t];
  ch[t] = n1 + n2;
  nreg[t] = uint32_t(c[1]) | (uint32_t(c[5]) << 16);
  nuni[t] = uint32_t(c[0]) | (uint32_t(c[4]) << 16);
  h1count[t] = int32_t(n1);
  (void)ts;
}

// The output arrays of the encoding.
struct TileOut {
  TileRec *rec;
  uint16_t *rec_slot;
  uint32_t *urec;
  uint16_t *ext;
  int32_t *halo;
  const int64_t *urs, *rs, *xs, *hs;  // exclusive prefixes of the per-tile counts
};

// every record, encoded again, checked and written at its place in the
// host builder's order.  One block per tile: its records are the sorted
// positions gs[tl * 9] .. gs[tl * 9 + 8].
template <int D>
__global__ void __launch_bounds__(BLOCK) k_emit(DTab<D> t, TileView v, const int64_t *__restrict__ sorted_p,
                                                const uint64_t *__restrict__ keys, const int64_t *__restrict__ gs,
                                                TileOut o, int *__restrict__ bad) {
  __shared__ int32_t hl[TileLimits<D>::LMAX];
  const int tl = int(blockIdx.x);
  TileLocal<D> loc;
  if (!stage_tile<D>(v, tl, hl, loc, bad)) return;
  const int64_t *G = gs + int64_t(tl) * 9;
  const int64_t p0 = v.rpos[tl];
  const int nr = (loc.e - loc.s) + loc.n1;
  for (int64_t k = G[0] + threadIdx.x; k < G[8]; k += BLOCK) {
    const int64_t ip = sorted_p[k] - p0;
    if (ip < 0 || ip >= nr) {
      flag(bad, BAD_BOUNDS);
      continue;
    }
    const int i = int(ip);
    const int g = int(keys[k] & 7u);
    const int64_t n = loc.node(i);
    TileRec r;
    uint16_t e[HangExt<D>::WORDS];
    const int nh = encode<D>(t, v, loc, n, r, e, bad);
    if (nh < 0) continue;
    verify<D>(t, v, n, r, e, bad);
    if (g == 0 || g == 4) {
      const int64_t u = o.urs[tl] + (g == 0 ? k - G[0] : (G[1] - G[0]) + (k - G[4]));
      uni_pack<D>(r.idx, uint32_t(i), r.geo & 15u, o.urec + u * UniRec<D>::WORDS);
      continue;
    }
    const int64_t q = o.rs[tl] + (g < 4 ? k - G[1] : (G[4] - G[1]) + (k - G[5]));
    o.rec[q] = r;
    o.rec_slot[q] = uint16_t(i);
    if (nh > 0) {
      const int64_t x = o.xs[tl] + (g < 4 ? k - G[2] : (G[4] - G[2]) + (k - G[6]));
      for (int w = 0; w < HangExt<D>::WORDS; ++w) o.ext[x * HangExt<D>::WORDS + w] = e[w];
    }
  }
}

// halo_idx: H1 then H2 of every tile
__global__ void k_emit_halo(const uint64_t *__restrict__ h, int64_t nh, const int64_t *__restrict__ seg,
                            const int64_t *__restrict__ h1seg, const int64_t *__restrict__ hs, bool second,
                            int32_t *__restrict__ out) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= nh) return;
  const int tl = int(h[q] >> 32);
  const int64_t pos = hs[tl] + (second ? (h1seg[tl + 1] - h1seg[tl]) : 0) + (q - seg[tl]);
  out[pos] = int32_t(uint32_t(h[q]));
}

// The independent bounds check of validate_tile_host, one block per tile.
template <int D>
__global__ void k_validate(TileArgs<float> A, int64_t N, int *__restrict__ bad) {
  constexpr int F = 2 * D, C = 1 << (D - 1), LMAX = TileLimits<D>::LMAX, DMAX = TileLimits<D>::DMAX;
  constexpr int WORDS = HangExt<D>::WORDS, SW = HangExt<D>::SW;
  __shared__ uint32_t seen[(DMAX + 31) / 32];
  const int tile = int(blockIdx.x);
  TileCtx<D> c;
  c.load(A, tile);
  for (int w = int(threadIdx.x); w < (DMAX + 31) / 32; w += BLOCK) seen[w] = 0;
  __syncthreads();
  const int nTn = c.nT - c.nTu, nH1n = c.nH1 - c.nH1u;
  const int nTreg = int(c.nreg & 0xffffu), nH1reg = int(c.nreg >> 16);
  const int nR = A.rec_start[tile + 1] - c.rb, nU = A.urec_start[tile + 1] - c.ub;
  if (c.nT <= 0 || c.nH1 < 0 || c.nH1 > c.nH || c.nT + c.nH > LMAX || c.nT + c.nH1 > DMAX ||
      nR + nU != c.nT + c.nH1 || c.nTu > c.nT || c.nH1u > c.nH1 || c.nTu + c.nH1u != nU || nTreg > nTn ||
      nH1reg > nH1n || A.ext_start[tile + 1] - c.ext0 != (nTn - nTreg) + (nH1n - nH1reg)) {
    if (threadIdx.x == 0) flag(bad, BAD_BOUNDS);
    return;
  }
  for (int k = int(threadIdx.x); k < c.nH; k += BLOCK) {
    const int32_t g = A.halo[c.hb + k];
    if (g < 0 || g >= N) flag(bad, BAD_BOUNDS);
  }
  for (int i = int(threadIdx.x); i < c.nT + c.nH1; i += BLOCK) {
    RecRegs r;
    const int js = fetch_record<D>(A, c, i, r);
    int slot = js;
    const TileRec rec = record_at<D>(c, i, r, slot);
    const bool own = i < c.nT;
    if (own ? (slot < 0 || slot >= c.nT) : (slot < c.nT || slot >= c.nT + c.nH1)) {
      flag(bad, BAD_BOUNDS);
      continue;
    }
    if (atomicOr(&seen[slot >> 5], 1u << (slot & 31)) & (1u << (slot & 31))) flag(bad, BAD_BOUNDS);
    const int bound = own ? c.nT + c.nH1 : c.nT + c.nH;
    for (int f = 0; f < F; ++f)
      if (rec.idx[f] >= bound) flag(bad, BAD_BOUNDS);
    if (!c.uniform(i)) {
      const int j = c.npos(i);
      const bool irregular = (rec.geo & GEO_IRREGULAR) != 0;
      if (irregular != (own ? (j >= nTreg) : (j >= nTn + nH1reg))) flag(bad, BAD_BOUNDS);
      if (irregular) {
        const uint16_t *e = c.ext_of(A.ext, i);
        const int h0 = D == 2 ? e[0] : (e[0] & 15), h1 = D == 2 ? int(HANG_NONE) : ((e[0] >> 4) & 15);
        if (h0 >= F || (h1 != HANG_NONE && (h1 >= F || h1 <= h0))) flag(bad, BAD_BOUNDS);
        for (int q = 0; q < (h1 == HANG_NONE ? 1 : 2); ++q)
          for (int cc = 0; cc < C; ++cc)
            if (e[1 + q * SW + cc] >= bound) flag(bad, BAD_BOUNDS);
      }
    }
  }
  (void)WORDS;
}

// ------------------------------------------------------------ host side
template <class Fn>
void cub_call(hipStream_t s, Fn fn) {
  size_t bytes = 0;
  HIP_CHECK(fn(nullptr, bytes));
  DevBuf<unsigned char> tmp;
  tmp.ensure(std::max<size_t>(bytes, 1));
  HIP_CHECK(fn(tmp.p, bytes));
  HIP_CHECK(hipStreamSynchronize(s));
}

template <class T>
T read_back(const T *p, hipStream_t s) {
  T v;
  HIP_CHECK(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return v;
}

std::string bad_reason(int b) {
  if (b & BAD_SCALE) return "arm cell size is not a power-of-two fraction of the coarsest cell";
  if (b & BAD_UNITS) return "transverse offset is not a multiple of the finest cell";
  if (b & BAD_ENCODE) return "a record the encoding cannot hold (corner outside the tile's halo or too many hanging faces)";
  if (b & BAD_VERIFY) return "tile record does not reproduce the tables";
  if (b & BAD_BOUNDS) return "tiled encoding failed its bounds check";
  return "tiled encoding failed";
}

// sorted unique keys (in place into `out`, count returned)
int64_t sort_unique(DevBuf<uint64_t> &keys, int64_t n, int end_bit, DevBuf<uint64_t> &out, hipStream_t s) {
  DevBuf<uint64_t> sorted;
  DevBuf<int64_t> cnt;
  sorted.ensure(size_t(std::max<int64_t>(n, 1)));
  out.ensure(size_t(std::max<int64_t>(n, 1)));
  cnt.ensure(1);
  if (n == 0) return 0;
  cub_call(s, [&](void *t, size_t &b) {
    return hipcub::DeviceRadixSort::SortKeys(t, b, keys.p, sorted.p, n, 0, end_bit, s);
  });
  cub_call(s, [&](void *t, size_t &b) { return hipcub::DeviceSelect::Unique(t, b, sorted.p, out.p, cnt.p, n, s); });
  return read_back(cnt.p, s);
}

int bit_width(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// exclusive prefix of n counts into out[0..n] (out[n] = total)
void exclusive(const int64_t *in, int64_t n, int64_t *out, hipStream_t s) {
  HIP_CHECK(hipMemsetAsync(out, 0, sizeof(int64_t), s));
  if (n == 0) return;
  cub_call(s, [&](void *t, size_t &b) { return hipcub::DeviceScan::InclusiveSum(t, b, in, out + 1, n, s); });
}

__global__ void k_i64_to_i32(const int64_t *__restrict__ a, int32_t *__restrict__ b, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = int32_t(a[i]);
}

// record positions of a tile: nT + nH1
__global__ void k_tile_records(const int32_t *__restrict__ ts, const int64_t *__restrict__ h1seg, int nt,
                               int64_t *__restrict__ out) {
  const int t = int(blockIdx.x) * blockDim.x + int(threadIdx.x);
  if (t < nt) out[t] = int64_t(ts[t + 1] - ts[t]) + (h1seg[t + 1] - h1seg[t]);
}

template <int D>
void build(const DTab<D> &tab, casl_stencil *st, hipStream_t s) {
  const int64_t N = tab.N;
  DevBuf<int> bad;
  bad.ensure(1);
  HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  // scales
  DevBuf<unsigned long long> amax;
  DevBuf<int> lmx;
  amax.ensure(D);
  lmx.ensure(1);
  HIP_CHECK(hipMemsetAsync(amax.p, 0, D * sizeof(unsigned long long), s));
  HIP_CHECK(hipMemsetAsync(lmx.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_arm_max<D>, dim3(reduce_grid(N)), dim3(BLOCK), 0, s, tab, amax.p, bad.p);
  HIP_CHECK(hipGetLastError());
  // as the host builder: one scale per axis (axes past D keep the box edge 1)
  unsigned long long bits[3] = {0, 0, 0};
  HIP_CHECK(hipMemcpyAsync(bits, amax.p, D * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  Scales sc;
  for (int a = 0; a < 3; ++a) sc.base[a] = 1.0;
  for (int a = 0; a < D; ++a) std::memcpy(&sc.base[a], &bits[a], sizeof(double));
  hipLaunchKernelGGL(k_level_max<D>, dim3(reduce_grid(N)), dim3(BLOCK), 0, s, tab, sc, lmx.p, bad.p);
  HIP_CHECK(hipGetLastError());
  sc.lmax = read_back(lmx.p, s);
  for (int a = 0; a < 3; ++a) sc.unit[a] = sc.base[a] * std::ldexp(1.0, -sc.lmax);
  if (!tile_scales_supported(sc.base, D))
    throw std::runtime_error("cells that are not cubes on a box whose edges are not dyadic");
  int b = read_back(bad.p, s);
  if (b) throw std::runtime_error(bad_reason(b));

  // tiles: candidates of TMAX nodes, halved where the halo does not fit
  constexpr int TMAX = TileLimits<D>::TMAX;
  std::vector<int32_t> starts;
  for (int64_t x = 0; x < N; x += TMAX) starts.push_back(int32_t(x));
  starts.push_back(int32_t(N));
  DevBuf<int32_t> ts;
  DevBuf<uint64_t> keys, h1, h2;
  DevBuf<int64_t> cnt, off, h1seg, h2seg;
  int64_t nh1 = 0, nh2 = 0;
  int nt = 0;
  for (int round = 0;; ++round) {
    nt = int(starts.size()) - 1;
    ts.ensure(starts.size());
    HIP_CHECK(hipMemcpyAsync(ts.p, starts.data(), starts.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    const int end_bit = 32 + bit_width(uint64_t(nt));
    // H1
    cnt.ensure(size_t(N));
    off.ensure(size_t(N) + 1);
    hipLaunchKernelGGL((k_h1<D, false>), dim3(grid_for(N)), dim3(BLOCK), 0, s, tab, ts.p, nt, cnt.p, nullptr,
                       nullptr);
    HIP_CHECK(hipGetLastError());
    exclusive(cnt.p, N, off.p, s);
    const int64_t k1 = read_back(off.p + N, s);
    keys.ensure(size_t(std::max<int64_t>(k1, 1)));
    hipLaunchKernelGGL((k_h1<D, true>), dim3(grid_for(N)), dim3(BLOCK), 0, s, tab, ts.p, nt, nullptr, off.p, keys.p);
    HIP_CHECK(hipGetLastError());
    nh1 = sort_unique(keys, k1, end_bit, h1, s);
    h1seg.ensure(size_t(nt) + 1);
    hipLaunchKernelGGL(k_segments, dim3(grid_for(nt + 1)), dim3(BLOCK), 0, s, h1.p, nh1, nt, h1seg.p);
    HIP_CHECK(hipGetLastError());
    // H2
    cnt.ensure(size_t(std::max<int64_t>(nh1, 1)));
    off.ensure(size_t(nh1) + 1);
    if (nh1 > 0)
      hipLaunchKernelGGL((k_h2<D, false>), dim3(grid_for(nh1)), dim3(BLOCK), 0, s, tab, ts.p, h1.p, nh1, h1seg.p,
                         cnt.p, nullptr, nullptr);
    HIP_CHECK(hipGetLastError());
    exclusive(cnt.p, nh1, off.p, s);
    const int64_t k2 = read_back(off.p + nh1, s);
    keys.ensure(size_t(std::max<int64_t>(k2, 1)));
    if (nh1 > 0)
      hipLaunchKernelGGL((k_h2<D, true>), dim3(grid_for(nh1)), dim3(BLOCK), 0, s, tab, ts.p, h1.p, nh1, h1seg.p,
                         nullptr, off.p, keys.p);
    HIP_CHECK(hipGetLastError());
    nh2 = sort_unique(keys, k2, end_bit, h2, s);
    h2seg.ensure(size_t(nt) + 1);
    hipLaunchKernelGGL(k_segments, dim3(grid_for(nt + 1)), dim3(BLOCK), 0, s, h2.p, nh2, nt, h2seg.p);
    HIP_CHECK(hipGetLastError());
    // which tiles do not fit
    DevBuf<uint8_t> fail;
    fail.ensure(size_t(nt));
    hipLaunchKernelGGL(k_tile_fits<D>, dim3(grid_for(nt)), dim3(BLOCK), 0, s, ts.p, nt, h1seg.p, h2seg.p, fail.p);
    HIP_CHECK(hipGetLastError());
    std::vector<uint8_t> hf(static_cast<size_t>(nt));
    HIP_CHECK(hipMemcpyAsync(hf.data(), fail.p, size_t(nt), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    bool any = false;
    std::vector<int32_t> next;
    next.reserve(starts.size() + 16);
    for (int t = 0; t < nt; ++t) {
      next.push_back(starts[size_t(t)]);
      if (hf[size_t(t)]) {
        const int32_t a = starts[size_t(t)], e = starts[size_t(t) + 1];
        if (e - a <= 1) throw std::runtime_error("a single node exceeds the tile limits");
        next.push_back(a + (e - a) / 2);
        any = true;
      }
    }
    next.push_back(int32_t(N));
    if (!any) break;
    starts.swap(next);
    if (round > 40) throw std::runtime_error("tile splitting did not converge");
  }
  keys.release();
  cnt.release();
  off.release();

  // records: positions, classification, the host builder's order
  DevBuf<int64_t> rcount, rpos;
  rcount.ensure(size_t(nt));
  rpos.ensure(size_t(nt) + 1);
  hipLaunchKernelGGL(k_tile_records, dim3(grid_for(nt)), dim3(BLOCK), 0, s, ts.p, h1seg.p, nt, rcount.p);
  HIP_CHECK(hipGetLastError());
  exclusive(rcount.p, nt, rpos.p, s);
  const int64_t R = read_back(rpos.p + nt, s);
  if (R >= int64_t(INT32_MAX) || nh1 + nh2 >= int64_t(INT32_MAX))
    throw std::runtime_error("tile tables too large for int32 offsets");
  TileView v{ts.p, rpos.p, h1.p, h2.p, h1seg.p, h2seg.p, nt, sc};
  DevBuf<uint64_t> rkeys, rkeys_sorted;
  DevBuf<int64_t> rvals, rvals_sorted;
  rkeys.ensure(size_t(R));
  rvals.ensure(size_t(R));
  rkeys_sorted.ensure(size_t(R));
  rvals_sorted.ensure(size_t(R));
  hipLaunchKernelGGL(k_classify<D>, dim3(unsigned(nt)), dim3(BLOCK), 0, s, tab, v, rkeys.p, rvals.p, bad.p);
  HIP_CHECK(hipGetLastError());
  b = read_back(bad.p, s);
  if (b) throw std::runtime_error(bad_reason(b));
  const int kbits = 3 + bit_width(uint64_t(nt));
  cub_call(s, [&](void *t, size_t &bb) {
    return hipcub::DeviceRadixSort::SortPairs(t, bb, rkeys.p, rkeys_sorted.p, rvals.p, rvals_sorted.p, R, 0, kbits,
                                              s);
  });
  rkeys.release();
  rvals.release();
  DevBuf<int64_t> gs, cu, cn, cx, ch, urs, rs, xs, hs;
  gs.ensure(size_t(nt) * 9);
  for (auto *x : {&cu, &cn, &cx, &ch}) x->ensure(size_t(nt));
  for (auto *x : {&urs, &rs, &xs, &hs}) x->ensure(size_t(nt) + 1);
  hipLaunchKernelGGL(k_group_starts, dim3(grid_for(nt)), dim3(BLOCK), 0, s, rkeys_sorted.p, R, nt, gs.p);
  st->t_nreg.ensure(size_t(nt));
  st->t_nuni.ensure(size_t(nt));
  st->t_h1count.ensure(size_t(nt));
  hipLaunchKernelGGL(k_tile_counts, dim3(grid_for(nt)), dim3(BLOCK), 0, s, gs.p, ts.p, h1seg.p, h2seg.p, nt, cu.p,
                     cn.p, cx.p, ch.p, st->t_nreg.p, st->t_nuni.p, st->t_h1count.p);
  HIP_CHECK(hipGetLastError());
  exclusive(cu.p, nt, urs.p, s);
  exclusive(cn.p, nt, rs.p, s);
  exclusive(cx.p, nt, xs.p, s);
  exclusive(ch.p, nt, hs.p, s);
  const int64_t n_uni = read_back(urs.p + nt, s), n_rec = read_back(rs.p + nt, s), n_ext = read_back(xs.p + nt, s);
  const int64_t n_halo = read_back(hs.p + nt, s);
  constexpr int UW = UniRec<D>::WORDS, XW = HangExt<D>::WORDS;
  st->t_rec.ensure(size_t(std::max<int64_t>(n_rec, 1)));
  st->t_rec_slot.ensure(size_t(std::max<int64_t>(n_rec, 1)));
  st->t_urec.ensure(size_t(std::max<int64_t>(n_uni * UW, 4)));
  st->t_ext.ensure(size_t(std::max<int64_t>(n_ext * XW, 16)));
  st->t_halo.ensure(size_t(std::max<int64_t>(n_halo, 1)));
  TileOut o{reinterpret_cast<TileRec *>(st->t_rec.p), st->t_rec_slot.p, st->t_urec.p, st->t_ext.p, st->t_halo.p,
            urs.p, rs.p, xs.p, hs.p};
  hipLaunchKernelGGL(k_emit<D>, dim3(unsigned(nt)), dim3(BLOCK), 0, s, tab, v, rvals_sorted.p, rkeys_sorted.p, gs.p,
                     o, bad.p);
  HIP_CHECK(hipGetLastError());
  if (nh1 > 0)
    hipLaunchKernelGGL(k_emit_halo, dim3(grid_for(nh1)), dim3(BLOCK), 0, s, h1.p, nh1, h1seg.p, h1seg.p, hs.p, false,
                       st->t_halo.p);
  if (nh2 > 0)
    hipLaunchKernelGGL(k_emit_halo, dim3(grid_for(nh2)), dim3(BLOCK), 0, s, h2.p, nh2, h2seg.p, h1seg.p, hs.p, true,
                       st->t_halo.p);
  HIP_CHECK(hipGetLastError());
  // per-tile arrays (int32)
  st->t_tile_start.ensure(size_t(nt) + 1);
  HIP_CHECK(hipMemcpyAsync(st->t_tile_start.p, ts.p, (size_t(nt) + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  auto to32 = [&](DevBuf<int64_t> &src, DevBuf<int32_t> &dst) {
    dst.ensure(size_t(nt) + 1);
    hipLaunchKernelGGL(k_i64_to_i32, dim3(grid_for(nt + 1)), dim3(BLOCK), 0, s, src.p, dst.p, int64_t(nt) + 1);
    HIP_CHECK(hipGetLastError());
  };
  to32(hs, st->t_halo_start);
  to32(rs, st->t_rec_start);
  to32(urs, st->t_urec_start);
  to32(xs, st->t_ext_start);
  b = read_back(bad.p, s);
  if (b) throw std::runtime_error(bad_reason(b));
  // geometry and counts
  st->n_tiles = nt;
  st->n_halo = n_halo;
  st->n_records = R;
  st->n_uniform = n_uni;
  st->n_irregular = n_ext;
  for (int a = 0; a < 3; ++a) {
    st->geom_base[a] = sc.base[a];
    st->geom_unit[a] = sc.unit[a];
  }
  st->tiled = true;
  // the independent bounds check before any kernel reads the encoding
  hipLaunchKernelGGL(k_validate<D>, dim3(unsigned(nt)), dim3(BLOCK), 0, s, tile_args<float>(st), N, bad.p);
  HIP_CHECK(hipGetLastError());
  b = read_back(bad.p, s);
  if (b) {
    st->tiled = false;
    throw std::runtime_error(bad_reason(b));
  }
}

}  // namespace

namespace casl {

// The grid-wide tables of a stencil from tables already on the device.
void upload_device_tables(const casl_flat_stencil &dtab, casl_stencil *st, hipStream_t s) {
  const int D = dtab.dim;
  if (D != 2 && D != 3) throw std::invalid_argument("dim must be 2 or 3");
  if (dtab.n_nodes <= 0 || dtab.n_nodes >= int64_t(INT32_MAX)) throw std::invalid_argument("bad n_nodes");
  const int64_t N = dtab.n_nodes;
  const int F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
  st->dim = D;
  st->N = N;
  st->zdt = dtab.zero_distance_threshold;
  HIP_CHECK(hipGetDevice(&st->device));
  st->nbr.ensure(size_t(F) * C * N);
  st->dist.ensure(size_t(F) * (1 + T) * N);
  st->wall.ensure(size_t(N));
  HIP_CHECK(hipMemcpyAsync(st->nbr.p, dtab.nbr, size_t(F) * C * N * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(st->wall.p, dtab.wall, size_t(N), hipMemcpyDeviceToDevice, s));
  DevBuf<int> bad;
  bad.ensure(1);
  HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  if (D == 2) {
    DTab<2> t{dtab.nbr, dtab.d, dtab.d_tr, dtab.wall, N};
    hipLaunchKernelGGL(k_check_convert<2>, dim3(grid_for(N)), dim3(BLOCK), 0, s, t, st->dist.p, bad.p);
  } else {
    DTab<3> t{dtab.nbr, dtab.d, dtab.d_tr, dtab.wall, N};
    hipLaunchKernelGGL(k_check_convert<3>, dim3(grid_for(N)), dim3(BLOCK), 0, s, t, st->dist.p, bad.p);
  }
  HIP_CHECK(hipGetLastError());
  const int b = read_back(bad.p, s);
  if (b & BAD_INDEX) throw std::runtime_error("stencil node index out of range");
  if (b & BAD_ZERO) throw std::runtime_error("zero face distance");
  if (b & BAD_INEXACT) throw std::domain_error("stencil distance not exact in FP32");
}

// The tiled encoding of device tables, built on the device into st; on
// failure the stencil keeps only the grid-wide tables and records why.
void build_tiles_device(const casl_flat_stencil &dtab, casl_stencil *st, hipStream_t s) {
  st->tiled = false;
  try {
    if (dtab.dim == 2)
      build<2>(DTab<2>{dtab.nbr, dtab.d, dtab.d_tr, dtab.wall, dtab.n_nodes}, st, s);
    else
      build<3>(DTab<3>{dtab.nbr, dtab.d, dtab.d_tr, dtab.wall, dtab.n_nodes}, st, s);
  } catch (const hip_error &) {
    throw;
  } catch (const std::exception &e) {
    st->tiled = false;
    st->tile_error = e.what();
    return;
  }
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, st->device));
  st->path = st->n_tiles >= int64_t(prop.multiProcessorCount) ? CASL_PATH_TILED : CASL_PATH_TWO_PASS;
}

}  // namespace casl
