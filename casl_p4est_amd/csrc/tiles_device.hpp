// Device side of the tiled stencil encoding (tiles.hpp): record decoding and
// the fused pseudo-time stage kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "reinit_device.hpp"
#include "tiles.hpp"

namespace casl {
namespace dev {

template <class R>
struct TileGeom {
  R base[3];      // coarsest arm-cell size per axis
  R unit[3];      // finest cell size per axis
  R inv_base[3];  // 1 / base, exact when pow2
  int pow2;       // every base[a] is a power of two: all arm lengths are, too
  int cubic;      // one scale on every axis (cubic cells)
  int lmax;       // the finest level: unit[a] = base[a] * 2^-lmax
  // Per face axis a (16 bits at 16a): e_t1 - e_a and e_t2 - e_a, each + 128
  // in 8 bits, where base[t] = 2^e_t, for its transverse axes t1 (y for x,
  // else x) and t2 (y for z, else z); zero differences when the cells are
  // cubes (non-cubic cells come on a pow2 geometry only).  A run-time face
  // axis picks them with one shift.
  uint64_t d_pack;
  // pow2 only: the bit pattern (FP32) or high word (FP64) of base and 1/base,
  // whose exponent field a level nibble moves down or up
  uint32_t base_bits[3], inv_bits[3];
};

__device__ __forceinline__ double pow2neg(double, int l) { return __longlong_as_double((1023ll - l) << 52); }
__device__ __forceinline__ float pow2neg(float, int l) { return __int_as_float((127 - l) << 23); }
__device__ __forceinline__ double pow2pos(double, int l) { return __longlong_as_double((1023ll + l) << 52); }
__device__ __forceinline__ float pow2pos(float, int l) { return __int_as_float((127 + l) << 23); }

// Face distances of a record: arm-cell sizes from the level nibbles, wall arms
// -(d_opposite) - (second-degree cell size) (construct_neighbors :609-719).
// iq[f] = 1 / q[f], exact when the geometry is pow2.
// Level nibble l of face f, moved to the exponent field of R's (high) word.
__device__ __forceinline__ uint32_t level_exp(float, uint32_t geo, int f) { return (geo << (23 - 4 * f)) & (15u << 23); }
__device__ __forceinline__ uint32_t level_exp(double, uint32_t geo, int f) { return (geo << (20 - 4 * f)) & (15u << 20); }
__device__ __forceinline__ float from_exp_bits(float, uint32_t b) { return __uint_as_float(b); }
__device__ __forceinline__ double from_exp_bits(double, uint32_t b) { return __hiloint2double(int(b), 0); }

template <int D, class R>
__device__ __forceinline__ void decode_dist(uint32_t geo, const TileGeom<R> &g, R *dist, R *q, R *iq) {
  constexpr int F = 2 * D;
  if (sizeof(R) == 8 && g.pow2) {
    // FP64: base * 2^-l and 2^l / base by exponent arithmetic on the high
    // word (same values, fewer two-word operations); FP32 keeps the
    // multiplications, which measured faster there
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const uint32_t m = level_exp(R(0), geo, f);
      q[f] = from_exp_bits(R(0), g.base_bits[f >> 1] - m);
      iq[f] = from_exp_bits(R(0), g.inv_bits[f >> 1] + m);
    }
  } else {
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const int l = int((geo >> (4 * f)) & 15u);
      q[f] = g.base[f >> 1] * pow2neg(R(0), l);
      iq[f] = g.inv_base[f >> 1] * pow2pos(R(0), l);
    }
  }
#pragma unroll
  for (int f = 0; f < F; ++f) dist[f] = q[f];
  // wall arms: selects per face in FP32; in FP64 (two-word selects) only
  // records with a wall bit take the loop
  if (sizeof(R) == 4 || (geo & (63u << 24))) {
#pragma unroll
    for (int f = 0; f < F; ++f)
      if ((geo >> (24 + f)) & 1u) dist[f] = -dist[f ^ 1] - q[f];
  }
}

// Face distances shared by both stencil kinds.  fast: pow2 geometry and no
// wall arm, so every face distance d is a power of two and x / d == x * iq
// exactly (both are the correctly rounded value of the same real number);
// FASTDIV lets the stencil arithmetic use that.
template <int D, class R>
struct DecodedDist {
  static constexpr bool FASTDIV = true;
  static constexpr bool SLOTS = false;
  static constexpr bool UNIFORM = false;
  static constexpr int F = 2 * D;
  R dist[F];
  R q[F];
  R iq[F];
  bool fast;
  __device__ __forceinline__ R d(int f, int64_t) const { return dist[f]; }
  __device__ __forceinline__ void decode(uint32_t geo, const TileGeom<R> &g) {
    decode_dist<D, R>(geo, g, dist, q, iq);
    fast = g.pow2 && ((geo >> 24) & 63u) == 0u;
  }
};

// Stencil of a regular node (no hanging arm): one LDS-local corner per face.
template <int D, class R>
struct RegStencil : DecodedDist<D, R> {
  static constexpr bool REGULAR = true;
  static constexpr int F = 2 * D;
  uint16_t id[F];
  __device__ __forceinline__ int32_t idx(int f, int, int64_t) const { return id[f]; }
  // transverse offsets (0, arm cell) -- only their product (0) enters the math
  __device__ __forceinline__ R dt(int f, int k, int64_t) const { return (k & 1) ? this->q[f] : R(0); }
  __device__ __forceinline__ void load(const TileRec &r, const TileGeom<R> &g) {
#pragma unroll
    for (int f = 0; f < F; ++f) id[f] = r.idx[f];
    this->decode(r.geo, g);
  }
};

// Stencil of a uniform node on a dyadic geometry of cubic cells (every arm
// cell a power of two, one scale on every axis; the kernels take this
// accessor only when TileGeom::pow2 and cubic, other geometries decode these
// records as RegStencil): one corner per face, all arms q = base * 2^-level
// long, no wall.  The per-face arrays hold the same value so the shared
// stencil code reads them as it reads a RegStencil's.
template <int D, class R>
struct UniStencil {
  static constexpr bool REGULAR = true;
  static constexpr bool FASTDIV = true;
  static constexpr bool SLOTS = false;
  static constexpr bool UNIFORM = true;
  static constexpr bool fast = true;
  static constexpr int F = 2 * D;
  uint16_t id[F];
  R dist[F], q[F], iq[F];
  R iq2[D];  // 1 / q^2 (per axis for the shared second difference)
  __device__ __forceinline__ int32_t idx(int f, int, int64_t) const { return id[f]; }
  __device__ __forceinline__ R d(int f, int64_t) const { return dist[f]; }
  __device__ __forceinline__ R dt(int f, int k, int64_t) const { return (k & 1) ? q[f] : R(0); }
  // w: the UniRec<D> words; returns the record's LDS slot
  __device__ __forceinline__ int load(const uint32_t *w, const TileGeom<R> &g) {
    int slot, level;
    uni_unpack<D>(w, id, slot, level);
    const R qq = g.base[0] * pow2neg(R(0), level), iqq = g.inv_base[0] * pow2pos(R(0), level);
    const R iqq2 = iqq * iqq;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      dist[f] = qq;
      q[f] = qq;
      iq[f] = iqq;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) iq2[a] = iqq2;
    return slot;
  }
};

// The 16-byte record of a uniform node (the general decoding's view of it).
template <int D>
__device__ __forceinline__ TileRec uni_as_rec(const uint32_t *w, int &slot) {
  TileRec r;
  int level;
  uni_unpack<D>(w, r.idx, slot, level);
  uint32_t geo = 0;
#pragma unroll
  for (int f = 0; f < 2 * D; ++f) geo |= uint32_t(level) << (4 * f);
  r.geo = geo;
  return r;
}

// Stencil of a node with hanging arms: the record's one corner per face, and
// for the faces that hang (at most D - 1) the extension record's corners and
// transverse offsets.  idx / dt serve any face (the per-face functions); the
// hanging-face functions of stencil_math.hpp use hface / hidx / hdt.
template <int D, class R>
struct HangStencil : DecodedDist<D, R> {
  static constexpr bool REGULAR = false;
  static constexpr bool SLOTS = true;
  static constexpr int F = 2 * D, C = 1 << (D - 1);
  static constexpr int NH = HangExt<D>::NH, WORDS = HangExt<D>::WORDS, SW = HangExt<D>::SW;
  static constexpr int NW = WORDS / 2;  // 32-bit words (the record is whole 8-byte words)
  uint16_t id[F];
  uint32_t e[NW];
  bool pow2;
  uint32_t geo;
  R base[D], inv_base[D];  // the geometry's (uniform values)
  uint64_t d_pack;         // TileGeom::d_pack and lmax
  int lmax;
  // The arm cell size, its inverse and the face distances of the axis of a
  // face f known only at run time, from the level nibbles and wall bits as
  // decode_dist computes them (no indexing of the per-face arrays, which
  // would move them to scratch memory), and the arm cell's extents along the
  // two transverse axes (x -> y, z; y -> x, z; z -> x, y): the arm cell's
  // own size when the cells are cubes, else (the tiled encoding then needs a
  // pow2 geometry) from the packed exponents.
  struct SlotGeom {
    R qf, iqf;        // arm cell of face f and its inverse
    R df;             // face distance of f
    R dp, dm;         // face distances of the + and - faces of f's axis
    R iqp;            // inverse arm cell of the + face
    R qt1, qt2;       // the arm cell's transverse extents
    R iqt1, iqt2;     // and their inverses
    R ut1, ut2;       // the finest cell along the transverse axes (offset unit)
  };
  __device__ __forceinline__ SlotGeom slot_geom(int f) const {
    const int a = f >> 1;
    R b = base[0], ib = inv_base[0];
#pragma unroll
    for (int k = 1; k < D; ++k) {
      b = (a == k) ? base[k] : b;
      ib = (a == k) ? inv_base[k] : ib;
    }
    const int le = int((geo >> (8 * a)) & 15u), lo = int((geo >> (8 * a + 4)) & 15u);
    const R qe = b * pow2neg(R(0), le), qo = b * pow2neg(R(0), lo);
    const R de = ((geo >> (24 + 2 * a)) & 1u) ? -qo - qe : qe;
    const R dO = ((geo >> (25 + 2 * a)) & 1u) ? -de - qo : qo;
    SlotGeom r;
    const int l = (f & 1) ? lo : le;
    r.qf = (f & 1) ? qo : qe;
    r.iqf = ib * pow2pos(R(0), l);
    r.df = (f & 1) ? dO : de;
    r.dp = dO;
    r.dm = de;
    r.iqp = ib * pow2pos(R(0), lo);
    // the transverse axes' scales are base * 2^d, d = e_t - e_a (0 for
    // cubic cells, whatever the base): exact power-of-two scalings
    const uint32_t dp = uint32_t(d_pack >> (16 * a));
    const int d1 = int(dp & 255u) - 128, d2 = int((dp >> 8) & 255u) - 128;
    const R ua = ldexp(b, -lmax);
    r.qt1 = ldexp(r.qf, d1);
    r.iqt1 = ldexp(r.iqf, -d1);
    r.ut1 = ldexp(ua, d1);
    r.qt2 = ldexp(r.qf, d2);
    r.iqt2 = ldexp(r.iqf, -d2);
    r.ut2 = ldexp(ua, d2);
    return r;
  }
  __device__ __forceinline__ uint32_t word(int i) const { return (e[i >> 1] >> (16 * (i & 1))) & 0xffffu; }
  // face of hanging slot q (HANG_NONE when the node has fewer)
  __device__ __forceinline__ int hface(int q) const {
    return D == 2 ? int(word(0)) : int((word(0) >> (4 * q)) & 15u);
  }
  __device__ __forceinline__ int32_t hidx(int q, int c) const { return int32_t(word(1 + q * SW + c)); }
  // transverse offset k (m0, p0, 0m, 0p) of slot q; sg: the slot's face geometry
  __device__ __forceinline__ R hdt(int q, int k, const SlotGeom &sg) const {
    const R o = R(word(1 + q * SW + C + (k >> 1))) * ((k >> 1) ? sg.ut2 : sg.ut1);
    return (k & 1) ? ((k >> 1) ? sg.qt2 : sg.qt1) - o : o;
  }
  __device__ __forceinline__ int32_t idx(int f, int c, int64_t) const {
    int32_t r = id[f];
#pragma unroll
    for (int q = 0; q < NH; ++q) r = (hface(q) == f) ? hidx(q, c) : r;
    return r;
  }
  // a regular face's offsets (0, q_f, 0, q_f) weight its one corner only
  __device__ __forceinline__ R dt(int f, int k, int64_t) const {
    R r = (k & 1) ? this->q[f] : R(0);
#pragma unroll
    for (int q = 0; q < NH; ++q) r = (hface(q) == f) ? hdt(q, k, slot_geom(f)) : r;
    return r;
  }
  // 1 / (the sum of face f's transverse offsets, times the second sum in 3D):
  // the interpolation's normalisation (pow2 geometry)
  __device__ __forceinline__ R inv_norm(int f) const {
    R r = D == 2 ? this->iq[f] : this->iq[f] * this->iq[f];
#pragma unroll
    for (int q = 0; q < NH; ++q)
      if (hface(q) == f) {
        const SlotGeom sg = slot_geom(f);
        r = D == 2 ? sg.iqt1 : sg.iqt1 * sg.iqt2;
      }
    return r;
  }
  // rec: the node's record; xrec: its extension record (WORDS uint16)
  __device__ __forceinline__ void load(const TileRec &rec, const uint16_t *__restrict__ xrec, const TileGeom<R> &g) {
#pragma unroll
    for (int f = 0; f < F; ++f) id[f] = rec.idx[f];
    const uint2 *src = reinterpret_cast<const uint2 *>(xrec);
#pragma unroll
    for (int i = 0; i < NW / 2; ++i) {
      const uint2 v = src[i];
      e[2 * i] = v.x;
      e[2 * i + 1] = v.y;
    }
    pow2 = g.pow2 != 0;
    geo = rec.geo;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      base[k] = g.base[k];
      inv_base[k] = g.inv_base[k];
    }
    d_pack = g.d_pack;
    lmax = g.lmax;
    this->decode(rec.geo, g);
  }
};

// Extension record of the irregular record at tile-local position i of a
// tile whose extension records start at ext0 (nreg: regular records of T in
// the low 16 bits, of H1 in the high 16 bits; nT tile nodes).
__device__ __forceinline__ const uint16_t *hang_record(const uint16_t *__restrict__ ext, int ext0, uint32_t nreg,
                                                      int nT, int i, int words) {
  const int nTreg = int(nreg & 0xffffu), nH1reg = int(nreg >> 16);
  const int rank = (i < nT) ? i - nTreg : (nT - nTreg) + (i - nT - nH1reg);
  return ext + (int64_t(ext0) + rank) * words;
}

// Blocks are dispatched round-robin over the 8 XCDs; give every XCD one
// contiguous run of tiles so neighbouring tiles (which share halo nodes)
// meet in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int X = 8;
  const int q = nb / X, r = nb % X;
  const int x = b % X, k = b / X;
  return x * q + (x < r ? x : r) + k;
}

template <class R>
struct TileArgs {
  const int32_t *__restrict__ tile_start;
  const int32_t *__restrict__ halo_start;
  const int32_t *__restrict__ h1count;
  const int32_t *__restrict__ halo;
  const int32_t *__restrict__ rec_start;
  const TileRec *__restrict__ rec;
  const uint16_t *__restrict__ rec_slot;
  const uint16_t *__restrict__ ext;
  const int32_t *__restrict__ ext_start;  // [tiles + 1] first extension record of each tile
  const uint32_t *__restrict__ nreg;      // [tiles] regular (non-uniform) records of T | of H1 << 16
  const uint32_t *__restrict__ urec;      // uniform records (UniRec<D> words)
  const int32_t *__restrict__ urec_start; // [tiles + 1] first uniform record of each tile
  const uint32_t *__restrict__ nuni;      // [tiles] uniform records of T | of H1 << 16
  TileGeom<R> geom;
  const uint16_t *__restrict__ flags;
  const R *__restrict__ sfix;
  const double *__restrict__ phi0;
  const R *__restrict__ phin;
  R *out;
  int64_t N;
};

// The ranges of one tile and the map from a record position i (0 <= i <
// nT + nH1: the tile's nodes, then its H1 nodes) to its record: uniform
// records first within T and within H1 (urec), the others after them (rec,
// rec_slot, and ext for those with hanging faces).
template <int D>
struct TileCtx {
  static constexpr int UW = UniRec<D>::WORDS;
  int s, nT, hb, nH, nH1, rb, ub, ext0, nTu, nH1u;
  uint32_t nreg;
  template <class R>
  __device__ __forceinline__ void load(const TileArgs<R> &A, int tile) {
    s = A.tile_start[tile];
    nT = A.tile_start[tile + 1] - s;
    hb = A.halo_start[tile];
    nH = A.halo_start[tile + 1] - hb;
    nH1 = A.h1count[tile];
    rb = A.rec_start[tile];
    ub = A.urec_start[tile];
    ext0 = A.ext_start[tile];
    nreg = A.nreg[tile];
    const uint32_t nu = A.nuni[tile];
    nTu = int(nu & 0xffffu);
    nH1u = int(nu >> 16);
  }
  __device__ __forceinline__ bool uniform(int i) const { return i < nT ? i < nTu : (i - nT) < nH1u; }
  // uniform position i: its record's first word
  __device__ __forceinline__ int64_t uword(int i) const {
    return (int64_t(ub) + (i < nT ? i : nTu + (i - nT))) * UW;
  }
  // other position i: its rank among the tile's non-uniform records
  __device__ __forceinline__ int npos(int i) const { return i < nT ? i - nTu : (nT - nTu) + (i - nT - nH1u); }
  __device__ __forceinline__ const uint16_t *ext_of(const uint16_t *__restrict__ ext, int i) const {
    return hang_record(ext, ext0, nreg, nT - nTu, npos(i), HangExt<D>::WORDS);
  }
};

// A record held in registers: the UniRec words of a uniform position, or the
// 16-byte TileRec of any other (whose LDS slot the caller keeps beside it).
struct RecRegs {
  uint32_t w[4];
};

// Loads the record at position i; returns its LDS slot for a non-uniform
// record (uniform records carry theirs: uni_slot).
template <int D, class R>
__device__ __forceinline__ int fetch_record(const TileArgs<R> &A, const TileCtx<D> &c, int i, RecRegs &r) {
  if (c.uniform(i)) {
    const uint32_t *p = A.urec + c.uword(i);
#pragma unroll
    for (int k = 0; k < UniRec<D>::WORDS; ++k) r.w[k] = p[k];
    return 0;
  }
  const int j = c.rb + c.npos(i);
  const uint4 v = reinterpret_cast<const uint4 *>(A.rec)[j];
  r.w[0] = v.x;
  r.w[1] = v.y;
  r.w[2] = v.z;
  r.w[3] = v.w;
  return A.rec_slot[j];
}

// The general view (TileRec) of the record at position i; slot: the slot
// fetch_record returned, replaced by the record's own for a uniform one.
template <int D>
__device__ __forceinline__ TileRec record_at(const TileCtx<D> &c, int i, const RecRegs &r, int &slot) {
  if (c.uniform(i)) return uni_as_rec<D>(r.w, slot);
  TileRec rec;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rec.idx[2 * k] = uint16_t(r.w[k] & 0xffffu);
    rec.idx[2 * k + 1] = uint16_t(r.w[k] >> 16);
  }
  rec.geo = r.w[3];
  return rec;
}

// The LDS slot of the record at position i.
template <int D>
__device__ __forceinline__ int slot_at(const TileCtx<D> &c, int i, const RecRegs &r, int fetched_slot) {
  return c.uniform(i) ? uni_slot<D>(r.w) : fetched_slot;
}

// Pass A (laplace) of the record at position i into dds[a * DMAX + slot].
// Uniform records on a dyadic geometry (ufast) take UniStencil, the others
// the general decoding.
template <int D, class R>
__device__ __forceinline__ void tile_pass_a(const TileArgs<R> &A, const TileCtx<D> &c, bool ufast, const RecRegs &r,
                                            int fetched_slot, int i, const R *__restrict__ phis,
                                            R *__restrict__ dds) {
  constexpr int DMAX = TileLimits<D>::DMAX;
  R f000, fnb[2 * D], dd[D];
  int j = fetched_slot;
  if (ufast && c.uniform(i)) {
    UniStencil<D, R> st;
    j = st.load(r.w, A.geom);
    laplace<D, R>(st, j, phis, f000, fnb, dd);
  } else {
    const TileRec rec = record_at<D>(c, i, r, j);
    if (!(rec.geo & GEO_IRREGULAR)) {
      RegStencil<D, R> st;
      st.load(rec, A.geom);
      laplace<D, R>(st, j, phis, f000, fnb, dd);
    } else {
      HangStencil<D, R> st;
      st.load(rec, c.ext_of(A.ext, i), A.geom);
      laplace<D, R>(st, j, phis, f000, fnb, dd);
    }
  }
#pragma unroll
  for (int a = 0; a < D; ++a) dds[a * DMAX + j] = dd[a];
}

// Threads per block of the tile kernels: 1024, two tile nodes per thread.
// The FP64 kernels hold 160 KB of LDS arrays, so one block runs per compute
// unit, at most 128 VGPRs for 4 waves per SIMD (a 512-thread block kept 2).
// The FP32 stage kernel holds 80 KB, so two blocks share a compute unit, at
// most 64 VGPRs for 8 waves per SIMD (two 512-thread blocks of four nodes
// per thread kept 4).
#ifndef CASL_TILE_BLOCK
#define CASL_TILE_BLOCK 1024
#endif
#ifndef CASL_TILE_BLOCK_F64
#define CASL_TILE_BLOCK_F64 1024
#endif
constexpr int TILE_BLOCK = CASL_TILE_BLOCK;
constexpr int TILE_BLOCK_F64 = CASL_TILE_BLOCK_F64;
template <class R>
constexpr int stage_block() {
  return sizeof(R) == 8 ? TILE_BLOCK_F64 : TILE_BLOCK;
}

// One pseudo-time stage of a tile: phi of T u H1 u H2 into LDS, pass A for
// T u H1 in LDS, pass B for T.  out = result, or (RK_AVG) out = (out + result)/2.
//
// Latency: the global loads of a node's stencil are issued before the first
// barrier -- each thread's tile records (its own nodes, used in both passes),
// their LDS slots and their flags (stored in record order by the prep pass,
// so the address does not wait for the slot) -- so after the phi gather the
// passes run on registers and LDS (crossing arms aside, whose s_fix is read
// in pass B).  The second RK2 stage's phi_n: see HOLD_PREV.
// Occupancy: the register budget is capped to keep TILE_WAVES_* waves per SIMD.
#ifndef CASL_TILE_WAVES_F32
#define CASL_TILE_WAVES_F32 8
#endif
#ifndef CASL_TILE_WAVES_F64
#define CASL_TILE_WAVES_F64 4
#endif
// The stage body for one tile, run by a block of BLOCK threads with LDS
// arrays phis[LMAX] and dds[D * DMAX].  It reads and writes global memory and
// the LDS arrays only; a caller running several tiles in a row separates them
// with a barrier (the next tile overwrites the LDS arrays).
template <int D, class R, bool SECOND_ORDER, bool RK_AVG, int BLOCK>
__device__ __forceinline__ void tile_stage(const TileArgs<R> &A, int tile, R *__restrict__ phis,
                                           R *__restrict__ dds) {
  constexpr int DMAX = TileLimits<D>::DMAX;
  constexpr int KT = TileLimits<D>::TMAX / BLOCK;  // tile nodes per thread, at most
  static_assert(TileLimits<D>::TMAX % BLOCK == 0, "tile size must be a multiple of the block");
  TileCtx<D> c;
  c.load(A, tile);
  const int s = c.s, nT = c.nT, hb = c.hb, nH = c.nH, nH1 = c.nH1;
  const bool ufast = A.geom.pow2 != 0 && A.geom.cubic != 0;
  RecRegs trec[KT];
  uint32_t tsf[KT];  // LDS slot of a non-uniform record (low 16 bits) and flags (high 16 bits)
  // the phi_n averaged with in the second RK2 stage: loaded before the first
  // barrier where registers allow (blocks below 1024 threads run under a
  // 128-VGPR cap), else read at the store (under the 64-VGPR cap of the
  // 1024-thread FP32 kernel holding it costs spills)
  constexpr bool HOLD_PREV = RK_AVG && BLOCK < 1024;
  R tprev[HOLD_PREV ? KT : 1];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int i = int(threadIdx.x) + k * BLOCK;
    if (i < nT) tsf[k] = uint32_t(fetch_record<D>(A, c, i, trec[k])) | (uint32_t(A.flags[s + i]) << 16);
  }
  // the first K1 rows of H1 records (pass A), also loaded before the barrier
  constexpr int K1 = 2;
  RecRegs hrec[K1];
  int hslot[K1];
  if constexpr (SECOND_ORDER) {
#pragma unroll
    for (int k = 0; k < K1; ++k) {
      const int i = nT + int(threadIdx.x) + k * BLOCK;
      if (i < nT + nH1) hslot[k] = fetch_record<D>(A, c, i, hrec[k]);
    }
  }
  // the phi gather, unrolled so that a thread's loads are in flight together:
  // the halo ids of the first KH rows, then every phi value, then the LDS
  // stores (the rare tail beyond KH rows of halo takes the plain loop)
  constexpr int KH0 = (TileLimits<D>::LMAX - TileLimits<D>::TMAX) / BLOCK / 2;
  constexpr int KH = KH0 > 0 ? KH0 : 1;
  int hid[KH];
#pragma unroll
  for (int k = 0; k < KH; ++k) {
    const int j = int(threadIdx.x) + k * BLOCK;
    if (j < nH) hid[k] = A.halo[hb + j];
  }
  R tv[KT], hv[KH];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int j = int(threadIdx.x) + k * BLOCK;
    if (j < nT) tv[k] = A.phin[s + j];
  }
  if constexpr (HOLD_PREV) {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int i = int(threadIdx.x) + k * BLOCK;
      if (i < nT) tprev[k] = A.out[s + slot_at<D>(c, i, trec[k], int(tsf[k] & 0xffffu))];
    }
  }
#pragma unroll
  for (int k = 0; k < KH; ++k) {
    const int j = int(threadIdx.x) + k * BLOCK;
    if (j < nH) hv[k] = A.phin[hid[k]];
  }
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int j = int(threadIdx.x) + k * BLOCK;
    if (j < nT) phis[j] = tv[k];
  }
#pragma unroll
  for (int k = 0; k < KH; ++k) {
    const int j = int(threadIdx.x) + k * BLOCK;
    if (j < nH) phis[nT + j] = hv[k];
  }
  for (int j = int(threadIdx.x) + KH * BLOCK; j < nH; j += BLOCK) phis[nT + j] = A.phin[A.halo[hb + j]];
  __syncthreads();
  // the phi_n averaged with: loaded here, its latency hidden by pass A
  constexpr bool LATE_PREV = RK_AVG && !HOLD_PREV;
  R lprev[LATE_PREV ? KT : 1];
  if constexpr (LATE_PREV) {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int i = int(threadIdx.x) + k * BLOCK;
      if (i < nT) lprev[k] = A.out[s + slot_at<D>(c, i, trec[k], int(tsf[k] & 0xffffu))];
    }
  }
  if constexpr (SECOND_ORDER) {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int i = int(threadIdx.x) + k * BLOCK;
      if (i < nT) tile_pass_a<D, R>(A, c, ufast, trec[k], int(tsf[k] & 0xffffu), i, phis, dds);
    }
#pragma unroll
    for (int k = 0; k < K1; ++k) {
      const int i = nT + int(threadIdx.x) + k * BLOCK;
      if (i < nT + nH1) tile_pass_a<D, R>(A, c, ufast, hrec[k], hslot[k], i, phis, dds);
    }
    for (int i = nT + int(threadIdx.x) + K1 * BLOCK; i < nT + nH1; i += BLOCK) {
      RecRegs r;
      const int js = fetch_record<D>(A, c, i, r);
      tile_pass_a<D, R>(A, c, ufast, r, js, i, phis, dds);
    }
    __syncthreads();
  }
  // pass B, the result of a tile node, stored (or averaged into out)
  auto store = [&](int k, int64_t n, R r) {
    if constexpr (HOLD_PREV)
      A.out[n] = R(0.5) * (tprev[k] + r);
    else if constexpr (LATE_PREV)
      A.out[n] = R(0.5) * (lprev[k] + r);
    else if constexpr (RK_AVG)
      A.out[n] = R(0.5) * (A.out[n] + r);
    else
      A.out[n] = r;
  };
  // Pass B in two loops: uniform records, then the others.  As branches of
  // one loop body the compiler sinks their common Godunov tail into shared
  // code with per-face registers, and the uniform records would run the
  // general tail.
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int i = int(threadIdx.x) + k * BLOCK;
    if (!(ufast && i < c.nTu)) continue;
    const uint16_t fl = uint16_t(tsf[k] >> 16);
    UniStencil<D, R> st;
    const int j = st.load(trec[k].w, A.geom);
    const R r = stage_value<D, R, SECOND_ORDER, UniStencil<D, R>, true>(st, j, fl, 0, A.sfix + (s + i), A.N,
                                                                          A.phi0 + (int64_t(s) + j), phis, dds, DMAX);
    store(k, int64_t(s) + j, r);
  }
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int i = int(threadIdx.x) + k * BLOCK;
    if (i >= nT || (ufast && i < c.nTu)) continue;
    const uint16_t fl = uint16_t(tsf[k] >> 16);
    int j = int(tsf[k] & 0xffffu);
    R r;
    const TileRec rec = record_at<D>(c, i, trec[k], j);
    const uint8_t w = uint8_t((rec.geo >> 24) & 63u);
    if (!(rec.geo & GEO_IRREGULAR)) {
      RegStencil<D, R> st;
      st.load(rec, A.geom);
      r = stage_value<D, R, SECOND_ORDER, RegStencil<D, R>, true>(st, j, fl, w, A.sfix + (s + i), A.N,
                                                                   A.phi0 + (int64_t(s) + j), phis, dds, DMAX);
    } else {
      HangStencil<D, R> st;
      st.load(rec, c.ext_of(A.ext, i), A.geom);
      r = stage_value<D, R, SECOND_ORDER, HangStencil<D, R>, true>(st, j, fl, w, A.sfix + (s + i), A.N,
                                                                    A.phi0 + (int64_t(s) + j), phis, dds, DMAX);
    }
    store(k, int64_t(s) + j, r);
  }
}

// The phi_0-only prep of one tile in FP64: phi_0 of T u H1 u H2 into LDS,
// pass A of phi_0 for T u H1 (second order), then flags and s_fix of the
// tile's nodes, written in record order (tile start + record position), the
// order the stage body reads them in.  A.phin is phi_0.
template <int D, class S, bool SECOND_ORDER, int BLOCK>
__device__ __forceinline__ void tile_prep(const TileArgs<double> &A, int tile, double *__restrict__ phis,
                                          double *__restrict__ dds, const uint8_t *__restrict__ mask,
                                          double lim_low, double lim_high, double zdt,
                                          uint16_t *__restrict__ flags, S *__restrict__ sfix) {
  constexpr int DMAX = TileLimits<D>::DMAX;
  TileCtx<D> c;
  c.load(A, tile);
  const int s = c.s, nT = c.nT, hb = c.hb, nH = c.nH, nH1 = c.nH1;
  const bool ufast = A.geom.pow2 != 0 && A.geom.cubic != 0;
  for (int j = threadIdx.x; j < nT; j += BLOCK) phis[j] = A.phin[s + j];
  for (int j = threadIdx.x; j < nH; j += BLOCK) phis[nT + j] = A.phin[A.halo[hb + j]];
  __syncthreads();
  if constexpr (SECOND_ORDER) {
    for (int i = int(threadIdx.x); i < nT + nH1; i += BLOCK) {
      RecRegs r;
      const int js = fetch_record<D>(A, c, i, r);
      tile_pass_a<D, double>(A, c, ufast, r, js, i, phis, dds);
    }
    __syncthreads();
  }
  for (int i = int(threadIdx.x); i < nT; i += BLOCK) {
    RecRegs r;
    int j = fetch_record<D>(A, c, i, r);
    uint16_t fl;
    if (ufast && c.uniform(i)) {
      UniStencil<D, double> st;
      j = st.load(r.w, A.geom);
      const bool masked = mask && mask[s + j] == 0;
      fl = prep_value<D, S, UniStencil<D, double>, SECOND_ORDER>(st, j, masked, phis, dds, DMAX, lim_low, lim_high,
                                                                  zdt, SECOND_ORDER, sfix + (s + i), A.N);
    } else {
      const TileRec rec = record_at<D>(c, i, r, j);
      const bool masked = mask && mask[s + j] == 0;
      if (!(rec.geo & GEO_IRREGULAR)) {
        RegStencil<D, double> st;
        st.load(rec, A.geom);
        fl = prep_value<D, S, RegStencil<D, double>, SECOND_ORDER>(st, j, masked, phis, dds, DMAX, lim_low,
                                                                   lim_high, zdt, SECOND_ORDER, sfix + (s + i), A.N);
      } else {
        HangStencil<D, double> st;
        st.load(rec, c.ext_of(A.ext, i), A.geom);
        fl = prep_value<D, S, HangStencil<D, double>, SECOND_ORDER>(st, j, masked, phis, dds, DMAX, lim_low,
                                                                    lim_high, zdt, SECOND_ORDER, sfix + (s + i),
                                                                    A.N);
      }
    }
    flags[s + i] = fl;
  }
}

template <int D, class S, bool SECOND_ORDER>
__global__ void __launch_bounds__(TILE_BLOCK_F64) __attribute__((amdgpu_waves_per_eu(CASL_TILE_WAVES_F64)))
    k_tile_prep(TileArgs<double> A, const uint8_t *__restrict__ mask,
                                                          double lim_low, double lim_high, double zdt,
                                                          uint16_t *__restrict__ flags, S *__restrict__ sfix) {
  __shared__ double phis[TileLimits<D>::LMAX];
  __shared__ double dds[SECOND_ORDER ? D * TileLimits<D>::DMAX : 1];
  tile_prep<D, S, SECOND_ORDER, TILE_BLOCK_F64>(A, xcd_tile(int(blockIdx.x), int(gridDim.x)), phis, dds, mask, lim_low,
                                            lim_high, zdt, flags, sfix);
}

template <int D, class R, bool SECOND_ORDER, bool RK_AVG>
__global__ void __launch_bounds__(stage_block<R>())
    __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? CASL_TILE_WAVES_F32 : CASL_TILE_WAVES_F64)))
    k_tile_stage(TileArgs<R> A) {
  __shared__ R phis[TileLimits<D>::LMAX];
  __shared__ R dds[SECOND_ORDER ? D * TileLimits<D>::DMAX : 1];
  tile_stage<D, R, SECOND_ORDER, RK_AVG, stage_block<R>()>(A, xcd_tile(int(blockIdx.x), int(gridDim.x)), phis, dds);
}

}  // namespace dev

// The tile arrays and geometry of a stencil's tiled encoding in arithmetic R;
// the per-call fields (flags, s_fix, phi_0, phi in and out) are left empty.
template <class R>
dev::TileArgs<R> tile_args(const casl_stencil *st);

}  // namespace casl
