// Tiled stencil encoding for the fused stage kernel (host and device view).
//
// Nodes are cut into tiles of consecutive Morton indices.  For a tile T the
// host precomputes
//   H1 = nodes outside T that carry a nonzero interpolation weight in the
//        stencil of a node of T (their second derivatives are needed),
//   H2 = nodes outside T u H1 in the stencils of H1 nodes,
// and the local numbering  T (n - tile_start) | H1 | H2  used for the
// tile's LDS arrays.  Every node of T u H1 gets a 16-byte record:
//   idx[f]     LDS-local index of the one corner with nonzero weight on
//              face f (a "regular" node: no arm hangs on a coarser cell)
//   geo        4-bit level of each face's arm cell, wall bits, irregular bit
// A node with a hanging arm ("irregular") keeps that record for its faces
// that do not hang and has an extension record (HangExt) for the faces that
// do: at most D - 1 of them, each with all its corner indices and transverse
// offsets (in units of the finest cell of the tree).  Extension records are
// not addressed from the record: within a tile the records are grouped (T
// regular, T hanging on one face, T hanging on two faces, then the same for
// H1), the extension records of a tile follow its irregular records in that
// order from ext_start[tile], and nreg[tile] counts the regular records of T
// and of H1, which gives each irregular record its rank.  The grouping also
// keeps wavefronts on one code path; rec_slot gives each record's LDS slot.
// Most nodes are "uniform": regular, off the walls, all arms at one level (81 %
// of the 34.2M-node octree of 64 spheres).  They get a compact record
// (UniRec) in a separate array instead, with their LDS slot and level packed
// beside the corner indices, and come first within T and within H1, so a
// tile's records run: T uniform (urec), T regular, T hanging on 1 .. D - 1
// faces (rec), H1 uniform (urec), H1 regular, H1 hanging (rec).  Record
// position i of a tile (0 <= i < nT + nH1: T first, then H1) therefore names
// a urec entry or a rec entry by the per-tile counts nuni.
// Decoding reproduces the flattened stencil of the node exactly (zero-weight
// corners aside, which contribute exactly 0 for finite values); the builder
// verifies this for every record.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/casl_reinit.h"

// functions shared by the host builder (g++) and the device kernels (hipcc)
#ifdef __HIPCC__
#define CASL_HD __host__ __device__
#else
#define CASL_HD
#endif

namespace casl {

struct alignas(16) TileRec {
  uint16_t idx[6];
  uint32_t geo;  // bits 0-23: face levels (4 bits each), 24-29: wall bits, 30: irregular
};
static_assert(sizeof(TileRec) == 16, "TileRec must be 16 bytes");

constexpr uint32_t GEO_IRREGULAR = 1u << 30;

// Compact record of a uniform node (all 2D arms at one level, no wall arm, no
// hanging face), in 32-bit words.
//   3D (3 words): word k = idx[2k] (bits 0-12) | idx[2k+1] (13-25) | x_k (26-31)
//                 with x_0, x_1 the low and high 6 bits of the LDS slot and
//                 x_2 the arm level (4 bits)
//   2D (2 words): word 0 = idx[0] (0-11) | idx[1] (12-23) | slot bits 0-7 (24-31)
//                 word 1 = idx[2] (0-11) | idx[3] (12-23) | slot bits 8-11 (24-27) | level (28-31)
template <int D>
struct UniRec;
template <>
struct UniRec<2> {
  static constexpr int WORDS = 2, IDX_BITS = 12, SLOT_BITS = 12;
  static constexpr int LO = 32 - 2 * IDX_BITS;  // slot bits in word 0
};
template <>
struct UniRec<3> {
  static constexpr int WORDS = 3, IDX_BITS = 13, SLOT_BITS = 12;
};

CASL_HD inline uint32_t uni_pack_word(uint32_t a, uint32_t b, int bits, uint32_t extra) {
  return a | (b << bits) | (extra << (2 * bits));
}
template <int D>
CASL_HD inline void uni_pack(const uint16_t *idx, uint32_t slot, uint32_t level, uint32_t *w) {
  if constexpr (D == 3) {
    w[0] = uni_pack_word(idx[0], idx[1], 13, slot & 63u);
    w[1] = uni_pack_word(idx[2], idx[3], 13, slot >> 6);
    w[2] = uni_pack_word(idx[4], idx[5], 13, level);
  } else {
    constexpr int B = UniRec<2>::IDX_BITS, LO = UniRec<2>::LO, S = UniRec<2>::SLOT_BITS;
    w[0] = uni_pack_word(idx[0], idx[1], B, slot & ((1u << LO) - 1u));
    w[1] = uni_pack_word(idx[2], idx[3], B, (slot >> LO) | (level << (S - LO)));
  }
}
// The LDS slot of a uniform record.
template <int D>
CASL_HD inline int uni_slot(const uint32_t *w) {
  if constexpr (D == 3)
    return int((w[0] >> 26) | ((w[1] >> 26) << 6));
  else
    return int((w[0] >> (2 * UniRec<2>::IDX_BITS)) |
               (((w[1] >> (2 * UniRec<2>::IDX_BITS)) & ((1u << (UniRec<2>::SLOT_BITS - UniRec<2>::LO)) - 1u))
                << UniRec<2>::LO));
}
// Unpacks a uniform record: corner indices, LDS slot, level.
template <int D>
CASL_HD inline void uni_unpack(const uint32_t *w, uint16_t *idx, int &slot, int &level) {
  if constexpr (D == 3) {
    constexpr uint32_t M = (1u << 13) - 1;
    for (int k = 0; k < 3; ++k) {
      idx[2 * k] = uint16_t(w[k] & M);
      idx[2 * k + 1] = uint16_t((w[k] >> 13) & M);
    }
    slot = uni_slot<3>(w);
    level = int((w[2] >> 26) & 15u);
  } else {
    constexpr int B = UniRec<2>::IDX_BITS;
    constexpr uint32_t M = (1u << B) - 1;
    for (int k = 0; k < 2; ++k) {
      idx[2 * k] = uint16_t(w[k] & M);
      idx[2 * k + 1] = uint16_t((w[k] >> B) & M);
    }
    slot = uni_slot<2>(w);
    level = int((w[1] >> (2 * B + UniRec<2>::SLOT_BITS - UniRec<2>::LO)) & 15u);
  }
}

// Extension record layout, in uint16 words:
//   2D (4 words):  face, c0, c1, m0
//   3D (16 words): face0 | face1 << 4 (face1 = HANG_NONE for one hanging
//                  face; face0 < face1), then per face c0 c1 c2 c3 m0 zm,
//                  then 3 words of padding
// c*: tile-local corner indices in the reference's corner order; m0: d_m0
// along the first transverse axis, zm: d_0m along the second, in units of
// the finest cell.
template <int D>
struct HangExt;
template <>
struct HangExt<2> {
  static constexpr int NH = 1, WORDS = 4, SW = 3;  // faces, words, words per face
};
template <>
struct HangExt<3> {
  static constexpr int NH = 2, WORDS = 16, SW = 6;
};
constexpr uint16_t HANG_NONE = 15;

// Tile sizing: at most TMAX tile nodes, LMAX nodes in T u H1 u H2 (LDS phi
// slots) and DMAX nodes in T u H1 (LDS second-derivative slots).  The
// defaults are build-time tunables (CASL_TILE_T2 / CASL_TILE_T3).
#ifndef CASL_TILE_T2
#define CASL_TILE_T2 2048
#endif
#ifndef CASL_TILE_T3
#define CASL_TILE_T3 2048
#endif
template <int D>
struct TileLimits;
template <>
struct TileLimits<2> {
  static constexpr int TMAX = CASL_TILE_T2, LMAX = 2 * CASL_TILE_T2, DMAX = 3 * CASL_TILE_T2 / 2;
};
template <>
struct TileLimits<3> {
  static constexpr int TMAX = CASL_TILE_T3, LMAX = 4 * CASL_TILE_T3, DMAX = 2 * CASL_TILE_T3;
};
// the uniform record's fields hold every LDS index and slot of a tile
static_assert(TileLimits<2>::LMAX <= (1 << UniRec<2>::IDX_BITS) && TileLimits<2>::DMAX <= (1 << UniRec<2>::SLOT_BITS),
              "2D tile limits exceed the uniform record's fields");
static_assert(TileLimits<3>::LMAX <= (1 << UniRec<3>::IDX_BITS) && TileLimits<3>::DMAX <= (1 << UniRec<3>::SLOT_BITS),
              "3D tile limits exceed the uniform record's fields");

// Every scale a normal power of two whose arm cells (down to 2^-15 of it) and
// their inverses stay normal: the tiled kernels' exact shortcuts hold.
inline bool dyadic_scales(const double *base, int dim) {
  for (int a = 0; a < dim; ++a) {
    int e = 0;
    if (std::frexp(base[a], &e) != 0.5 || e < -90 || e > 90) return false;
  }
  return true;
}
// Scales the tiled kernels take: dyadic, or one scale on every axis (cubic
// cells, whose hanging arms' transverse extents are the arm cell's size).
inline bool tile_scales_supported(const double *base, int dim) {
  bool cubic = true;
  for (int a = 1; a < dim; ++a) cubic = cubic && base[a] == base[0];
  return cubic || dyadic_scales(base, dim);
}

struct TileHost {
  int dim = 0;
  int lmax = 0;             // finest arm level present
  double size[3] = {1, 1, 1};  // tree size per axis (box edge)
  std::vector<int32_t> tile_start;  // [ntiles + 1] node ranges
  std::vector<int32_t> halo_start;  // [ntiles + 1] ranges in halo_idx
  std::vector<int32_t> h1count;     // [ntiles] number of H1 entries (first in each halo range)
  std::vector<int32_t> halo_idx;    // global node ids of H1 then H2, per tile
  std::vector<int32_t> rec_start;   // [ntiles + 1] ranges in rec (the non-uniform records of T u H1)
  std::vector<TileRec> rec;         // per tile: T regular, T hanging, H1 regular, H1 hanging
  std::vector<uint16_t> rec_slot;   // LDS slot (local index) of each record
  std::vector<int32_t> urec_start;  // [ntiles + 1] ranges in urec, in records
  std::vector<uint32_t> urec;       // UniRec<D> records (WORDS each): T uniform, then H1 uniform
  std::vector<uint32_t> nuni;       // [ntiles] uniform records of T (low 16 bits) and of H1 (high 16 bits)
  std::vector<uint16_t> ext;        // HangExt<D> records, as raw uint16
  std::vector<int32_t> ext_start;   // [ntiles + 1] first extension record of each tile
  std::vector<uint32_t> nreg;       // [ntiles] regular records of T (low 16 bits) and of H1 (high 16 bits)
  int64_t n_irregular = 0;
};

// Builds the tiled encoding of a flattened stencil; throws std::runtime_error
// when the stencil cannot be encoded exactly (then the grid-wide kernels are
// used instead).  box sizes are recovered from the arm lengths.
void build_tiles(const casl_flat_stencil &tab, TileHost &out);

}  // namespace casl
