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
// Decoding reproduces the flattened stencil of the node exactly (zero-weight
// corners aside, which contribute exactly 0 for finite values); the builder
// verifies this for every record.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/casl_reinit.h"

namespace casl {

struct alignas(16) TileRec {
  uint16_t idx[6];
  uint32_t geo;  // bits 0-23: face levels (4 bits each), 24-29: wall bits, 30: irregular
};
static_assert(sizeof(TileRec) == 16, "TileRec must be 16 bytes");

constexpr uint32_t GEO_IRREGULAR = 1u << 30;

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
#define CASL_TILE_T2 1024
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

struct TileHost {
  int dim = 0;
  int lmax = 0;             // finest arm level present
  double size[3] = {1, 1, 1};  // tree size per axis (box edge)
  std::vector<int32_t> tile_start;  // [ntiles + 1] node ranges
  std::vector<int32_t> halo_start;  // [ntiles + 1] ranges in halo_idx
  std::vector<int32_t> h1count;     // [ntiles] number of H1 entries (first in each halo range)
  std::vector<int32_t> halo_idx;    // global node ids of H1 then H2, per tile
  std::vector<int32_t> rec_start;   // [ntiles + 1] ranges in rec (T u H1 of each tile)
  std::vector<TileRec> rec;         // per tile: T regular, T hanging, H1 regular, H1 hanging
  std::vector<uint16_t> rec_slot;   // LDS slot (local index) of each record
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
