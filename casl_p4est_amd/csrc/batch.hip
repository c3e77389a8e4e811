// Batched reinitialization of many independent small trees in one
// persistent kernel (reinitialize_2nd_order, src/my_p4est_level_set.h:468,
// applied to every tree).
//
// The trees' stencils are concatenated into one table set (node indices
// offset per tree).  A persistent grid of workgroups pulls trees from a
// device work queue, largest first, and runs the whole call for its tree --
// phi_0 copy, pass A on phi_0, the phi_0-only prep, then every pseudo-time
// stage -- separated by workgroup barriers only: stencils never cross
// trees, so no grid-wide synchronisation is needed.  The per-node arithmetic
// is the same device code as the grid-wide kernels, so each tree's result is
// bit-identical to casl_reinit on that tree alone.
//
// Two kernels share that structure.  k_batch gathers through the grid-wide
// tables, NPT nodes per thread.  k_batch_tiles runs every stage of its tree
// as a sequence of the tree's LDS-staged tiles (the fused stage body of the
// single-tree path, compact 16-byte records): each tree's tiled encoding is
// built on the host, checked, and concatenated.  The launch shape -- kernel,
// threads per block, nodes per thread -- is picked by casl_batch_autotune;
// casl_batch_reinit_graph records the queue reset and the persistent kernel
// into a hipGraph once and replays it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "reinit_device.hpp"
#include "tiles_device.hpp"

using namespace casl;
using namespace casl::dev;

namespace {

template <int D, class R>
struct BatchArgs {
  Tables<D> t;
  const int64_t *__restrict__ off;    // [n_trees + 1]
  const int32_t *__restrict__ order;  // processing order, largest tree first
  int32_t n_trees;
  unsigned int *__restrict__ queue;   // work counter
  double *__restrict__ phi;           // in/out, FP64
  double *__restrict__ phi0;
  double *__restrict__ dd0;
  uint16_t *__restrict__ flags;
  R *__restrict__ sfix;
  R *__restrict__ cur;
  R *__restrict__ nxt;
  R *__restrict__ ddn;
  int n_iter;
  double zdt;
};

// Runs body(n) for every node n of [b, e): each thread takes NPT nodes a
// blockDim apart per sweep so their gathers are in flight together.
template <int BT, int NPT, class Body>
__device__ __forceinline__ void for_nodes(int64_t b, int64_t e, Body body) {
  for (int64_t base = b + threadIdx.x; base < e; base += int64_t(BT) * NPT) {
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      const int64_t n = base + int64_t(j) * BT;
      if (n < e) body(n);
    }
  }
}

template <int D, class R, int BT, int NPT>
__global__ void __launch_bounds__(BT) k_batch(BatchArgs<D, R> A) {
  __shared__ int32_t s_tree;
  const Tables<D> &t = A.t;
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned k = atomicAdd(A.queue, 1u);
      s_tree = (int32_t(k) < A.n_trees) ? A.order[k] : -1;
    }
    __syncthreads();
    const int32_t tree = s_tree;
    __syncthreads();  // every thread has read s_tree before thread 0 rewrites it
    if (tree < 0) break;
    const int64_t b = A.off[tree], e = A.off[tree + 1];
    for_nodes<BT, NPT>(b, e, [&](int64_t n) {
      const double v = A.phi[n];
      A.phi0[n] = v;
      A.cur[n] = R(v);
    });
    __syncthreads();
    for_nodes<BT, NPT>(b, e, [&](int64_t n) { node_dd<D, double>(t, n, A.phi0, A.dd0); });
    __syncthreads();
    for_nodes<BT, NPT>(b, e, [&](int64_t n) {
      node_prep<D, R>(t, n, A.phi0, A.dd0, nullptr, -DBL_MAX, DBL_MAX, A.zdt, true, A.flags, A.sfix);
    });
    __syncthreads();
    for (int it = 0; it < A.n_iter; ++it) {
      for_nodes<BT, NPT>(b, e, [&](int64_t n) { node_dd<D, R>(t, n, A.cur, A.ddn); });
      __syncthreads();
      for_nodes<BT, NPT>(b, e, [&](int64_t n) {
        A.nxt[n] = node_stage<D, R, true>(t, n, A.flags[n], A.sfix, A.phi0, A.cur, A.ddn);
      });
      __syncthreads();
      for_nodes<BT, NPT>(b, e, [&](int64_t n) { node_dd<D, R>(t, n, A.nxt, A.ddn); });
      __syncthreads();
      for_nodes<BT, NPT>(b, e, [&](int64_t n) {
        const R r = node_stage<D, R, true>(t, n, A.flags[n], A.sfix, A.phi0, A.nxt, A.ddn);
        A.cur[n] = R(0.5) * (A.cur[n] + r);
      });
      __syncthreads();
    }
    for_nodes<BT, NPT>(b, e, [&](int64_t n) { A.phi[n] = double(A.cur[n]); });
    __syncthreads();
  }
}

// Arguments of the tiled persistent kernel: the concatenated tiled encoding of
// all trees (one geometry) and the per-tree tile ranges.
template <class R>
struct BatchTileArgs {
  TileArgs<R> T;
  const int64_t *__restrict__ off;
  const int32_t *__restrict__ order;
  const int32_t *__restrict__ tree_tile;  // [n_trees + 1]
  unsigned int *__restrict__ queue;
  double *__restrict__ phi;
  R *__restrict__ cur;
  R *__restrict__ nxt;
  int32_t n_trees;
  int n_iter;
};

// phi_0 and the working copy of every batch node (before the tiled prep and
// the persistent kernel).
template <class R>
__global__ void k_batch_copy(const double *__restrict__ phi, double *__restrict__ phi0, R *__restrict__ cur,
                             int64_t N) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double v = phi[n];
  phi0[n] = v;
  cur[n] = R(v);
}

// Register caps of k_batch_tiles (waves per SIMD): its 256- and 512-thread
// blocks run several tile nodes per thread, which need up to 128 VGPRs in
// FP32 and 168 in FP64; the 1024-thread blocks of 2D trees (two nodes of a
// 2048-node tile per thread) keep FP32 at 64 VGPRs, two blocks per compute
// unit, and FP64 at 128.
constexpr int BATCH_WAVES_F32 = 4, BATCH_WAVES_F64 = 3;
template <class R, int BLOCK>
constexpr int batch_waves() {
  if (BLOCK == 1024) return sizeof(R) == 4 ? 8 : 4;
  return sizeof(R) == 4 ? BATCH_WAVES_F32 : BATCH_WAVES_F64;
}

// Persistent kernel over trees whose pseudo-time stages run tile by tile in
// LDS; every tree's result is its single-tree tiled result.
template <int D, class R, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(batch_waves<R, BLOCK>())))
    k_batch_tiles(BatchTileArgs<R> A) {
  __shared__ R phis[TileLimits<D>::LMAX];
  __shared__ R dds[D * TileLimits<D>::DMAX];
  int32_t *s_tree = reinterpret_cast<int32_t *>(phis);  // between trees phis is free
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned k = atomicAdd(A.queue, 1u);
      *s_tree = (int32_t(k) < A.n_trees) ? A.order[k] : -1;
    }
    __syncthreads();
    const int32_t tree = *s_tree;
    __syncthreads();  // every thread has read the tree before the tiles overwrite phis
    if (tree < 0) break;
    const int t0 = A.tree_tile[tree], t1 = A.tree_tile[tree + 1];
    TileArgs<R> a = A.T;
    for (int it = 0; it < A.n_iter; ++it) {
      a.phin = A.cur;
      a.out = A.nxt;
      for (int tile = t0; tile < t1; ++tile) {
        tile_stage<D, R, true, false, BLOCK>(a, tile, phis, dds);
        __syncthreads();
      }
      a.phin = A.nxt;
      a.out = A.cur;
      for (int tile = t0; tile < t1; ++tile) {
        tile_stage<D, R, true, true, BLOCK>(a, tile, phis, dds);
        __syncthreads();
      }
    }
    const int64_t b = A.off[tree], e = A.off[tree + 1];
    for (int64_t n = b + threadIdx.x; n < e; n += BLOCK) A.phi[n] = double(A.cur[n]);
    __syncthreads();
  }
}

__global__ void k_reset_queue(unsigned int *q) { *q = 0u; }

// npt >= 1: k_batch with that many nodes per thread; npt == 0: k_batch_tiles.
struct Schedule {
  int block = 256;
  int npt = 1;
};

const int kBlocks[] = {64, 128, 256, 512};
const int kNpts[] = {1, 2};

template <int D, class R>
using KernelFn = void (*)(BatchArgs<D, R>);
template <int D, class R>
using TileKernelFn = void (*)(BatchTileArgs<R>);

template <int D, class R>
KernelFn<D, R> pick_kernel(const Schedule &s) {
#define CASL_PICK(BT, NPT) \
  if (s.block == BT && s.npt == NPT) return &k_batch<D, R, BT, NPT>;
  CASL_PICK(64, 1) CASL_PICK(64, 2) CASL_PICK(128, 1) CASL_PICK(128, 2) CASL_PICK(256, 1) CASL_PICK(256, 2)
  CASL_PICK(512, 1) CASL_PICK(512, 2)
#undef CASL_PICK
  throw std::invalid_argument("unsupported batch schedule (block in {64,128,256,512}, nodes/thread in {1,2}; "
                              "nodes/thread 0: tiled kernel with block 256 (2D) or 512)");
}

// block sizes of the tiled kernel: a tile's nodes per thread stay small
template <int D>
constexpr bool tile_block_ok(int block) {
  return block == 512 || (D == 2 && (block == 256 || block == 1024));
}

template <int D, class R>
TileKernelFn<D, R> pick_tile_kernel(int block) {
  if (block == 512) return &k_batch_tiles<D, R, 512>;
  if constexpr (D == 2) {
    if (block == 256) return &k_batch_tiles<D, R, 256>;
    if (block == 1024) return &k_batch_tiles<D, R, 1024>;
  }
  throw std::invalid_argument("unsupported tiled batch block size");
}

}  // namespace

struct casl_batch {
  casl_stencil st;  // concatenated tables + scratch (phi0, dd0, flags); tiled encodings of all trees
  int32_t n_trees = 0;
  int32_t max_tree = 0;
  bool tiled = false;  // every tree has a tiled encoding (k_batch_tiles usable)
  std::string tile_error;
  casl::DevBuf<int32_t> tree_tile;  // [n_trees + 1] tile ranges
  casl::DevBuf<int64_t> off;
  casl::DevBuf<int32_t> order;
  casl::DevBuf<unsigned int> queue;
  Schedule sched;
  int cus = 0;
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double *g_phi = nullptr;
  int g_iter = -1, g_prec = -1, g_block = 0, g_npt = 0;

  ~casl_batch() { drop_graph(); }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

template <int D, class R>
void launch_batch(casl_batch *B, double *phi, int n_iter, hipStream_t s, bool allocate_only) {
  casl_stencil &st = B->st;
  const int64_t N = st.N;
  constexpr int F = 2 * D;
  st.phi0.ensure(N);
  st.dd0.ensure(size_t(D) * N);
  st.flags.ensure(N);
  BatchArgs<D, R> A;
  A.t = st.tables<D>();
  A.off = B->off.p;
  A.order = B->order.p;
  A.n_trees = B->n_trees;
  A.queue = B->queue.p;
  A.phi = phi;
  A.phi0 = st.phi0.p;
  A.dd0 = st.dd0.p;
  A.flags = st.flags.p;
  A.n_iter = n_iter;
  A.zdt = st.zdt;
  // scratch: arms, phi_np1 and phi_n working copies, pass-A output
  if constexpr (sizeof(R) == 8) {
    st.sfix64.ensure(size_t(F) * N);
    st.tmp64.ensure(2 * size_t(N));
    st.ddn64.ensure(size_t(D) * N);
    A.sfix = st.sfix64.p;
    A.nxt = st.tmp64.p;
    A.cur = st.tmp64.p + N;
    A.ddn = st.ddn64.p;
  } else {
    st.sfixf.ensure(size_t(F) * N);
    st.tmpf.ensure(2 * size_t(N));
    st.ddnf.ensure(size_t(D) * N);
    A.sfix = st.sfixf.p;
    A.nxt = st.tmpf.p;
    A.cur = st.tmpf.p + N;
    A.ddn = st.ddnf.p;
  }
  if (allocate_only) return;
  if (B->sched.npt == 0) {
    if (!B->tiled) throw std::invalid_argument("the tiled batch kernel is not available: " + B->tile_error);
    BatchTileArgs<R> P{};
    TileArgs<R> &T = P.T;
    T = tile_args<R>(&st);
    T.flags = st.flags.p;
    T.sfix = A.sfix;
    T.phi0 = st.phi0.p;
    T.N = N;
    P.off = B->off.p;
    P.order = B->order.p;
    P.tree_tile = B->tree_tile.p;
    P.queue = B->queue.p;
    P.phi = phi;
    P.cur = A.cur;
    P.nxt = A.nxt;
    P.n_trees = B->n_trees;
    P.n_iter = n_iter;
    const unsigned g = unsigned((N + 255) / 256);
    hipLaunchKernelGGL((k_batch_copy<R>), dim3(g), dim3(256), 0, s, phi, st.phi0.p, A.cur, N);
    // flags and s_fix of every tree in tile-record order (FP64 tiled prep)
    TileArgs<double> tp = tile_args<double>(&st);
    tp.phin = st.phi0.p;
    hipLaunchKernelGGL((k_tile_prep<D, R, true>), dim3(unsigned(st.n_tiles)), dim3(TILE_BLOCK_F64), 0, s, tp,
                       (const uint8_t *)nullptr, -DBL_MAX, DBL_MAX, st.zdt, st.flags.p, A.sfix);
    HIP_CHECK(hipGetLastError());
    TileKernelFn<D, R> fn = pick_tile_kernel<D, R>(B->sched.block);
    int occ = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn, B->sched.block, 0));
    occ = std::max(occ, 1);
    const int grid = std::max(1, std::min<int>(B->n_trees, B->cus * occ));
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(1), 0, s, B->queue.p);
    prof_begin("k_batch", s);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(B->sched.block), 0, s, P);
    prof_end("k_batch", s);
    HIP_CHECK(hipGetLastError());
    return;
  }
  KernelFn<D, R> fn = pick_kernel<D, R>(B->sched);
  int occ = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn, B->sched.block, 0));
  occ = std::max(occ, 1);
  const int grid = std::max(1, std::min<int>(B->n_trees, B->cus * occ));
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(1), 0, s, B->queue.p);
  prof_begin("k_batch", s);
  hipLaunchKernelGGL(fn, dim3(grid), dim3(B->sched.block), 0, s, A);
  prof_end("k_batch", s);
  HIP_CHECK(hipGetLastError());
}

// Tiled encodings of every tree (built in parallel), checked one by one and
// concatenated: node, halo, record and extension indices offset per tree.
void build_batch_tiles(casl_batch *B, const casl_flat_stencil *trees, const std::vector<int64_t> &off) {
  const int32_t nt = B->n_trees;
  const int D = trees[0].dim;
  const int WORDS = D == 2 ? HangExt<2>::WORDS : HangExt<3>::WORDS, SW = D == 2 ? HangExt<2>::SW : HangExt<3>::SW;
  const int NH = D == 2 ? HangExt<2>::NH : HangExt<3>::NH;
  std::vector<TileHost> hosts(static_cast<size_t>(nt));
  std::vector<std::string> errs(static_cast<size_t>(nt));
  parallel_for(nt, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) {
      try {
        build_tiles(trees[k], hosts[size_t(k)]);
        validate_tile_host(hosts[size_t(k)], D, trees[k].n_nodes);
      } catch (const std::exception &ex) {
        errs[size_t(k)] = ex.what();
      }
    }
  }, 1);
  for (int32_t k = 0; k < nt; ++k)
    if (!errs[size_t(k)].empty()) {
      B->tile_error = "tree " + std::to_string(k) + ": " + errs[size_t(k)];
      return;
    }
  // one geometry for the batch: equal coarsest arm cells, extension offsets
  // rescaled to the finest unit of all trees (exact: powers of two)
  TileHost all;
  all.dim = D;
  for (int32_t k = 0; k < nt; ++k) {
    for (int a = 0; a < 3; ++a)
      if (hosts[size_t(k)].size[a] != hosts[0].size[a]) {
        B->tile_error = "trees differ in their coarsest cell size";
        return;
      }
    all.lmax = std::max(all.lmax, hosts[size_t(k)].lmax);
  }
  for (int a = 0; a < 3; ++a) all.size[a] = hosts[0].size[a];
  all.tile_start.push_back(0);
  all.halo_start.push_back(0);
  all.rec_start.push_back(0);
  all.ext_start.push_back(0);
  all.urec_start.push_back(0);
  const int UW = D == 2 ? UniRec<2>::WORDS : UniRec<3>::WORDS;
  std::vector<int32_t> tree_tile(static_cast<size_t>(nt) + 1, 0);
  const int C = 1 << (D - 1);
  for (int32_t k = 0; k < nt; ++k) {
    const TileHost &H = hosts[size_t(k)];
    const int32_t node_off = int32_t(off[size_t(k)]);
    const int32_t halo_off = int32_t(all.halo_idx.size());
    const int32_t rec_off = int32_t(all.rec.size());
    const int32_t ext_off = int32_t(all.ext.size() / size_t(WORDS));
    const int32_t urec_off = int32_t(all.urec.size() / size_t(UW));
    const int64_t ntile = int64_t(H.tile_start.size()) - 1;
    tree_tile[size_t(k)] = int32_t(all.tile_start.size()) - 1;
    for (int64_t t = 0; t < ntile; ++t) {
      all.tile_start.push_back(H.tile_start[size_t(t) + 1] + node_off);
      all.halo_start.push_back(H.halo_start[size_t(t) + 1] + halo_off);
      all.rec_start.push_back(H.rec_start[size_t(t) + 1] + rec_off);
      all.h1count.push_back(H.h1count[size_t(t)]);
      all.ext_start.push_back(H.ext_start[size_t(t) + 1] + ext_off);
      all.nreg.push_back(H.nreg[size_t(t)]);
      all.urec_start.push_back(H.urec_start[size_t(t) + 1] + urec_off);
      all.nuni.push_back(H.nuni[size_t(t)]);
    }
    // uniform records hold tile-local indices and levels of the shared
    // coarsest cell only: they need no offset
    all.urec.insert(all.urec.end(), H.urec.begin(), H.urec.end());
    for (int32_t g : H.halo_idx) all.halo_idx.push_back(g + node_off);
    all.rec.insert(all.rec.end(), H.rec.begin(), H.rec.end());
    all.rec_slot.insert(all.rec_slot.end(), H.rec_slot.begin(), H.rec_slot.end());
    const uint32_t scale = 1u << (all.lmax - H.lmax);
    for (size_t q = 0; q < H.ext.size(); ++q) {
      uint32_t v = H.ext[q];
      const int w = int(q % size_t(WORDS));  // header, then SW words per hanging face
      if (w >= 1 && w - 1 < NH * SW && (w - 1) % SW >= C) {  // transverse offset words
        v *= scale;
        if (v > 0xffffu) {
          B->tile_error = "extension offset does not fit the batch's finest unit";
          return;
        }
      }
      all.ext.push_back(uint16_t(v));
    }
    all.n_irregular += H.n_irregular;
  }
  tree_tile[size_t(nt)] = int32_t(all.tile_start.size()) - 1;
  const int64_t N = off[size_t(nt)];
  validate_tile_host(all, D, N);
  upload_tile_host(all, N, &B->st);
  // k_batch_tiles keeps the dyadic shortcuts in both precisions
  if (!B->st.dyadic_geometry()) {
    B->tile_error = "the trees' coarsest cells are not powers of two";
    return;
  }
  B->tree_tile.ensure(tree_tile.size());
  HIP_CHECK(hipMemcpy(B->tree_tile.p, tree_tile.data(), tree_tile.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  B->tiled = true;
}

void batch_run(casl_batch *B, double *phi, int n_iter, int precision, hipStream_t s, bool allocate_only = false) {
  if (B->st.dim == 2) {
    if (precision == CASL_PRECISION_FP64)
      launch_batch<2, double>(B, phi, n_iter, s, allocate_only);
    else
      launch_batch<2, float>(B, phi, n_iter, s, allocate_only);
  } else {
    if (precision == CASL_PRECISION_FP64)
      launch_batch<3, double>(B, phi, n_iter, s, allocate_only);
    else
      launch_batch<3, float>(B, phi, n_iter, s, allocate_only);
  }
}

void check_run_args(casl_batch *B, double *phi, int n_iter, int precision) {
  if (!B || !phi) throw std::invalid_argument("null argument");
  if (n_iter < 0) throw std::invalid_argument("negative number of iterations");
  if (precision != CASL_PRECISION_FP64 && precision != CASL_PRECISION_MIXED)
    throw std::invalid_argument("unknown precision mode");
}

}  // namespace

extern "C" {

int casl_batch_create(const casl_flat_stencil *trees, int32_t n_trees, casl_batch_t **out) {
  return guarded([&]() {
    if (!out || !trees || n_trees <= 0) throw std::invalid_argument("bad arguments");
    *out = nullptr;
    const int D = trees[0].dim;
    if (D != 2 && D != 3) throw std::invalid_argument("dim must be 2 or 3");
    const int F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
    std::vector<int64_t> off(size_t(n_trees) + 1, 0);
    double zdt = trees[0].zero_distance_threshold;
    for (int32_t k = 0; k < n_trees; ++k) {
      if (trees[k].dim != D) throw std::invalid_argument("all trees must have the same dimension");
      if (trees[k].n_nodes <= 0) throw std::invalid_argument("empty tree");
      if (trees[k].zero_distance_threshold != zdt)
        throw std::invalid_argument("all trees must share the zero distance threshold");
      off[k + 1] = off[k] + trees[k].n_nodes;
    }
    const int64_t N = off[n_trees];
    if (N >= int64_t(INT32_MAX)) throw std::invalid_argument("batch too large for int32 node indices");
    std::vector<int32_t> nbr(size_t(F) * C * N);
    std::vector<double> d(size_t(F) * N), dtr(size_t(F) * T * N);
    std::vector<uint8_t> wall(N);
    int32_t max_tree = 0;
    for (int32_t k = 0; k < n_trees; ++k) {
      const casl_flat_stencil &tr = trees[k];
      const int64_t n = tr.n_nodes, o = off[k];
      max_tree = std::max<int32_t>(max_tree, int32_t(n));
      for (int fc = 0; fc < F * C; ++fc)
        for (int64_t i = 0; i < n; ++i) {
          const int32_t v = tr.nbr[size_t(fc) * n + i];
          if (v < 0 || v >= n) throw std::runtime_error("tree stencil index out of range");
          nbr[size_t(fc) * N + o + i] = int32_t(o + v);
        }
      for (int f = 0; f < F; ++f) std::memcpy(&d[size_t(f) * N + o], tr.d + size_t(f) * n, n * sizeof(double));
      for (int ft = 0; ft < F * T; ++ft)
        std::memcpy(&dtr[size_t(ft) * N + o], tr.d_tr + size_t(ft) * n, n * sizeof(double));
      std::memcpy(&wall[o], tr.wall, n);
    }
    auto B = std::make_unique<casl_batch>();
    casl_flat_stencil all{};
    all.dim = D;
    all.n_nodes = N;
    all.nbr = nbr.data();
    all.d = d.data();
    all.d_tr = dtr.data();
    all.wall = wall.data();
    all.zero_distance_threshold = zdt;
    upload_tables(&all, &B->st, false, false);  // the batch kernels read FP32 distances only
    B->n_trees = n_trees;
    B->max_tree = max_tree;
    std::vector<int32_t> order(n_trees);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
    B->off.ensure(off.size());
    B->order.ensure(order.size());
    B->queue.ensure(1);
    HIP_CHECK(hipMemcpy(B->off.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(B->order.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    build_batch_tiles(B.get(), trees, off);
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, B->st.device));
    B->cus = prop.multiProcessorCount;
    *out = B.release();
  });
}

int casl_batch_destroy(casl_batch_t *b) {
  delete b;
  return CASL_OK;
}

int casl_batch_info(const casl_batch_t *b, int64_t *total_nodes, int32_t *n_trees, int32_t *max_tree_nodes) {
  return guarded([&]() {
    if (!b) throw std::invalid_argument("null batch");
    if (total_nodes) *total_nodes = b->st.N;
    if (n_trees) *n_trees = b->n_trees;
    if (max_tree_nodes) *max_tree_nodes = b->max_tree;
  });
}

int casl_batch_set_schedule(casl_batch_t *b, int32_t block_threads, int32_t nodes_per_thread) {
  return guarded([&]() {
    if (!b) throw std::invalid_argument("null batch");
    if (nodes_per_thread < 0) throw std::invalid_argument("nodes per thread must be >= 0");
    Schedule s;
    s.block = block_threads > 0 ? block_threads : (nodes_per_thread == 0 ? 512 : 256);
    s.npt = nodes_per_thread;
    if (s.npt == 0) {
      if (!b->tiled) throw std::invalid_argument("the tiled batch kernel is not available: " + b->tile_error);
      if (!(b->st.dim == 2 ? tile_block_ok<2>(s.block) : tile_block_ok<3>(s.block)))
        throw std::invalid_argument("tiled batch kernel: block must be 512 (or 256 in 2D)");
    } else {
      (void)pick_kernel<2, double>(s);  // validates the pair
    }
    b->sched = s;
  });
}

int casl_batch_tile_info(const casl_batch_t *b, int32_t *tiled, int64_t *n_tiles) {
  return guarded([&]() {
    if (!b) throw std::invalid_argument("null batch");
    if (tiled) *tiled = b->tiled ? 1 : 0;
    if (n_tiles) *n_tiles = b->tiled ? b->st.n_tiles : 0;
  });
}

int casl_batch_get_schedule(const casl_batch_t *b, int32_t *block_threads, int32_t *nodes_per_thread) {
  return guarded([&]() {
    if (!b) throw std::invalid_argument("null batch");
    if (block_threads) *block_threads = b->sched.block;
    if (nodes_per_thread) *nodes_per_thread = b->sched.npt;
  });
}

int casl_batch_reinit(casl_batch_t *b, double *phi, int n_iter, int precision, void *stream) {
  return guarded([&]() {
    check_run_args(b, phi, n_iter, precision);
    batch_run(b, phi, n_iter, precision, static_cast<hipStream_t>(stream));
  });
}

int casl_batch_reinit_graph(casl_batch_t *b, double *phi, int n_iter, int precision, void *stream) {
  return guarded([&]() {
    check_run_args(b, phi, n_iter, precision);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!b->exec || b->g_phi != phi || b->g_iter != n_iter || b->g_prec != precision ||
        b->g_block != b->sched.block || b->g_npt != b->sched.npt) {
      b->drop_graph();
      // scratch must exist before capture (no allocation inside the graph)
      hipStream_t cap;
      HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      try {
        batch_run(b, phi, n_iter, precision, cap, true);  // scratch only, nothing launched
        HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        batch_run(b, phi, n_iter, precision, cap);
        HIP_CHECK(hipStreamEndCapture(cap, &b->graph));
        HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      } catch (...) {
        (void)hipStreamDestroy(cap);
        b->drop_graph();
        throw;
      }
      HIP_CHECK(hipStreamDestroy(cap));
      b->g_phi = phi;
      b->g_iter = n_iter;
      b->g_prec = precision;
      b->g_block = b->sched.block;
      b->g_npt = b->sched.npt;
    }
    HIP_CHECK(hipGraphLaunch(b->exec, s));
  });
}

int casl_batch_autotune(casl_batch_t *b, const double *phi, int n_iter, int precision, void *stream,
                        int32_t *best_block, int32_t *best_npt, float *best_ms) {
  return guarded([&]() {
    check_run_args(b, const_cast<double *>(phi), n_iter, precision);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t N = b->st.N;
    double *work = nullptr;
    HIP_CHECK(hipMalloc(&work, sizeof(double) * N));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    Schedule best = b->sched;
    float best_t = 1e30f;
    std::vector<Schedule> cands;
    for (int bt : kBlocks)
      for (int npt : kNpts) cands.push_back(Schedule{bt, npt});
    if (b->tiled) {
      for (int bt : {256, 512, 1024})
        if (b->st.dim == 2 ? tile_block_ok<2>(bt) : tile_block_ok<3>(bt)) cands.push_back(Schedule{bt, 0});
    }
    try {
      for (const Schedule &c : cands) {
        {
          const int bt = c.block, npt = c.npt;
          b->sched.block = bt;
          b->sched.npt = npt;
          float tsum = 0;
          for (int rep = 0; rep < 3; ++rep) {
            HIP_CHECK(hipMemcpyAsync(work, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
            HIP_CHECK(hipEventRecord(e0, s));
            batch_run(b, work, n_iter, precision, s);
            HIP_CHECK(hipEventRecord(e1, s));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tsum += ms;  // first run warms caches
          }
          const float t = tsum / 2;
          if (t < best_t) {
            best_t = t;
            best.block = bt;
            best.npt = npt;
          }
        }
      }
    } catch (...) {
      (void)hipFree(work);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      b->sched = best;
      throw;
    }
    (void)hipFree(work);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->sched = best;
    if (best_block) *best_block = best.block;
    if (best_npt) *best_npt = best.npt;
    if (best_ms) *best_ms = best_t;
  });
}

}  // extern "C"
