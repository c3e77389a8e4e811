// Grid construction on the GPU: the single-tree forest refined by
// refine_levelset_cf, p4est's node numbering and construct_neighbors, so a
// changed level set does not send the grid through the host.
//
//   refinement      refine_levelset_cf applied level by level
//                   (src/my_p4est_refine_coarsen.cpp:17-89): one decision per
//                   quadrant, an exclusive scan of the child counts, children
//                   written in place of their parent -- the leaves stay in
//                   Morton order, the order p4est's refine leaves them in
//   node numbering  my_p4est_nodes_new for one tree (src/my_p4est_nodes.c:
//                   383-446, 614-671): clamped corner keys, radix sort, unique;
//                   a node's index is its rank in that Morton order
//   neighbours      construct_neighbors (src/my_p4est_node_neighbors.cpp:114-1201):
//                   the cells around every node (leaf-corner scatter, then point
//                   location by binary search over the leaf keys), the finest
//                   cell per face, corner ids, distances, wall arms
// The arithmetic is that of the host construction (csrc/forest.cpp), so the
// tables are the same (the level-set values may differ in the last unit in
// the last place where they use sin, cos or atan2).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <memory>
#include <vector>

#include "reinit_device.hpp"

using namespace casl;

namespace {

constexpr int BLOCK = 256;
// One thread per item; a launch holds fewer than 2^32 work-items.
inline unsigned grid_for(int64_t n) {
  if (n >= (int64_t(1) << 32) - BLOCK) throw std::runtime_error("grid too large for one launch");
  return unsigned((n + BLOCK - 1) / BLOCK);
}

struct DQuad {
  int32_t c[3];
  int32_t level;
};

struct DevLS {
  int kind;
  int n_items;
  const double *p;  // device parameters
};

// The level sets of casl_levelset_eval, evaluated on the device.
__device__ double dev_levelset(const DevLS &ls, int dim, const double *x) {
  const double *p = ls.p;
  switch (ls.kind) {
    case CASL_LS_STAR: {
      double dx = x[0] - p[0], dy = x[1] - p[1];
      double r2 = dx * dx + dy * dy;
      double th = atan2(dy, dx);
      double r = p[2] + p[3] * cos(p[4] * th);
      return (sqrt(r2) - r) * (1.0 + p[5] * r2);
    }
    case CASL_LS_SPHERES: {
      double m = p[0];
      double best = HUGE_VAL;
      for (int i = 0; i < ls.n_items; ++i) {
        const double *s = p + 1 + 4 * i;
        double dx = x[0] - s[0], dy = x[1] - s[1], dz = x[2] - s[2];
        double v = sqrt(dx * dx + dy * dy + dz * dz) - s[3];
        if (v < best) best = v;
      }
      return best * (1.0 + m * sin(2.0 * M_PI * x[0]) * cos(2.0 * M_PI * x[1]));
    }
    case CASL_LS_GYROID: {
      double w = p[0];
      double g = sin(w * x[0]) * cos(w * x[1]) + sin(w * x[1]) * cos(w * x[2]) + sin(w * x[2]) * cos(w * x[0]) - p[1];
      return g * (1.0 + p[2] * x[0] * x[1] * x[2]);
    }
    case CASL_LS_SPHERE: {
      double r2 = 0.0;
      for (int a = 0; a < dim; ++a) r2 += (x[a] - p[a]) * (x[a] - p[a]);
      return sqrt(r2) - p[dim];
    }
    case CASL_LS_PLANE: {
      double v = 0.0;
      for (int a = 0; a < dim; ++a) v += p[a] * x[a];
      return v + p[dim];
    }
    case CASL_LS_MODULATED: {
      double dx = x[0] - p[0], dy = x[1] - p[1];
      double e = sqrt(dx * dx + dy * dy) - p[2];
      return e * (1.0 + p[3] * sin(5.0 * x[0]) * cos(4.0 * x[1]));
    }
  }
  return 0.0;
}

template <int D>
struct Geo {
  static constexpr int MAXLEVEL = (D == 2 ? 30 : 19);
  static constexpr int32_t ROOT_LEN = int32_t(1) << MAXLEVEL;
  static constexpr int CH = 1 << D, F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
  double bmin[3], bmax[3];
  __host__ __device__ static int32_t qlen(int level) { return int32_t(1) << (MAXLEVEL - level); }
  __host__ __device__ double size(int a) const { return bmax[a] - bmin[a]; }
};

__device__ __forceinline__ uint64_t dspread2(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
__device__ __forceinline__ uint32_t dcompact2(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return uint32_t(x);
}
__device__ __forceinline__ uint64_t dspread3(uint32_t v) {
  uint64_t x = v & 0x1FFFFF;
  x = (x | (x << 32)) & 0x001F00000000FFFFull;
  x = (x | (x << 16)) & 0x001F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__device__ __forceinline__ uint32_t dcompact3(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x | (x >> 2)) & 0x10C30C30C30C30C3ull;
  x = (x | (x >> 4)) & 0x100F00F00F00F00Full;
  x = (x | (x >> 8)) & 0x001F0000FF0000FFull;
  x = (x | (x >> 16)) & 0x001F00000000FFFFull;
  x = (x | (x >> 32)) & 0x00000000001FFFFFull;
  return uint32_t(x);
}
// bit b of axis a -> bit D*b + a (z outranks y outranks x at equal b)
template <int D>
__device__ __forceinline__ uint64_t dmorton(const int32_t *c) {
  if constexpr (D == 2)
    return dspread2(uint32_t(c[0])) | (dspread2(uint32_t(c[1])) << 1);
  else
    return dspread3(uint32_t(c[0])) | (dspread3(uint32_t(c[1])) << 1) | (dspread3(uint32_t(c[2])) << 2);
}
template <int D>
__device__ __forceinline__ void ddemorton(uint64_t k, int32_t *c) {
  if constexpr (D == 2) {
    c[0] = int32_t(dcompact2(k));
    c[1] = int32_t(dcompact2(k >> 1));
  } else {
    c[0] = int32_t(dcompact3(k));
    c[1] = int32_t(dcompact3(k >> 1));
    c[2] = int32_t(dcompact3(k >> 2));
  }
}

// ---------------------------------------------------------------- refinement
// refine_levelset_cf's decision (csrc/forest.cpp decide(), same arithmetic)
template <int D>
__device__ bool decide(const Geo<D> &G, const DQuad &q, int min_lvl, int max_lvl, double lip, double band_factor,
                       const DevLS &ls) {
  using GG = Geo<D>;
  if (q.level < min_lvl) return true;
  if (q.level >= max_lvl) return false;
  double dmin = (double)GG::qlen(q.level) / (double)GG::ROOT_LEN;
  double dx[3], x[3];
  for (int a = 0; a < D; ++a) dx[a] = G.size(a) * dmin;
  double smallest = fmax(G.size(0), G.size(1)) * ((double)GG::qlen(max_lvl)) / ((double)GG::ROOT_LEN);
  if (D == 3) smallest = fmax(smallest, G.size(2) * ((double)GG::qlen(max_lvl)) / ((double)GG::ROOT_LEN));
  double d2 = dx[0] * dx[0] + dx[1] * dx[1];
  if (D == 3) d2 = d2 + dx[2] * dx[2];
  double d = sqrt(d2);
  for (int a = 0; a < D; ++a) x[a] = G.size(a) * (double)q.c[a] / (double)GG::ROOT_LEN + G.bmin[a];
  double band = band_factor * smallest;
  double f[8];
  for (int k = 0; k < GG::CH; ++k) {
    double xc[3];
    for (int a = 0; a < D; ++a) xc[a] = x[a] + ((k >> a) & 1) * dx[a];
    f[k] = dev_levelset(ls, D, xc);
    if (fabs(f[k]) - band <= 0.5 * lip * d) return true;
  }
  if (D == 3) {
    if (f[0] * f[1] < 0 || f[0] * f[2] < 0 || f[1] * f[3] < 0 || f[2] * f[3] < 0 || f[3] * f[4] < 0 ||
        f[4] * f[5] < 0 || f[5] * f[6] < 0 || f[6] * f[7] < 0)
      return true;
  } else {
    if (f[0] * f[1] < 0 || f[0] * f[2] < 0 || f[1] * f[3] < 0 || f[2] * f[3] < 0) return true;
  }
  return false;
}

// Pass p decides the quadrants of level p: a coarser leaf was decided in an
// earlier pass, and the decision depends on the quadrant alone.
template <int D>
__global__ void k_decide(Geo<D> G, const DQuad *__restrict__ q, int64_t L, int pass, int min_lvl, int max_lvl,
                         double lip, double band, DevLS ls, int64_t *__restrict__ cnt) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= L) return;
  const bool r = q[l].level == pass && q[l].level < Geo<D>::MAXLEVEL - 1 &&
                 decide<D>(G, q[l], min_lvl, max_lvl, lip, band, ls);
  cnt[l] = r ? Geo<D>::CH : 1;
}

template <int D>
__global__ void k_emit(const DQuad *__restrict__ q, int64_t L, const int64_t *__restrict__ cnt,
                       const int64_t *__restrict__ off, DQuad *__restrict__ out) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= L) return;
  const DQuad p = q[l];
  if (cnt[l] == 1) {
    out[off[l]] = p;
    return;
  }
  const int32_t h = Geo<D>::qlen(p.level + 1);
  for (int k = 0; k < Geo<D>::CH; ++k) {
    DQuad c = p;
    c.level = p.level + 1;
    for (int a = 0; a < D; ++a) c.c[a] = p.c[a] + ((k >> a) & 1) * h;
    out[off[l] + k] = c;
  }
}

// ------------------------------------------------------------------ nodes
template <int D>
__global__ void k_corner_keys(const DQuad *__restrict__ q, int64_t L, uint64_t *__restrict__ keys) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= L) return;
  const int32_t h = Geo<D>::qlen(q[l].level);
  for (int k = 0; k < Geo<D>::CH; ++k) {
    int32_t p[3];
    for (int a = 0; a < D; ++a) {
      p[a] = q[l].c[a] + ((k >> a) & 1) * h;
      if (p[a] == Geo<D>::ROOT_LEN) p[a] = Geo<D>::ROOT_LEN - 1;  // p4est_node_clamp_inside
    }
    keys[l * Geo<D>::CH + k] = dmorton<D>(p);
  }
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t *a, int64_t n, uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int64_t upper_bound_u64(const uint64_t *a, int64_t n, uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_local_nodes(const uint64_t *__restrict__ keys, int64_t K, const uint64_t *__restrict__ nodes,
                              int64_t N, int32_t *__restrict__ local) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < K) local[i] = int32_t(lower_bound_u64(nodes, N, keys[i]));
}

template <int D>
__global__ void k_node_ixyz(const uint64_t *__restrict__ nodes, int64_t N, int32_t *__restrict__ ixyz) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int32_t c[3];
  ddemorton<D>(nodes[n], c);
  for (int a = 0; a < D; ++a) ixyz[n * D + a] = (c[a] == Geo<D>::ROOT_LEN - 1 ? Geo<D>::ROOT_LEN : c[a]);
}

// ------------------------------------------------------------- neighbours
template <int D>
__global__ void k_leaf_keys(const DQuad *__restrict__ q, int64_t L, uint64_t *__restrict__ keys) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l < L) keys[l] = dmorton<D>(q[l].c);
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// the cell at corner k of leaf l is the cell in direction k^(CH-1) of that corner's node
template <int D>
__global__ void k_scatter_cells(const int32_t *__restrict__ local, int64_t L, int32_t *__restrict__ cell) {
  constexpr int CH = Geo<D>::CH;
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= L) return;
  for (int k = 0; k < CH; ++k) cell[int64_t(local[l * CH + k]) * CH + (k ^ (CH - 1))] = int32_t(l);
}

template <int D>
__global__ void k_locate(const int32_t *__restrict__ ixyz, int64_t N, const uint64_t *__restrict__ leaf_keys,
                         int64_t L, int32_t *__restrict__ cell) {
  constexpr int CH = Geo<D>::CH;
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= N) return;
  for (int dir = 0; dir < CH; ++dir) {
    int32_t &slot = cell[n * CH + dir];
    if (slot != -2) continue;
    int32_t p[3];
    bool inside = true;
    for (int a = 0; a < D; ++a) {
      p[a] = ixyz[n * D + a] + (((dir >> a) & 1) ? 1 : -1);
      if (p[a] < 0 || p[a] > Geo<D>::ROOT_LEN) inside = false;
    }
    if (!inside) {
      slot = -1;
      continue;
    }
    slot = int32_t(upper_bound_u64(leaf_keys, L, dmorton<D>(p)) - 1);
  }
}

template <int D>
struct NbrOut {
  int32_t *nbr;
  double *dface, *dtr;
  uint8_t *wall;
  int *bad;
};

template <int D>
__device__ void fill_from_cell(const Geo<D> &G, const DQuad *leaves, const int32_t *local, const int32_t *ixyz,
                               int64_t N, int64_t n, int f, int32_t qi, int side, const NbrOut<D> &o) {
  using GG = Geo<D>;
  const int a = f >> 1;
  int t1, t2;
  if (D == 2) {
    t1 = 1 - a;
    t2 = -1;
  } else {
    t1 = (a == 0 ? 1 : 0);
    t2 = (a == 2 ? 1 : 2);
  }
  const DQuad Q = leaves[qi];
  for (int c = 0; c < GG::C; ++c) {
    int corner = (side << a) | ((c & 1) << t1);
    if (D == 3) corner |= ((c >> 1) & 1) << t2;
    o.nbr[(int64_t(f) * GG::C + c) * N + n] = local[int64_t(qi) * GG::CH + corner];
  }
  const double qh = GG::qlen(Q.level) / (double)GG::ROOT_LEN;
  for (int t = 0; t < D - 1; ++t) {
    const int ax = (t == 0 ? t1 : t2);
    const double thr = G.size(ax) * ((double)1) / ((double)GG::ROOT_LEN);
    const double X = G.size(ax) * (ixyz[n * D + ax] / (double)GG::ROOT_LEN) + G.bmin[ax];
    const double QX = G.size(ax) * (Q.c[ax] / (double)GG::ROOT_LEN) + G.bmin[ax];
    const double m0 = (fabs(X - QX) > thr) ? (X - QX) : 0.0;
    const double p0 = (fabs(G.size(ax) * qh - m0) > thr) ? (G.size(ax) * qh - m0) : 0.0;
    o.dtr[(int64_t(f) * GG::T + 2 * t + 0) * N + n] = m0;
    o.dtr[(int64_t(f) * GG::T + 2 * t + 1) * N + n] = p0;
  }
}

template <int D>
__global__ void k_construct(Geo<D> G, const DQuad *__restrict__ leaves, const int32_t *__restrict__ local,
                            const int32_t *__restrict__ ixyz, const int32_t *__restrict__ cell, int64_t N,
                            NbrOut<D> o) {
  using GG = Geo<D>;
  constexpr int F = GG::F, CH = GG::CH;
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int32_t *cn = cell + n * CH;
  auto lvl = [&](int dir) { return cn[dir] < 0 ? -1 : leaves[cn[dir]].level; };
  bool is_wall[6] = {false, false, false, false, false, false};
  for (int f = 0; f < F; ++f) {
    const int a = f >> 1, s = f & 1;
    int best = -1;
    bool any = false;
    for (int dir = 0; dir < CH; ++dir) {
      if (((dir >> a) & 1) != s) continue;
      if (cn[dir] >= 0) any = true;
      if (best < 0) {
        best = dir;
        continue;
      }
      if (lvl(best) < lvl(dir)) best = dir;  // finest cell, first wins on ties
    }
    if (!any) {
      is_wall[f] = true;
      continue;
    }
    const int32_t qi = cn[best];
    o.dface[int64_t(f) * N + n] = G.size(a) * (GG::qlen(leaves[qi].level) / (double)GG::ROOT_LEN);
    fill_from_cell<D>(G, leaves, local, ixyz, N, n, f, qi, s == 0 ? 0 : 1, o);
  }
  for (int f = 0; f < F; ++f) {
    if (!is_wall[f]) continue;
    const int a = f >> 1, s = f & 1, so = 1 - s;
    int dmin = -1, lbest = -1;
    for (int dir = 0; dir < CH; ++dir) {
      if (((dir >> a) & 1) != so) continue;
      if (lbest < lvl(dir)) {
        lbest = lvl(dir);
        dmin = dir;
      }
    }
    if (dmin < 0) {
      *o.bad = 1;
      return;
    }
    const int32_t qmin = cn[dmin];
    int corner = so << a;
    for (int b = 0; b < D; ++b)
      if (b != a) corner |= (((dmin >> b) & 1) ? 0 : 1) << b;
    const int32_t ntmp = local[int64_t(qmin) * CH + corner];
    const int32_t qtmp = cell[int64_t(ntmp) * CH + dmin];
    if (qtmp < 0) {
      *o.bad = 1;
      return;
    }
    const int fo = f ^ 1;
    o.dface[int64_t(f) * N + n] =
        -o.dface[int64_t(fo) * N + n] - G.size(a) * (GG::qlen(leaves[qtmp].level) / (double)GG::ROOT_LEN);
    fill_from_cell<D>(G, leaves, local, ixyz, N, n, f, qtmp, so, o);
  }
  uint8_t w = 0;
  for (int a = 0; a < D; ++a) {
    const int32_t x = ixyz[n * D + a];
    if (x == 0) w |= uint8_t(1u << (2 * a));
    if (x == GG::ROOT_LEN - 1 || x == GG::ROOT_LEN) w |= uint8_t(1u << (2 * a + 1));
  }
  o.wall[n] = w;
}

template <int D>
__global__ void k_sample(Geo<D> G, const int32_t *__restrict__ ixyz, int64_t N, DevLS ls, double *__restrict__ out) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double x[3];
  for (int a = 0; a < D; ++a) x[a] = G.size(a) * (ixyz[n * D + a] / (double)Geo<D>::ROOT_LEN) + G.bmin[a];
  out[n] = dev_levelset(ls, D, x);
}

int ls_param_count(const casl_levelset &ls, int dim) {
  switch (ls.kind) {
    case CASL_LS_STAR:
      if (dim != 2) throw std::invalid_argument("star level set is 2D");
      return 6;
    case CASL_LS_SPHERES:
      if (dim != 3) throw std::invalid_argument("spheres level set is 3D");
      if (ls.n_items < 0) throw std::invalid_argument("negative sphere count");
      return 1 + 4 * ls.n_items;
    case CASL_LS_GYROID:
      if (dim != 3) throw std::invalid_argument("gyroid level set is 3D");
      return 3;
    case CASL_LS_SPHERE:
    case CASL_LS_PLANE:
      return dim + 1;
    case CASL_LS_MODULATED:
      if (dim != 2) throw std::invalid_argument("modulated level set is 2D");
      return 4;
    case CASL_LS_CALLBACK:
      throw std::invalid_argument("host callbacks cannot be evaluated on the device");
  }
  throw std::invalid_argument("unknown level-set kind");
}

// device copy of a level set's parameters
struct DevLSHolder {
  casl::DevBuf<double> p;
  DevLS ls{};
  DevLSHolder(const casl_levelset *src, int dim) {
    if (!src || !src->params) throw std::invalid_argument("null level set");
    const int np = ls_param_count(*src, dim);
    p.ensure(size_t(np));
    HIP_CHECK(hipMemcpy(p.p, src->params, size_t(np) * sizeof(double), hipMemcpyHostToDevice));
    ls.kind = src->kind;
    ls.n_items = src->n_items;
    ls.p = p.p;
  }
};

}  // namespace

struct casl_device_grid {
  int dim = 0;
  int64_t L = 0, N = 0;
  int max_level = 0;
  double bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};
  double zdt = 0;
  casl::DevBuf<DQuad> leaves;
  casl::DevBuf<int32_t> ixyz, nbr;
  casl::DevBuf<double> dface, dtr;
  casl::DevBuf<uint8_t> wall;
};

namespace {

// A hipCUB device-wide call with its own temporary storage, which stays
// allocated until the call has finished on the stream.
template <class Fn>
void cub_call(hipStream_t s, Fn fn) {
  size_t bytes = 0;
  HIP_CHECK(fn(nullptr, bytes));
  casl::DevBuf<unsigned char> tmp;
  tmp.ensure(std::max<size_t>(bytes, 1));
  HIP_CHECK(fn(tmp.p, bytes));
  HIP_CHECK(hipStreamSynchronize(s));
}

template <int D>
void build_grid(casl_device_grid *g, int min_lvl, int max_lvl, double lip, double band, const DevLS &ls,
                hipStream_t s) {
  using GG = Geo<D>;
  constexpr int CH = GG::CH;
  if (max_lvl >= GG::MAXLEVEL - 1) throw std::invalid_argument("max_lvl too large for this dimension");
  if (min_lvl < 0 || max_lvl < 0) throw std::invalid_argument("negative levels");
  Geo<D> G;
  for (int a = 0; a < 3; ++a) {
    G.bmin[a] = g->bmin[a];
    G.bmax[a] = g->bmax[a];
  }
  // refinement, level by level
  casl::DevBuf<DQuad> q, q2;
  casl::DevBuf<int64_t> cnt, off;
  q.ensure(1);
  DQuad root{};
  HIP_CHECK(hipMemcpyAsync(q.p, &root, sizeof(DQuad), hipMemcpyHostToDevice, s));
  int64_t L = 1;
  int finest = 0;
  for (int pass = 0; pass < GG::MAXLEVEL; ++pass) {
    HIP_CHECK(hipStreamSynchronize(s));  // the last pass's emit has read cnt and off
    cnt.ensure(size_t(L));
    off.ensure(size_t(L));
    hipLaunchKernelGGL(k_decide<D>, dim3(grid_for(L)), dim3(BLOCK), 0, s, G, q.p, L, pass, min_lvl, max_lvl, lip, band,
                       ls, cnt.p);
    HIP_CHECK(hipGetLastError());
    cub_call(s, [&](void *t, size_t &b) { return hipcub::DeviceScan::ExclusiveSum(t, b, cnt.p, off.p, int(L), s); });
    int64_t last_off = 0, last_cnt = 0;
    HIP_CHECK(hipMemcpyAsync(&last_off, off.p + (L - 1), sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&last_cnt, cnt.p + (L - 1), sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const int64_t L2 = last_off + last_cnt;
    if (L2 == L) break;  // nothing refined: done
    if (L2 >= int64_t(INT32_MAX)) throw std::runtime_error("too many leaves for int32 indices");
    q2.ensure(size_t(L2));
    hipLaunchKernelGGL(k_emit<D>, dim3(grid_for(L)), dim3(BLOCK), 0, s, q.p, L, cnt.p, off.p, q2.p);
    HIP_CHECK(hipGetLastError());
    q.swap(q2);
    L = L2;
    finest = pass + 1;
  }
  HIP_CHECK(hipStreamSynchronize(s));  // the last emit has finished with q2, cnt and off
  q2.release();
  cnt.release();
  off.release();
  // node numbering: sort the leaves' corner keys (in place, between two
  // buffers) and keep the distinct ones
  const int64_t K = L * CH;  // may exceed 2^31: the sort and unique take 64-bit counts
  int64_t N = 0;
  casl::DevBuf<int32_t> local;
  {
    casl::DevBuf<uint64_t> ka, kb;
    casl::DevBuf<int64_t> nsel;
    ka.ensure(size_t(K));
    kb.ensure(size_t(K));
    nsel.ensure(1);
    hipLaunchKernelGGL(k_corner_keys<D>, dim3(grid_for(L)), dim3(BLOCK), 0, s, q.p, L, ka.p);
    HIP_CHECK(hipGetLastError());
    // Every corner coordinate is a multiple of 2^z, z = MAXLEVEL - finest,
    // or ROOT_LEN - 1 (clamped): its low z bits are all 0 or all 1, so bit
    // z - 1 fixes the z - 1 bits below it.  Keys that agree from bit
    // D (z - 1) up are equal, and the sort can start there.
    const int end_bit = D * GG::MAXLEVEL, begin_bit = D * (GG::MAXLEVEL - finest - 1);
    hipcub::DoubleBuffer<uint64_t> db;
    cub_call(s, [&](void *t, size_t &b) {
      db = hipcub::DoubleBuffer<uint64_t>(ka.p, kb.p);  // the size query and the sort both start from ka
      return hipcub::DeviceRadixSort::SortKeys(t, b, db, K, begin_bit, end_bit, s);
    });
    uint64_t *sorted = db.Current(), *nodes = db.Alternate();
    cub_call(s, [&](void *t, size_t &b) { return hipcub::DeviceSelect::Unique(t, b, sorted, nodes, nsel.p, K, s); });
    HIP_CHECK(hipMemcpyAsync(&N, nsel.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    // each leaf corner's node: the corner keys again, in the sorted copy's
    // place, looked up among the distinct ones
    hipLaunchKernelGGL(k_corner_keys<D>, dim3(grid_for(L)), dim3(BLOCK), 0, s, q.p, L, sorted);
    local.ensure(size_t(K));
    hipLaunchKernelGGL(k_local_nodes, dim3(grid_for(K)), dim3(BLOCK), 0, s, sorted, K, nodes, N, local.p);
    g->ixyz.ensure(size_t(N) * D);
    hipLaunchKernelGGL(k_node_ixyz<D>, dim3(grid_for(N)), dim3(BLOCK), 0, s, nodes, N, g->ixyz.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));  // the key buffers are freed here
  }
  // neighbours
  casl::DevBuf<int32_t> cell;
  casl::DevBuf<uint64_t> leaf_keys;
  casl::DevBuf<int> bad;
  cell.ensure(size_t(N) * CH);
  leaf_keys.ensure(size_t(L));
  bad.ensure(1);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N * CH)), dim3(BLOCK), 0, s, cell.p, N * CH, -2);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, reinterpret_cast<int32_t *>(bad.p), int64_t(1), 0);
  hipLaunchKernelGGL(k_scatter_cells<D>, dim3(grid_for(L)), dim3(BLOCK), 0, s, local.p, L, cell.p);
  hipLaunchKernelGGL(k_leaf_keys<D>, dim3(grid_for(L)), dim3(BLOCK), 0, s, q.p, L, leaf_keys.p);
  hipLaunchKernelGGL(k_locate<D>, dim3(grid_for(N)), dim3(BLOCK), 0, s, g->ixyz.p, N, leaf_keys.p, L, cell.p);
  HIP_CHECK(hipGetLastError());
  g->nbr.ensure(size_t(GG::F) * GG::C * N);
  g->dface.ensure(size_t(GG::F) * N);
  g->dtr.ensure(size_t(GG::F) * GG::T * N);
  g->wall.ensure(size_t(N));
  NbrOut<D> o{g->nbr.p, g->dface.p, g->dtr.p, g->wall.p, bad.p};
  hipLaunchKernelGGL(k_construct<D>, dim3(grid_for(N)), dim3(BLOCK), 0, s, G, q.p, local.p, g->ixyz.p, cell.p, N,
                     o);
  HIP_CHECK(hipGetLastError());
  int hbad = 0;
  HIP_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (hbad) throw std::runtime_error("construct_neighbors: a wall arm has no second-degree cell");
  // keep the leaves (the refinement's buffer, handed over); the finest level present
  g->leaves.release();
  g->leaves.swap(q);
  g->L = L;
  g->N = N;
  g->max_level = finest;
}

}  // namespace

extern "C" {

int casl_device_grid_build(int dim, const double *box_min, const double *box_max, int min_lvl, int max_lvl,
                           double lip, double band, const casl_levelset *ls, void *stream,
                           casl_device_grid_t **out) {
  return guarded([&]() {
    if (!out) throw std::invalid_argument("null output");
    *out = nullptr;
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    auto g = std::make_unique<casl_device_grid>();
    g->dim = dim;
    double m = 0;
    for (int a = 0; a < dim; ++a) {
      g->bmin[a] = box_min ? box_min[a] : 0.0;
      g->bmax[a] = box_max ? box_max[a] : 1.0;
      if (!(g->bmax[a] > g->bmin[a])) throw std::invalid_argument("empty box");
      const double sz = g->bmax[a] - g->bmin[a];
      m = (a == 0) ? sz : std::min(m, sz);
    }
    g->zdt = 1e-13 * m;
    DevLSHolder h(ls, dim);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dim == 2)
      build_grid<2>(g.get(), min_lvl, max_lvl, lip, band, h.ls, s);
    else
      build_grid<3>(g.get(), min_lvl, max_lvl, lip, band, h.ls, s);
    *out = g.release();
  });
}

int casl_device_grid_destroy(casl_device_grid_t *g) {
  delete g;
  return CASL_OK;
}

int casl_device_grid_sizes(const casl_device_grid_t *g, int64_t *n_leaves, int64_t *n_nodes, int32_t *max_level) {
  return guarded([&]() {
    if (!g) throw std::invalid_argument("null grid");
    if (n_leaves) *n_leaves = g->L;
    if (n_nodes) *n_nodes = g->N;
    if (max_level) *max_level = g->max_level;
  });
}

int casl_device_grid_tables(const casl_device_grid_t *g, casl_flat_stencil *out) {
  return guarded([&]() {
    if (!g || !out) throw std::invalid_argument("null argument");
    out->dim = g->dim;
    out->n_nodes = g->N;
    out->nbr = g->nbr.p;
    out->d = g->dface.p;
    out->d_tr = g->dtr.p;
    out->wall = g->wall.p;
    out->zero_distance_threshold = g->zdt;
  });
}

int casl_device_grid_download(const casl_device_grid_t *g, int32_t *nbr, double *d, double *d_tr, uint8_t *wall,
                              int32_t *ixyz) {
  return guarded([&]() {
    if (!g) throw std::invalid_argument("null grid");
    const int D = g->dim, F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
    const size_t N = size_t(g->N);
    if (nbr) HIP_CHECK(hipMemcpy(nbr, g->nbr.p, size_t(F) * C * N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (d) HIP_CHECK(hipMemcpy(d, g->dface.p, size_t(F) * N * sizeof(double), hipMemcpyDeviceToHost));
    if (d_tr) HIP_CHECK(hipMemcpy(d_tr, g->dtr.p, size_t(F) * T * N * sizeof(double), hipMemcpyDeviceToHost));
    if (wall) HIP_CHECK(hipMemcpy(wall, g->wall.p, N, hipMemcpyDeviceToHost));
    if (ixyz) HIP_CHECK(hipMemcpy(ixyz, g->ixyz.p, size_t(D) * N * sizeof(int32_t), hipMemcpyDeviceToHost));
  });
}

int casl_device_grid_sample(const casl_device_grid_t *g, const casl_levelset *ls, double *phi, void *stream) {
  return guarded([&]() {
    if (!g || !phi) throw std::invalid_argument("null argument");
    DevLSHolder h(ls, g->dim);
    Geo<2> G2;
    Geo<3> G3;
    for (int a = 0; a < 3; ++a) {
      G2.bmin[a] = G3.bmin[a] = g->bmin[a];
      G2.bmax[a] = G3.bmax[a] = g->bmax[a];
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (g->dim == 2)
      hipLaunchKernelGGL(k_sample<2>, dim3(grid_for(g->N)), dim3(BLOCK), 0, s, G2, g->ixyz.p, g->N, h.ls, phi);
    else
      hipLaunchKernelGGL(k_sample<3>, dim3(grid_for(g->N)), dim3(BLOCK), 0, s, G3, g->ixyz.p, g->N, h.ls, phi);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));  // the parameter copy lives until here
  });
}

int casl_device_grid_stencil(const casl_device_grid_t *g, casl_stencil_t **out) {
  return guarded([&]() {
    if (!g || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    // the grid-wide tables and the tiled encoding, both from the grid's device
    // tables: nothing travels to the host but a few counts
    casl_flat_stencil tab{};
    tab.dim = g->dim;
    tab.n_nodes = g->N;
    tab.nbr = g->nbr.p;
    tab.d = g->dface.p;
    tab.d_tr = g->dtr.p;
    tab.wall = g->wall.p;
    tab.zero_distance_threshold = g->zdt;
    auto st = std::make_unique<casl_stencil>();
    upload_device_tables(tab, st.get(), nullptr);
    build_tiles_device(tab, st.get(), nullptr);
    HIP_CHECK(hipStreamSynchronize(nullptr));
    *out = st.release();
  });
}

}  // extern "C"
