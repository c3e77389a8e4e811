// Host-side forest: the part of the grid that the reference takes from p4est
// and my_p4est_node_neighbors_t, for a brick of trees on a box, periodic or
// not along each axis.
//
//   brick           my_p4est_brick_new         src/my_p4est_tools.c:37-170
//   refinement      refine_levelset_cf        src/my_p4est_refine_coarsen.cpp:17-89
//                   iterated max_lvl times    examples/reinitialization/main_2d.cpp:169-173
//   node numbering  my_p4est_nodes_new        src/my_p4est_nodes.c:383-446, 614-671
//                   (clamped corner nodes in their canonical tree,
//                    de-duplicated, sorted by p4est_quadrant_compare_piggy:
//                    tree, then Morton order)
//   cell search     find_neighbor_cell_of_node src/my_p4est_hierarchy.cpp:569-601
//   stencils        construct_neighbors        src/my_p4est_node_neighbors.cpp:114-1201
//   sampling        sample_cf_on_nodes, node_x_fr_n
//                                             src/my_p4est_utils.cpp:2129-2170,
//                                             src/my_p4est_utils.h:1346-1368
//
// Design: leaves are produced depth-first in child (Morton) order, tree by
// tree, so the leaf array is a sorted linear forest.  Positions across trees
// are brick coordinates (tree position * ROOT_LEN + logical coordinate), so
// neighbours across tree faces and periodic boundaries need no face
// transforms: a brick has none.  The 2^D cells around a node are found by
// scattering every leaf to its corner nodes (the cell that holds the corner
// in the direction pointing into the leaf); only slots where the node hangs
// on a coarser cell fall back to a Morton point location (binary search over
// the leaf start keys of the tree that holds the point), which is the
// hierarchy descent of the reference for a complete linear tree.
#include "../../include/casl_reinit.h"
#include "common.hpp"

#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <chrono>
#include <cstdio>
#include <stdexcept>

namespace casl {

// Phase timing on stderr when CASL_VERBOSE is set.
struct PhaseTimer {
  bool on = std::getenv("CASL_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  void mark(const char *what) {
    if (!on) return;
    auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[casl] %-28s %8.3f s\n", what, std::chrono::duration<double>(t1 - t0).count());
    t0 = t1;
  }
};

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *get_error() { return g_last_error.c_str(); }

int num_threads() {
  static int n = [] {
    const char *e = std::getenv("CASL_NUM_THREADS");
    if (!e) e = std::getenv("OMP_NUM_THREADS");
    int v = e ? std::atoi(e) : 0;
    if (v <= 0) v = int(std::thread::hardware_concurrency());
    if (v <= 0) v = 1;
    return std::min(v, 64);
  }();
  return n;
}

// ------------------------------------------------------------ level sets
static double eval_levelset(const casl_levelset &ls, int dim, const double *x) {
  const double *p = ls.params;
  switch (ls.kind) {
    case CASL_LS_CALLBACK:
      return ls.fn(x, ls.ctx);
    case CASL_LS_STAR: {  // 2D star, geom::Star-like r(theta) = r0 + A cos(p theta)
      double dx = x[0] - p[0], dy = x[1] - p[1];
      double r2 = dx * dx + dy * dy;
      double th = std::atan2(dy, dx);
      double r = p[2] + p[3] * std::cos(p[4] * th);
      return (std::sqrt(r2) - r) * (1.0 + p[5] * r2);
    }
    case CASL_LS_SPHERES: {  // 3D union of spheres, modulated
      double m = p[0];
      double best = HUGE_VAL;
      for (int i = 0; i < ls.n_items; ++i) {
        const double *s = p + 1 + 4 * i;
        double dx = x[0] - s[0], dy = x[1] - s[1], dz = x[2] - s[2];
        double v = std::sqrt(dx * dx + dy * dy + dz * dz) - s[3];
        if (v < best) best = v;
      }
      return best * (1.0 + m * std::sin(2.0 * M_PI * x[0]) * std::cos(2.0 * M_PI * x[1]));
    }
    case CASL_LS_GYROID: {
      double w = p[0];
      double g = std::sin(w * x[0]) * std::cos(w * x[1]) + std::sin(w * x[1]) * std::cos(w * x[2]) +
                 std::sin(w * x[2]) * std::cos(w * x[0]) - p[1];
      return g * (1.0 + p[2] * x[0] * x[1] * x[2]);
    }
    case CASL_LS_SPHERE: {
      double r2 = 0.0;
      for (int a = 0; a < dim; ++a) r2 += (x[a] - p[a]) * (x[a] - p[a]);
      return std::sqrt(r2) - p[dim];
    }
    case CASL_LS_PLANE: {
      double v = 0.0;
      for (int a = 0; a < dim; ++a) v += p[a] * x[a];
      return v + p[dim];
    }
    case CASL_LS_MODULATED: {  // examples/reinitialization/main_2d.cpp:112-116 (2D)
      double dx = x[0] - p[0], dy = x[1] - p[1];
      double e = std::sqrt(dx * dx + dy * dy) - p[2];
      return e * (1.0 + p[3] * std::sin(5.0 * x[0]) * std::cos(4.0 * x[1]));
    }
  }
  throw std::invalid_argument("unknown level-set kind");
}

static void check_levelset(const casl_levelset *ls, int dim) {
  if (!ls) throw std::invalid_argument("null level set");
  switch (ls->kind) {
    case CASL_LS_CALLBACK:
      if (!ls->fn) throw std::invalid_argument("callback level set without fn");
      return;
    case CASL_LS_STAR:
    case CASL_LS_MODULATED:
      if (dim != 2) throw std::invalid_argument("star/modulated level sets are 2D");
      break;
    case CASL_LS_SPHERES:
      if (dim != 3 || ls->n_items < 1) throw std::invalid_argument("spheres level set is 3D with n_items >= 1");
      break;
    case CASL_LS_GYROID:
      if (dim != 3) throw std::invalid_argument("gyroid level set is 3D");
      break;
    case CASL_LS_SPHERE:
    case CASL_LS_PLANE:
      break;
    default:
      throw std::invalid_argument("unknown level-set kind");
  }
  if (!ls->params) throw std::invalid_argument("level set without params");
}

// ------------------------------------------------------------ the forest
struct Quad {
  int32_t c[3];  // logical coordinates in its tree
  int32_t level;
  int32_t tree;
};

// A brick of nt[0] x nt[1] (x nt[2]) trees over [bmin, bmax], periodic or
// not per axis (my_p4est_brick_new, src/my_p4est_tools.c:37-170).  Every
// tree spans one (bmax - bmin) / nt cell of the brick; its vertex
// coordinates are bmin + i * dx and bmin + (i + 1) * dx, and its extent is
// their difference, as the connectivity's vertices give them.
template <int D>
struct ForestT {
  static constexpr int MAXLEVEL = (D == 2 ? 30 : 19);  // P4EST_MAXLEVEL / P8EST_MAXLEVEL
  static constexpr int32_t ROOT_LEN = int32_t(1) << MAXLEVEL;
  static constexpr int CH = 1 << D;        // P4EST_CHILDREN
  static constexpr int F = 2 * D;          // P4EST_FACES
  static constexpr int C = 1 << (D - 1);   // corners per face
  static constexpr int T = 2 * (D - 1);    // transverse distances per face
  static int32_t qlen(int level) { return int32_t(1) << (MAXLEVEL - level); }

  double bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};
  int nt[3] = {1, 1, 1};
  bool per[3] = {false, false, false};
  // per tree (in tree order): brick position, vertex min and extent
  std::vector<std::array<int32_t, 3>> tree_pos;
  std::vector<std::array<double, 3>> tmin, text;
  std::vector<int32_t> tree_at;  // brick position (i + nt0 * (j + nt1 * k)) -> tree
  double dom_min[3] = {0, 0, 0}, dom_max[3] = {1, 1, 1};  // domain corners from the vertices

  std::vector<Quad> leaves;
  std::vector<int64_t> tree_first;  // leaves of tree t: [tree_first[t], tree_first[t + 1])
  bool nodes_built = false, ngbd_built = false;
  int64_t n_nodes = 0;
  std::vector<int32_t> node_tree;    // [N] tree of the node's canonical position
  std::vector<int32_t> node_ixyz;    // [N][D] unclamped logical coordinates in that tree
  std::vector<int32_t> local_nodes;  // [L][CH]
  std::vector<int32_t> nbr;          // [F*C][N]
  std::vector<double> dface;         // [F][N]
  std::vector<double> dtr;           // [F*T][N]
  std::vector<uint8_t> wall;         // [N]

  int n_trees() const { return int(tree_pos.size()); }
  // logical length of the brick along axis a
  int64_t span(int a) const { return int64_t(nt[a]) * ROOT_LEN; }

  // Tree order: the brick positions along a Morton curve over the smallest
  // power-of-two box per axis that holds the brick, positions outside the
  // brick skipped (each axis contributes one bit per level while it has
  // bits, x first).  The order grows with every coordinate, so the tree of
  // smallest index that holds a point is the one of smallest position.
  void setup(const double *lo, const double *hi, const int *n, const int *periodic) {
    int lg[3] = {0, 0, 0};
    int64_t total = 1;
    for (int a = 0; a < D; ++a) {
      bmin[a] = lo[a];
      bmax[a] = hi[a];
      nt[a] = n[a];
      per[a] = periodic[a] != 0;
      while ((1 << lg[a]) < nt[a]) ++lg[a];
      total *= nt[a];
    }
    tree_pos.clear();
    tree_at.assign(size_t(total), -1);
    const int bits = lg[0] + lg[1] + (D == 3 ? lg[2] : 0);
    for (uint64_t m = 0; m < (uint64_t(1) << bits); ++m) {
      std::array<int32_t, 3> p = {0, 0, 0};
      int bit = 0;
      for (int b = 0; b < 16; ++b)
        for (int a = 0; a < D; ++a)
          if (b < lg[a]) p[a] |= int32_t((m >> bit++) & 1u) << b;
      bool in = true;
      for (int a = 0; a < D; ++a) in = in && p[a] < nt[a];
      if (!in) continue;
      tree_at[size_t(p[0] + int64_t(nt[0]) * (p[1] + int64_t(nt[1]) * p[2]))] = int32_t(tree_pos.size());
      tree_pos.push_back(p);
    }
    double dx[3];
    for (int a = 0; a < D; ++a) dx[a] = (bmax[a] - bmin[a]) / double(nt[a]);
    tmin.assign(tree_pos.size(), {0, 0, 0});
    text.assign(tree_pos.size(), {1, 1, 1});
    for (size_t t = 0; t < tree_pos.size(); ++t)
      for (int a = 0; a < D; ++a) {
        const double v0 = bmin[a] + tree_pos[t][a] * dx[a], v1 = bmin[a] + (tree_pos[t][a] + 1) * dx[a];
        tmin[t][a] = v0;
        text[t][a] = v1 - v0;
      }
    for (int a = 0; a < D; ++a) {
      dom_min[a] = bmin[a] + 0 * dx[a];
      dom_max[a] = bmin[a] + nt[a] * dx[a];
    }
    leaves.clear();
    for (size_t t = 0; t < tree_pos.size(); ++t) {
      Quad root{};
      root.level = 0;
      root.tree = int32_t(t);
      leaves.push_back(root);
    }
    index_trees();
    nodes_built = ngbd_built = false;
  }

  void index_trees() {
    tree_first.assign(tree_pos.size() + 1, 0);
    for (auto &q : leaves) ++tree_first[size_t(q.tree) + 1];
    for (size_t t = 0; t < tree_pos.size(); ++t) tree_first[t + 1] += tree_first[t];
  }

  // ---- refine_levelset_cf (refine_only_inside = false)
  bool decide(const Quad &q, int min_lvl, int max_lvl, double lip, double band_factor,
              const casl_levelset &ls) const {
    if (q.level < min_lvl) return true;
    if (q.level >= max_lvl) return false;
    const auto &e = text[q.tree];
    const auto &m = tmin[q.tree];
    double dmin = (double)qlen(q.level) / (double)ROOT_LEN;
    double dx[3], x[3];
    for (int a = 0; a < D; ++a) dx[a] = e[a] * dmin;
    double smallest = std::max(e[0], e[1]) * ((double)qlen(max_lvl)) / ((double)ROOT_LEN);
    if (D == 3) smallest = std::max(smallest, e[2] * ((double)qlen(max_lvl)) / ((double)ROOT_LEN));
    double d2 = dx[0] * dx[0] + dx[1] * dx[1];
    if (D == 3) d2 = d2 + dx[2] * dx[2];
    double d = std::sqrt(d2);
    for (int a = 0; a < D; ++a) x[a] = e[a] * (double)q.c[a] / (double)ROOT_LEN + m[a];
    double band = band_factor * smallest;
    double f[8];
    for (int k = 0; k < CH; ++k) {
      double xc[3];
      for (int a = 0; a < D; ++a) xc[a] = x[a] + ((k >> a) & 1) * dx[a];
      f[k] = eval_levelset(ls, D, xc);
      if (std::fabs(f[k]) - band <= 0.5 * lip * d) return true;
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

  static Quad child(const Quad &q, int k) {
    Quad c = q;
    c.level = q.level + 1;
    int32_t h = qlen(c.level);
    for (int a = 0; a < D; ++a) c.c[a] = q.c[a] + ((k >> a) & 1) * h;
    return c;
  }

  void dfs(const Quad &q, int min_lvl, int max_lvl, double lip, double band, const casl_levelset &ls,
           std::vector<Quad> &out) const {
    if (q.level < MAXLEVEL - 1 && decide(q, min_lvl, max_lvl, lip, band, ls)) {
      for (int k = 0; k < CH; ++k) dfs(child(q, k), min_lvl, max_lvl, lip, band, ls, out);
    } else {
      out.push_back(q);
    }
  }

  void refine(int min_lvl, int max_lvl, double lip, double band, const casl_levelset &ls) {
    if (max_lvl >= MAXLEVEL - 1) throw std::invalid_argument("max_lvl too large for this dimension");
    struct Item {
      Quad q;
      bool final_leaf;
    };
    std::vector<Item> items;
    items.reserve(leaves.size());
    for (auto &q : leaves) items.push_back({q, false});
    const size_t want = size_t(16) * num_threads();
    for (int round = 0; round < 8; ++round) {
      size_t open = 0;
      for (auto &it : items) open += !it.final_leaf;
      if (open == 0 || open >= want) break;
      std::vector<Item> next;
      next.reserve(items.size() * 2);
      for (auto &it : items) {
        if (it.final_leaf) {
          next.push_back(it);
        } else if (it.q.level < MAXLEVEL - 1 && decide(it.q, min_lvl, max_lvl, lip, band, ls)) {
          for (int k = 0; k < CH; ++k) next.push_back({child(it.q, k), false});
        } else {
          next.push_back({it.q, true});
        }
      }
      items.swap(next);
    }
    std::vector<std::vector<Quad>> parts(items.size());
    std::atomic<size_t> cursor{0};
    int nt = num_threads();
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> err(size_t(nt), nullptr);
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&, t]() {
        try {
          for (;;) {
            size_t i = cursor.fetch_add(1);
            if (i >= items.size()) break;
            if (items[i].final_leaf)
              parts[i].push_back(items[i].q);
            else
              dfs(items[i].q, min_lvl, max_lvl, lip, band, ls, parts[i]);
          }
        } catch (...) {
          err[size_t(t)] = std::current_exception();
          cursor.store(items.size());
        }
      });
    for (auto &th : pool) th.join();
    for (auto &x : err)
      if (x) std::rethrow_exception(x);
    size_t total = 0;
    for (auto &p : parts) total += p.size();
    std::vector<Quad> out;
    out.reserve(total);
    for (auto &p : parts) out.insert(out.end(), p.begin(), p.end());
    leaves.swap(out);
    index_trees();
    nodes_built = ngbd_built = false;
  }


  // ---- my_p4est_nodes_new on one rank
  // Brick coordinate of a point: tree position * ROOT_LEN + logical
  // coordinate.  A node's canonical owner (p4est_node_canonicalize: the tree
  // of smallest index that holds it) is the tree of smallest position, so a
  // node on a face between trees belongs to the lower one, at logical
  // coordinate ROOT_LEN there, and a node on a periodic boundary to the
  // first tree along that axis, at 0.
  void canonical(const int64_t *g, int32_t &tree, int32_t *loc) const {
    int64_t p[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) {
      int64_t v = g[a];
      if (per[a]) v = ((v % span(a)) + span(a)) % span(a);
      p[a] = v == 0 ? 0 : (v - 1) / ROOT_LEN;
      loc[a] = int32_t(v - p[a] * ROOT_LEN);
    }
    tree = tree_at[size_t(p[0] + nt[0] * (p[1] + int64_t(nt[1]) * p[2]))];
  }
  // node keys: (tree, Morton key of the clamped coordinates), the order of
  // p4est_quadrant_compare_piggy; one tree fits a 64-bit key
  template <class K>
  static K node_key(int32_t tree, const int32_t *loc) {
    int32_t p[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) p[a] = loc[a] == ROOT_LEN ? ROOT_LEN - 1 : loc[a];  // p4est_node_clamp_inside
    const uint64_t m = morton<D>(p);
    if constexpr (sizeof(K) == 8)
      return K(m);
    else
      return (K(uint32_t(tree)) << 64) | K(m);
  }
  template <class K>
  static void key_split(K k, int32_t &tree, uint64_t &m) {
    m = uint64_t(k);
    if constexpr (sizeof(K) == 8)
      tree = 0;
    else
      tree = int32_t(uint64_t(k >> 64));
  }

  void build_nodes() {
    if (n_trees() == 1)
      build_nodes_k<uint64_t>();
    else
      build_nodes_k<unsigned __int128>();
  }

  template <class K>
  void build_nodes_k() {
    const int64_t L = int64_t(leaves.size());
    std::vector<K> keys(size_t(L) * CH);
    parallel_for(L, [&](int64_t b, int64_t e) {
      for (int64_t l = b; l < e; ++l) {
        const Quad &q = leaves[l];
        int32_t h = qlen(q.level);
        for (int k = 0; k < CH; ++k) {
          int64_t g[3] = {0, 0, 0};
          for (int a = 0; a < D; ++a)
            g[a] = int64_t(tree_pos[q.tree][a]) * ROOT_LEN + q.c[a] + ((k >> a) & 1) * h;
          int32_t tree, loc[3] = {0, 0, 0};
          canonical(g, tree, loc);
          keys[size_t(l) * CH + k] = node_key<K>(tree, loc);
        }
      }
    });
    // sort + unique, chunk-parallel then pairwise merges
    std::vector<K> sorted(keys);
    {
      int nth = std::max(1, std::min(num_threads(), int(sorted.size() / 65536) + 1));
      std::vector<size_t> bounds(nth + 1);
      for (int t = 0; t <= nth; ++t) bounds[t] = sorted.size() * size_t(t) / size_t(nth);
      {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t)
          pool.emplace_back([&, t]() {
            std::sort(sorted.begin() + bounds[t], sorted.begin() + bounds[t + 1]);
            auto end = std::unique(sorted.begin() + bounds[t], sorted.begin() + bounds[t + 1]);
            // keep chunk length; pad duplicates with the chunk's last unique key
            if (end != sorted.begin() + bounds[t + 1] && end != sorted.begin() + bounds[t]) {
              K last = *(end - 1);
              std::fill(end, sorted.begin() + bounds[t + 1], last);
            }
          });
        for (auto &th : pool) th.join();
      }
      std::vector<K> tmp(sorted.size());
      for (size_t width = 1; width < size_t(nth); width *= 2) {
        std::vector<std::thread> pool;
        for (size_t t = 0; t < size_t(nth); t += 2 * width) {
          size_t lo = bounds[t], mid = bounds[std::min(size_t(nth), t + width)],
                 hi = bounds[std::min(size_t(nth), t + 2 * width)];
          pool.emplace_back([&, lo, mid, hi]() {
            std::merge(sorted.begin() + lo, sorted.begin() + mid, sorted.begin() + mid, sorted.begin() + hi,
                       tmp.begin() + lo);
          });
        }
        for (auto &th : pool) th.join();
        sorted.swap(tmp);
      }
      sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
    }
    const int64_t N = int64_t(sorted.size());
    if (N >= int64_t(INT32_MAX)) throw std::runtime_error("too many nodes for int32 indices");
    local_nodes.assign(size_t(L) * CH, -1);
    parallel_for(L * CH, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        auto it = std::lower_bound(sorted.begin(), sorted.end(), keys[size_t(i)]);
        local_nodes[size_t(i)] = int32_t(it - sorted.begin());
      }
    });
    node_ixyz.assign(size_t(N) * D, 0);
    node_tree.assign(size_t(N), 0);
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n) {
        int32_t c[3], tree;
        uint64_t m;
        key_split(sorted[size_t(n)], tree, m);
        demorton<D>(m, c);
        node_tree[size_t(n)] = tree;
        for (int a = 0; a < D; ++a) node_ixyz[size_t(n) * D + a] = (c[a] == ROOT_LEN - 1 ? ROOT_LEN : c[a]);
      }
    });
    n_nodes = N;
    nodes_built = true;
    ngbd_built = false;
  }

  // construct_neighbors :131-152 (the unclamped node in its canonical tree)
  double node_coord(int64_t n, int a) const {
    const int32_t t = node_tree[size_t(n)];
    return text[t][a] * (node_ixyz[size_t(n) * D + a] / (double)ROOT_LEN) + tmin[t][a];
  }
  int64_t node_pos(int64_t n, int a) const {
    return int64_t(tree_pos[node_tree[size_t(n)]][a]) * ROOT_LEN + node_ixyz[size_t(n) * D + a];
  }

  // ---- construct_neighbors for every node
  void init_neighbors() {
    if (!nodes_built) throw std::logic_error("nodes not built");
    PhaseTimer timer;
    const int64_t L = int64_t(leaves.size());
    const int64_t N = n_nodes;
    // cells around each node, indexed by direction bits (bit a set <=> +a side)
    std::vector<int32_t> cell(size_t(N) * CH, -2);
    parallel_for(L, [&](int64_t b, int64_t e) {
      for (int64_t l = b; l < e; ++l)
        for (int k = 0; k < CH; ++k) cell[size_t(local_nodes[size_t(l) * CH + k]) * CH + (k ^ (CH - 1))] = int32_t(l);
    });
    timer.mark("scatter cells");
    std::vector<uint64_t> leaf_keys(L);
    parallel_for(L, [&](int64_t b, int64_t e) {
      for (int64_t l = b; l < e; ++l) leaf_keys[l] = morton<D>(leaves[l].c);
    });
    // find_neighbor_cell_of_node: the node moved by one logical unit per axis
    // (is_node_in_domain: off a non-periodic boundary there is no cell,
    // across a periodic one the point wraps), then the leaf of that point's
    // tree that holds it
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n) {
        for (int dir = 0; dir < CH; ++dir) {
          int32_t &slot = cell[size_t(n) * CH + dir];
          if (slot != -2) continue;
          int64_t p[3] = {0, 0, 0};
          bool inside = true;
          for (int a = 0; a < D; ++a) {
            p[a] = node_pos(n, a) + (((dir >> a) & 1) ? 1 : -1);
            if (per[a])
              p[a] = ((p[a] % span(a)) + span(a)) % span(a);
            else if (p[a] < 0 || p[a] > span(a))
              inside = false;
          }
          if (!inside) {
            slot = -1;  // NOT_A_VALID_QUADRANT -> "root" (level -1)
            continue;
          }
          // p is odd on every axis (nodes sit on even coordinates): strictly inside one tree
          int64_t tp[3] = {0, 0, 0};
          int32_t loc[3] = {0, 0, 0};
          for (int a = 0; a < D; ++a) {
            tp[a] = p[a] / ROOT_LEN;
            loc[a] = int32_t(p[a] - tp[a] * ROOT_LEN);
          }
          const int32_t t = tree_at[size_t(tp[0] + nt[0] * (tp[1] + int64_t(nt[1]) * tp[2]))];
          const uint64_t key = morton<D>(loc);
          auto first = leaf_keys.begin() + tree_first[t], last = leaf_keys.begin() + tree_first[t + 1];
          auto it = std::upper_bound(first, last, key);
          slot = int32_t((it - leaf_keys.begin()) - 1);
        }
      }
    });

    timer.mark("point location");
    nbr.assign(size_t(F) * C * N, -1);
    dface.assign(size_t(F) * N, 0.0);
    dtr.assign(size_t(F) * T * N, 0.0);
    wall.assign(size_t(N), 0);
    timer.mark("allocate tables");
    std::atomic<int> bad{0};
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n)
        if (!construct(n, cell)) bad.store(1);
    });
    timer.mark("construct stencils");
    if (bad.load()) throw std::runtime_error("construct_neighbors: a wall arm has no second-degree cell");
    ngbd_built = true;
  }

  static void transverse_axes(int a, int &t1, int &t2) {
    if (D == 2) {
      t1 = 1 - a;
      t2 = -1;
    } else {
      t1 = (a == 0 ? 1 : 0);
      t2 = (a == 2 ? 1 : 2);
    }
  }

  // Fills nbr/dtr for face f from leaf qi, whose face with logical bit
  // `side` along axis a carries the neighbor values.  The transverse
  // offsets use the leaf's tree; on a periodic axis a leaf that wrapped
  // around from the far end (the node lies more than a quarter cell below
  // it) is moved down by the domain length (construct_neighbors :345-357).
  void fill_from_cell(int64_t n, int f, int32_t qi, int side) {
    const int64_t N = n_nodes;
    const int a = f >> 1;
    int t1, t2;
    transverse_axes(a, t1, t2);
    const Quad &Q = leaves[qi];
    for (int c = 0; c < C; ++c) {
      int corner = (side << a) | ((c & 1) << t1);
      if (D == 3) corner |= ((c >> 1) & 1) << t2;
      nbr[(size_t(f) * C + c) * N + n] = local_nodes[size_t(qi) * CH + corner];
    }
    const double qh = qlen(Q.level) / (double)ROOT_LEN;
    for (int t = 0; t < D - 1; ++t) {
      const int ax = (t == 0 ? t1 : t2);
      const double ext = text[Q.tree][ax];
      const double thr = ext * ((double)1) / ((double)ROOT_LEN);  // QUADRANT_LEN(P4EST_MAXLEVEL) == 1
      const double X = node_coord(n, ax);
      const double QX = ext * (Q.c[ax] / (double)ROOT_LEN) + tmin[Q.tree][ax];
      double m0 = (std::fabs(X - QX) > thr) ? (X - QX) : 0.0;
      if (per[ax] && X < QX - ext * qh / 4.0) {
        const double w = m0 + (dom_max[ax] - dom_min[ax]);
        m0 = (std::fabs(w) > thr) ? w : 0.0;
      }
      const double p0 = (std::fabs(ext * qh - m0) > thr) ? (ext * qh - m0) : 0.0;
      dtr[(size_t(f) * T + 2 * t + 0) * N + n] = m0;
      dtr[(size_t(f) * T + 2 * t + 1) * N + n] = p0;
    }
  }

  bool construct(int64_t n, const std::vector<int32_t> &cell) {
    const int64_t N = n_nodes;
    const int32_t *cn = &cell[size_t(n) * CH];
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
      dface[size_t(f) * N + n] = text[leaves[qi].tree][a] * (qlen(leaves[qi].level) / (double)ROOT_LEN);
      fill_from_cell(n, f, qi, s == 0 ? 0 : 1);
    }
    // walls: second-degree neighbor on the opposite side, negative distance
    for (int f = 0; f < F; ++f) {
      if (!is_wall[f]) continue;
      const int a = f >> 1, s = f & 1, so = 1 - s;
      int dmin = -1;
      int lbest = -1;  // root level
      for (int dir = 0; dir < CH; ++dir) {
        if (((dir >> a) & 1) != so) continue;
        if (lbest < lvl(dir)) {
          lbest = lvl(dir);
          dmin = dir;
        }
      }
      if (dmin < 0) return false;
      const int32_t qmin = cn[dmin];
      int corner = so << a;
      for (int b = 0; b < D; ++b)
        if (b != a) corner |= (((dmin >> b) & 1) ? 0 : 1) << b;  // d_b = (c_b != 1)
      const int32_t ntmp = local_nodes[size_t(qmin) * CH + corner];
      const int32_t qtmp = cell[size_t(ntmp) * CH + dmin];
      if (qtmp < 0) return false;
      const int fo = f ^ 1;
      dface[size_t(f) * N + n] =
          -dface[size_t(fo) * N + n] - text[leaves[qtmp].tree][a] * (qlen(leaves[qtmp].level) / (double)ROOT_LEN);
      fill_from_cell(n, f, qtmp, so);
    }
    // is_node_*Wall: on a domain face of a non-periodic axis
    uint8_t w = 0;
    for (int a = 0; a < D; ++a) {
      if (per[a]) continue;
      const int64_t x = node_pos(n, a);
      if (x == 0) w |= uint8_t(1u << (2 * a));
      if (x == span(a)) w |= uint8_t(1u << (2 * a + 1));
    }
    wall[n] = w;
    return true;
  }

  void sample(const casl_levelset &ls, double *out) const {
    parallel_for(n_nodes, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n) {
        double x[3];
        // node_x_fr_n: the unclamped coordinate in the node's tree
        for (int a = 0; a < D; ++a) x[a] = node_coord(n, a);
        out[n] = eval_levelset(ls, D, x);
      }
    }, 1024);
  }

  int max_level() const {
    int m = 0;
    for (auto &q : leaves) m = std::max(m, int(q.level));
    return m;
  }
};

}  // namespace casl

struct casl_forest {
  int dim;
  std::unique_ptr<casl::ForestT<2>> f2;
  std::unique_ptr<casl::ForestT<3>> f3;
};

using namespace casl;

template <class Fn>
static int guarded(Fn fn) {
  try {
    fn();
    return CASL_OK;
  } catch (const std::invalid_argument &e) {
    set_error(e.what());
    return CASL_ERR_INVALID_ARGUMENT;
  } catch (const std::logic_error &e) {
    set_error(e.what());
    return CASL_ERR_NOT_BUILT;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return CASL_ERR_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    set_error(e.what());
    return CASL_ERR_INVALID_STENCIL;
  }
}

template <int D>
static void get_leaves_t(const ForestT<D> &F, int32_t *xyz, int8_t *level, int32_t *local) {
  for (size_t l = 0; l < F.leaves.size(); ++l) {
    if (xyz)
      for (int a = 0; a < D; ++a) xyz[l * D + a] = F.leaves[l].c[a];
    if (level) level[l] = int8_t(F.leaves[l].level);
  }
  if (local) {
    if (!F.nodes_built) throw std::logic_error("nodes not built");
    std::memcpy(local, F.local_nodes.data(), F.local_nodes.size() * sizeof(int32_t));
  }
}

template <int D>
static void get_nodes_t(const ForestT<D> &F, int32_t *ixyz, double *xyz) {
  if (!F.nodes_built) throw std::logic_error("nodes not built");
  const size_t N = size_t(F.n_nodes);
  if (ixyz) std::memcpy(ixyz, F.node_ixyz.data(), N * D * sizeof(int32_t));
  if (xyz)
    for (size_t n = 0; n < N; ++n)
      for (int a = 0; a < D; ++a) xyz[n * D + a] = F.node_coord(int64_t(n), a);
}

template <int D>
static void init_neighbors_t(ForestT<D> &F, casl_flat_stencil *out) {
  if (!F.ngbd_built) F.init_neighbors();
  out->dim = D;
  out->n_nodes = F.n_nodes;
  out->nbr = F.nbr.data();
  out->d = F.dface.data();
  out->d_tr = F.dtr.data();
  out->wall = F.wall.data();
  double m = F.text[0][0];
  for (int a = 1; a < D; ++a) m = std::min(m, F.text[0][a]);
  out->zero_distance_threshold = 1e-13 * m;  // EPS*MIN(tree dimensions), the trees of a brick being equal
}

extern "C" {

const char *casl_last_error(void) { return get_error(); }

int casl_levelset_eval(const casl_levelset *ls, int dim, int64_t n, const double *xyz, double *out) {
  return guarded([&]() {
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    check_levelset(ls, dim);
    if (n < 0 || (n > 0 && (!xyz || !out))) throw std::invalid_argument("bad arrays");
    for (int64_t i = 0; i < n; ++i) out[i] = eval_levelset(*ls, dim, xyz + i * dim);
  });
}

int casl_forest_new_brick(int dim, const double *box_min, const double *box_max, const int32_t *n_trees,
                          const int32_t *periodic, casl_forest_t **out) {
  return guarded([&]() {
    if (!out) throw std::invalid_argument("null out");
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    int n[3] = {1, 1, 1}, per[3] = {0, 0, 0};
    int64_t total = 1;
    for (int a = 0; a < dim; ++a) {
      lo[a] = box_min ? box_min[a] : 0.0;
      hi[a] = box_max ? box_max[a] : 1.0;
      if (!(hi[a] > lo[a])) throw std::invalid_argument("empty box");
      n[a] = n_trees ? n_trees[a] : 1;
      if (n[a] < 1 || n[a] > 1024) throw std::invalid_argument("trees per axis must be in [1, 1024]");
      per[a] = periodic ? (periodic[a] != 0) : 0;
      total *= n[a];
    }
    if (total > (int64_t(1) << 20)) throw std::invalid_argument("more than 2^20 trees");
    auto f = std::make_unique<casl_forest>();
    f->dim = dim;
    if (dim == 2) {
      f->f2 = std::make_unique<ForestT<2>>();
      f->f2->setup(lo, hi, n, per);
    } else {
      f->f3 = std::make_unique<ForestT<3>>();
      f->f3->setup(lo, hi, n, per);
    }
    *out = f.release();
  });
}

int casl_forest_new(int dim, const double *box_min, const double *box_max, casl_forest_t **out) {
  return casl_forest_new_brick(dim, box_min, box_max, nullptr, nullptr, out);
}

int casl_forest_trees(const casl_forest_t *f, int32_t *n_trees, int32_t *tree_pos, int32_t *leaf_tree,
                      int32_t *node_tree) {
  return guarded([&]() {
    if (!f) throw std::invalid_argument("null forest");
    auto get = [&](const auto &F) {
      if (n_trees) *n_trees = F.n_trees();
      if (tree_pos)
        for (int t = 0; t < F.n_trees(); ++t)
          for (int a = 0; a < 3; ++a) tree_pos[3 * t + a] = F.tree_pos[size_t(t)][a];
      if (leaf_tree)
        for (size_t l = 0; l < F.leaves.size(); ++l) leaf_tree[l] = F.leaves[l].tree;
      if (node_tree) {
        if (!F.nodes_built) throw std::logic_error("nodes not built");
        std::memcpy(node_tree, F.node_tree.data(), F.node_tree.size() * sizeof(int32_t));
      }
    };
    if (f->dim == 2)
      get(*f->f2);
    else
      get(*f->f3);
  });
}

int casl_forest_destroy(casl_forest_t *f) {
  delete f;
  return CASL_OK;
}

int casl_forest_refine_levelset(casl_forest_t *f, int min_lvl, int max_lvl, double lip, double band,
                                const casl_levelset *phi) {
  return guarded([&]() {
    if (!f) throw std::invalid_argument("null forest");
    if (min_lvl < 0 || max_lvl < 0) throw std::invalid_argument("negative levels");
    check_levelset(phi, f->dim);
    if (f->dim == 2)
      f->f2->refine(min_lvl, max_lvl, lip, band, *phi);
    else
      f->f3->refine(min_lvl, max_lvl, lip, band, *phi);
  });
}

int casl_forest_build_nodes(casl_forest_t *f) {
  return guarded([&]() {
    if (!f) throw std::invalid_argument("null forest");
    if (f->dim == 2)
      f->f2->build_nodes();
    else
      f->f3->build_nodes();
  });
}

int casl_forest_sizes(const casl_forest_t *f, int64_t *n_leaves, int64_t *n_nodes, int32_t *max_level) {
  return guarded([&]() {
    if (!f) throw std::invalid_argument("null forest");
    if (f->dim == 2) {
      if (n_leaves) *n_leaves = int64_t(f->f2->leaves.size());
      if (n_nodes) *n_nodes = f->f2->nodes_built ? f->f2->n_nodes : -1;
      if (max_level) *max_level = f->f2->max_level();
    } else {
      if (n_leaves) *n_leaves = int64_t(f->f3->leaves.size());
      if (n_nodes) *n_nodes = f->f3->nodes_built ? f->f3->n_nodes : -1;
      if (max_level) *max_level = f->f3->max_level();
    }
  });
}


int casl_forest_get_leaves(const casl_forest_t *f, int32_t *xyz, int8_t *level, int32_t *local_nodes) {
  return guarded([&]() {
    if (!f) throw std::invalid_argument("null forest");
    if (f->dim == 2)
      get_leaves_t(*f->f2, xyz, level, local_nodes);
    else
      get_leaves_t(*f->f3, xyz, level, local_nodes);
  });
}


int casl_forest_get_nodes(const casl_forest_t *f, int32_t *ixyz, double *xyz) {
  return guarded([&]() {
    if (!f) throw std::invalid_argument("null forest");
    if (f->dim == 2)
      get_nodes_t(*f->f2, ixyz, xyz);
    else
      get_nodes_t(*f->f3, ixyz, xyz);
  });
}

int casl_forest_sample(const casl_forest_t *f, const casl_levelset *phi, double *out) {
  return guarded([&]() {
    if (!f || !out) throw std::invalid_argument("null argument");
    check_levelset(phi, f->dim);
    if (f->dim == 2) {
      if (!f->f2->nodes_built) throw std::logic_error("nodes not built");
      f->f2->sample(*phi, out);
    } else {
      if (!f->f3->nodes_built) throw std::logic_error("nodes not built");
      f->f3->sample(*phi, out);
    }
  });
}


int casl_forest_init_neighbors(casl_forest_t *f, casl_flat_stencil *out) {
  return guarded([&]() {
    if (!f || !out) throw std::invalid_argument("null argument");
    if (f->dim == 2)
      init_neighbors_t(*f->f2, out);
    else
      init_neighbors_t(*f->f3, out);
  });
}

}  // extern "C"
