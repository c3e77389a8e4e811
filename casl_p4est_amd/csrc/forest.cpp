// Host-side forest: the part of the grid that the reference takes from p4est
// and my_p4est_node_neighbors_t, for one non-periodic tree on a box.
//
//   refinement      refine_levelset_cf        src/my_p4est_refine_coarsen.cpp:17-89
//                   iterated max_lvl times    examples/reinitialization/main_2d.cpp:169-173
//   node numbering  my_p4est_nodes_new        src/my_p4est_nodes.c:383-446, 614-671
//                   (clamped corner nodes, de-duplicated, sorted by
//                    p4est_quadrant_compare_piggy = Morton order in one tree)
//   cell search     find_neighbor_cell_of_node src/my_p4est_hierarchy.cpp:569-601
//   stencils        construct_neighbors        src/my_p4est_node_neighbors.cpp:114-1201
//   sampling        sample_cf_on_nodes, node_x_fr_n
//                                             src/my_p4est_utils.cpp:2129-2170,
//                                             src/my_p4est_utils.h:1346-1368
//
// Design: leaves are produced depth-first in child (Morton) order, so the
// leaf array is a sorted linear tree.  The 2^D cells around a node are found
// by scattering every leaf to its corner nodes (the cell that holds the
// corner in the direction pointing into the leaf); only slots where the node
// hangs on a coarser cell fall back to a Morton point location (binary search
// over leaf start keys), which is the hierarchy descent of the reference for
// a complete linear tree.
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
  int32_t c[3];
  int32_t level;
};

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
  std::vector<Quad> leaves;
  bool nodes_built = false, ngbd_built = false;
  std::vector<uint64_t> node_keys;   // Morton keys of clamped coords, sorted
  std::vector<int32_t> node_ixyz;    // [N][D] unclamped logical coordinates
  std::vector<int32_t> local_nodes;  // [L][CH]
  std::vector<int32_t> nbr;          // [F*C][N]
  std::vector<double> dface;         // [F][N]
  std::vector<double> dtr;           // [F*T][N]
  std::vector<uint8_t> wall;         // [N]

  double size(int a) const { return bmax[a] - bmin[a]; }

  ForestT() {
    Quad root{};
    root.level = 0;
    leaves.push_back(root);
  }

  // ---- refine_levelset_cf (refine_only_inside = false)
  bool decide(const Quad &q, int min_lvl, int max_lvl, double lip, double band_factor,
              const casl_levelset &ls) const {
    if (q.level < min_lvl) return true;
    if (q.level >= max_lvl) return false;
    double dmin = (double)qlen(q.level) / (double)ROOT_LEN;
    double dx[3], x[3];
    for (int a = 0; a < D; ++a) dx[a] = size(a) * dmin;
    double smallest = std::max(size(0), size(1)) * ((double)qlen(max_lvl)) / ((double)ROOT_LEN);
    if (D == 3) smallest = std::max(smallest, size(2) * ((double)qlen(max_lvl)) / ((double)ROOT_LEN));
    double d2 = dx[0] * dx[0] + dx[1] * dx[1];
    if (D == 3) d2 = d2 + dx[2] * dx[2];
    double d = std::sqrt(d2);
    for (int a = 0; a < D; ++a) x[a] = size(a) * (double)q.c[a] / (double)ROOT_LEN + bmin[a];
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
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&]() {
        for (;;) {
          size_t i = cursor.fetch_add(1);
          if (i >= items.size()) break;
          if (items[i].final_leaf)
            parts[i].push_back(items[i].q);
          else
            dfs(items[i].q, min_lvl, max_lvl, lip, band, ls, parts[i]);
        }
      });
    for (auto &th : pool) th.join();
    size_t total = 0;
    for (auto &p : parts) total += p.size();
    std::vector<Quad> out;
    out.reserve(total);
    for (auto &p : parts) out.insert(out.end(), p.begin(), p.end());
    leaves.swap(out);
    nodes_built = ngbd_built = false;
  }

  // ---- my_p4est_nodes_new for one tree on one rank
  void build_nodes() {
    const int64_t L = int64_t(leaves.size());
    std::vector<uint64_t> keys(size_t(L) * CH);
    parallel_for(L, [&](int64_t b, int64_t e) {
      for (int64_t l = b; l < e; ++l) {
        const Quad &q = leaves[l];
        int32_t h = qlen(q.level);
        for (int k = 0; k < CH; ++k) {
          int32_t p[3];
          for (int a = 0; a < D; ++a) {
            p[a] = q.c[a] + ((k >> a) & 1) * h;
            if (p[a] == ROOT_LEN) p[a] = ROOT_LEN - 1;  // p4est_node_clamp_inside
          }
          keys[size_t(l) * CH + k] = morton<D>(p);
        }
      }
    });
    // sort + unique, chunk-parallel then pairwise merges
    std::vector<uint64_t> sorted(keys);
    {
      int nt = std::max(1, std::min(num_threads(), int(sorted.size() / 65536) + 1));
      std::vector<size_t> bounds(nt + 1);
      for (int t = 0; t <= nt; ++t) bounds[t] = sorted.size() * size_t(t) / size_t(nt);
      {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
          pool.emplace_back([&, t]() {
            std::sort(sorted.begin() + bounds[t], sorted.begin() + bounds[t + 1]);
            auto end = std::unique(sorted.begin() + bounds[t], sorted.begin() + bounds[t + 1]);
            // keep chunk length; pad duplicates with the chunk's last unique key
            if (end != sorted.begin() + bounds[t + 1] && end != sorted.begin() + bounds[t]) {
              uint64_t last = *(end - 1);
              std::fill(end, sorted.begin() + bounds[t + 1], last);
            }
          });
        for (auto &th : pool) th.join();
      }
      std::vector<uint64_t> tmp(sorted.size());
      for (size_t width = 1; width < size_t(nt); width *= 2) {
        std::vector<std::thread> pool;
        for (size_t t = 0; t < size_t(nt); t += 2 * width) {
          size_t lo = bounds[t], mid = bounds[std::min(size_t(nt), t + width)],
                 hi = bounds[std::min(size_t(nt), t + 2 * width)];
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
    node_keys.swap(sorted);
    const int64_t N = int64_t(node_keys.size());
    if (N >= int64_t(INT32_MAX)) throw std::runtime_error("too many nodes for int32 indices");
    local_nodes.assign(size_t(L) * CH, -1);
    parallel_for(L * CH, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        auto it = std::lower_bound(node_keys.begin(), node_keys.end(), keys[i]);
        local_nodes[i] = int32_t(it - node_keys.begin());
      }
    });
    node_ixyz.assign(size_t(N) * D, 0);
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n) {
        int32_t c[3];
        demorton<D>(node_keys[n], c);
        for (int a = 0; a < D; ++a) node_ixyz[size_t(n) * D + a] = (c[a] == ROOT_LEN - 1 ? ROOT_LEN : c[a]);
      }
    });
    nodes_built = true;
    ngbd_built = false;
  }

  double node_coord(int64_t n, int a) const {  // construct_neighbors :131-136 (unclamped node)
    return size(a) * (node_ixyz[size_t(n) * D + a] / (double)ROOT_LEN) + bmin[a];
  }

  // ---- construct_neighbors for every node
  void init_neighbors() {
    if (!nodes_built) throw std::logic_error("nodes not built");
    PhaseTimer timer;
    const int64_t L = int64_t(leaves.size());
    const int64_t N = int64_t(node_keys.size());
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
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n) {
        for (int dir = 0; dir < CH; ++dir) {
          int32_t &slot = cell[size_t(n) * CH + dir];
          if (slot != -2) continue;
          int32_t p[3];
          bool inside = true;
          for (int a = 0; a < D; ++a) {
            p[a] = node_ixyz[size_t(n) * D + a] + (((dir >> a) & 1) ? 1 : -1);
            if (p[a] < 0 || p[a] > ROOT_LEN) inside = false;  // is_node_in_domain, one tree
          }
          if (!inside) {
            slot = -1;  // NOT_A_VALID_QUADRANT -> "root" (level -1)
            continue;
          }
          uint64_t key = morton<D>(p);
          auto it = std::upper_bound(leaf_keys.begin(), leaf_keys.end(), key);
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
  // `side` along axis a carries the neighbor values.
  void fill_from_cell(int64_t n, int f, int32_t qi, int side) {
    const int64_t N = int64_t(node_keys.size());
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
      const double thr = size(ax) * ((double)1) / ((double)ROOT_LEN);  // QUADRANT_LEN(P4EST_MAXLEVEL) == 1
      const double X = node_coord(n, ax);
      const double QX = size(ax) * (Q.c[ax] / (double)ROOT_LEN) + bmin[ax];
      const double m0 = (std::fabs(X - QX) > thr) ? (X - QX) : 0.0;
      const double p0 = (std::fabs(size(ax) * qh - m0) > thr) ? (size(ax) * qh - m0) : 0.0;
      dtr[(size_t(f) * T + 2 * t + 0) * N + n] = m0;
      dtr[(size_t(f) * T + 2 * t + 1) * N + n] = p0;
    }
  }

  bool construct(int64_t n, const std::vector<int32_t> &cell) {
    const int64_t N = int64_t(node_keys.size());
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
      dface[size_t(f) * N + n] = size(a) * (qlen(leaves[qi].level) / (double)ROOT_LEN);
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
          -dface[size_t(fo) * N + n] - size(a) * (qlen(leaves[qtmp].level) / (double)ROOT_LEN);
      fill_from_cell(n, f, qtmp, so);
    }
    uint8_t w = 0;
    for (int a = 0; a < D; ++a) {
      int32_t x = node_ixyz[size_t(n) * D + a];  // clamped ROOT_LEN-1 reads as ROOT_LEN here
      if (x == 0) w |= uint8_t(1u << (2 * a));
      if (x == ROOT_LEN - 1 || x == ROOT_LEN) w |= uint8_t(1u << (2 * a + 1));
    }
    wall[n] = w;
    return true;
  }

  void sample(const casl_levelset &ls, double *out) const {
    const int64_t N = int64_t(node_keys.size());
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t n = b; n < e; ++n) {
        double x[3];
        for (int a = 0; a < D; ++a) {
          // node_x_fr_n: clamped ROOT_LEN-1 maps to 1.0, i.e. the unclamped coordinate
          x[a] = size(a) * (node_ixyz[size_t(n) * D + a] / (double)ROOT_LEN) + bmin[a];
        }
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
  const size_t N = F.node_keys.size();
  if (ixyz) std::memcpy(ixyz, F.node_ixyz.data(), N * D * sizeof(int32_t));
  if (xyz)
    for (size_t n = 0; n < N; ++n)
      for (int a = 0; a < D; ++a) xyz[n * D + a] = F.node_coord(int64_t(n), a);
}

template <int D>
static void init_neighbors_t(ForestT<D> &F, casl_flat_stencil *out) {
  if (!F.ngbd_built) F.init_neighbors();
  out->dim = D;
  out->n_nodes = int64_t(F.node_keys.size());
  out->nbr = F.nbr.data();
  out->d = F.dface.data();
  out->d_tr = F.dtr.data();
  out->wall = F.wall.data();
  double m = F.size(0);
  for (int a = 1; a < D; ++a) m = std::min(m, F.size(a));
  out->zero_distance_threshold = 1e-13 * m;  // EPS*MIN(tree dimensions)
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

int casl_forest_new(int dim, const double *box_min, const double *box_max, casl_forest_t **out) {
  return guarded([&]() {
    if (!out) throw std::invalid_argument("null out");
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    auto f = std::make_unique<casl_forest>();
    f->dim = dim;
    double *bmin, *bmax;
    if (dim == 2) {
      f->f2 = std::make_unique<ForestT<2>>();
      bmin = f->f2->bmin;
      bmax = f->f2->bmax;
    } else {
      f->f3 = std::make_unique<ForestT<3>>();
      bmin = f->f3->bmin;
      bmax = f->f3->bmax;
    }
    for (int a = 0; a < dim; ++a) {
      bmin[a] = box_min ? box_min[a] : 0.0;
      bmax[a] = box_max ? box_max[a] : 1.0;
      if (!(bmax[a] > bmin[a])) throw std::invalid_argument("empty box");
    }
    *out = f.release();
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
      if (n_nodes) *n_nodes = f->f2->nodes_built ? int64_t(f->f2->node_keys.size()) : -1;
      if (max_level) *max_level = f->f2->max_level();
    } else {
      if (n_leaves) *n_leaves = int64_t(f->f3->leaves.size());
      if (n_nodes) *n_nodes = f->f3->nodes_built ? int64_t(f->f3->node_keys.size()) : -1;
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
