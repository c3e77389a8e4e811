// Host builder of the tiled stencil encoding (see tiles.hpp).
#include "tiles.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "common.hpp"

namespace casl {

namespace {

template <int D>
struct Builder {
  static constexpr int F = 2 * D, C = 1 << (D - 1), T = 2 * (D - 1);
  const casl_flat_stencil &tab;
  const int64_t N;
  double base[3] = {0, 0, 0};  // coarsest arm-cell size per axis
  int Lrel = 0;                // finest relative level
  double unit[3] = {0, 0, 0};

  explicit Builder(const casl_flat_stencil &t) : tab(t), N(t.n_nodes) {}

  int32_t nbr(int f, int c, int64_t n) const { return tab.nbr[(int64_t(f) * C + c) * N + n]; }
  double d(int f, int64_t n) const { return tab.d[int64_t(f) * N + n]; }
  double dt(int f, int k, int64_t n) const { return tab.d_tr[(int64_t(f) * T + k) * N + n]; }
  bool wall(int f, int64_t n) const { return (tab.wall[n] >> f) & 1; }
  static void transverse(int a, int &t1, int &t2) {
    if (D == 2) {
      t1 = 1 - a;
      t2 = -1;
    } else {
      t1 = (a == 0 ? 1 : 0);
      t2 = (a == 2 ? 1 : 2);
    }
  }
  // size of the cell that carries face f's neighbors
  double arm_cell(int f, int64_t n) const {
    return wall(f, n) ? (-d(f, n) - d(f ^ 1, n)) : d(f, n);
  }
  // weights of the C corners (get_linear_interpolation_weights)
  void weights(int f, int64_t n, double *w) const {
    const double m0 = dt(f, 0, n), p0 = dt(f, 1, n);
    if (D == 2) {
      w[0] = p0;
      w[1] = m0;
    } else {
      const double zm = dt(f, 2, n), zp = dt(f, 3, n);
      w[0] = p0 * zp;
      w[1] = m0 * zp;
      w[2] = p0 * zm;
      w[3] = m0 * zm;
    }
  }
  // nodes with nonzero weight in n's stencil
  int needed(int64_t n, int32_t *out) const {
    int k = 0;
    for (int f = 0; f < F; ++f) {
      double w[4];
      weights(f, n, w);
      for (int c = 0; c < C; ++c)
        if (w[c] != 0.0) out[k++] = nbr(f, c, n);
    }
    return k;
  }

  int level_of(int a, double qh) const {
    if (!(qh > 0)) throw std::runtime_error("non-positive arm cell size");
    const int l = int(std::lround(std::log2(base[a] / qh)));
    if (l < 0 || l > 15 || base[a] * std::ldexp(1.0, -l) != qh)
      throw std::runtime_error("arm cell size is not a power-of-two fraction of the coarsest cell");
    return l;
  }

  void scan_scales() {
    std::vector<double> bmax(D, 0.0);
    std::mutex mu;
    parallel_for(N, [&](int64_t b, int64_t e) {
      double loc[3] = {0, 0, 0};
      for (int64_t n = b; n < e; ++n)
        for (int f = 0; f < F; ++f) loc[f >> 1] = std::max(loc[f >> 1], arm_cell(f, n));
      std::lock_guard<std::mutex> g(mu);
      for (int a = 0; a < D; ++a) bmax[a] = std::max(bmax[a], loc[a]);
    });
    // one scale for all axes: cells are cubes in logical space, so the
    // transverse extent of an arm cell equals its extent along the face axis
    // (boxes whose edges differ by non-power-of-two ratios fail verify()).
    double b = 0.0;
    for (int a = 0; a < D; ++a) b = std::max(b, bmax[a]);
    for (int a = 0; a < D; ++a) base[a] = b;
    std::atomic<int> lr{0};
    parallel_for(N, [&](int64_t b, int64_t e) {
      int m = 0;
      for (int64_t n = b; n < e; ++n)
        for (int f = 0; f < F; ++f) m = std::max(m, level_of(f >> 1, arm_cell(f, n)));
      int cur = lr.load();
      while (m > cur && !lr.compare_exchange_weak(cur, m)) {
      }
    });
    Lrel = lr.load();
    for (int a = 0; a < D; ++a) unit[a] = base[a] * std::ldexp(1.0, -Lrel);
  }

  uint16_t units(int ax, double v) const {
    const double k = v / unit[ax];
    if (!(k >= 0) || k > 65535 || std::floor(k) != k || double(k) * unit[ax] != v)
      throw std::runtime_error("transverse offset is not a multiple of the finest cell");
    return uint16_t(k);
  }

  struct Tile {
    int32_t s, e;
    std::vector<int32_t> h1, h2;
    std::vector<TileRec> rec;
    std::vector<uint16_t> slot;
    std::vector<uint16_t> ext;  // extensions of this tile's irregular records
  };

  // local index of global node g in tile (s, e, h1, h2)
  static int32_t local(const Tile &t, int32_t g) {
    if (g >= t.s && g < t.e) return g - t.s;
    auto it = std::lower_bound(t.h1.begin(), t.h1.end(), g);
    if (it != t.h1.end() && *it == g) return int32_t(t.e - t.s) + int32_t(it - t.h1.begin());
    auto jt = std::lower_bound(t.h2.begin(), t.h2.end(), g);
    if (jt != t.h2.end() && *jt == g)
      return int32_t(t.e - t.s) + int32_t(t.h1.size()) + int32_t(jt - t.h2.begin());
    return -1;
  }

  bool halo(Tile &t) const {
    const int64_t nT = t.e - t.s;
    std::vector<int32_t> buf(F * C);
    t.h1.clear();
    t.h2.clear();
    for (int64_t n = t.s; n < t.e; ++n) {
      const int k = needed(n, buf.data());
      for (int i = 0; i < k; ++i)
        if (buf[i] < t.s || buf[i] >= t.e) t.h1.push_back(buf[i]);
    }
    std::sort(t.h1.begin(), t.h1.end());
    t.h1.erase(std::unique(t.h1.begin(), t.h1.end()), t.h1.end());
    if (nT + int64_t(t.h1.size()) > TileLimits<D>::DMAX) return false;
    for (int32_t h : t.h1) {
      const int k = needed(h, buf.data());
      for (int i = 0; i < k; ++i) {
        const int32_t g = buf[i];
        if (g >= t.s && g < t.e) continue;
        if (std::binary_search(t.h1.begin(), t.h1.end(), g)) continue;
        t.h2.push_back(g);
      }
    }
    std::sort(t.h2.begin(), t.h2.end());
    t.h2.erase(std::unique(t.h2.begin(), t.h2.end()), t.h2.end());
    return nT + int64_t(t.h1.size()) + int64_t(t.h2.size()) <= TileLimits<D>::LMAX;
  }

  TileRec encode(const Tile &t, int64_t n, std::vector<uint16_t> &ext) const {
    TileRec r{};
    uint32_t geo = 0;
    bool irregular = false;
    for (int f = 0; f < F; ++f) {
      geo |= uint32_t(level_of(f >> 1, arm_cell(f, n))) << (4 * f);
      if (wall(f, n)) geo |= 1u << (24 + f);
      double w[4];
      weights(f, n, w);
      int nz = 0;
      for (int c = 0; c < C; ++c) nz += (w[c] != 0.0);
      if (nz != 1) irregular = true;
    }
    if (!irregular) {
      for (int f = 0; f < F; ++f) {
        double w[4];
        weights(f, n, w);
        for (int c = 0; c < C; ++c)
          if (w[c] != 0.0) {
            const int32_t l = local(t, nbr(f, c, n));
            if (l < 0 || l > 65535) throw std::runtime_error("tile-local index out of range");
            r.idx[f] = uint16_t(l);
          }
      }
    } else {
      geo |= GEO_IRREGULAR;
      const size_t id = ext.size() / (F * (D == 2 ? 3 : 6));
      if (id > 0xFFFFFFFFull) throw std::runtime_error("too many irregular nodes");
      r.idx[0] = uint16_t(id & 0xFFFF);
      r.idx[1] = uint16_t(id >> 16);
      for (int f = 0; f < F; ++f) {
        int t1, t2;
        transverse(f >> 1, t1, t2);
        double w[4];
        weights(f, n, w);
        int first = -1;
        for (int c = 0; c < C; ++c)
          if (w[c] != 0.0 && first < 0) first = c;
        for (int c = 0; c < C; ++c) {
          const int32_t l = local(t, nbr(f, w[c] != 0.0 ? c : first, n));
          if (l < 0 || l > 65535) throw std::runtime_error("tile-local index out of range");
          ext.push_back(uint16_t(l));
        }
        ext.push_back(units(t1, dt(f, 0, n)));
        if (D == 3) ext.push_back(units(t2, dt(f, 2, n)));
      }
    }
    r.geo = geo;
    verify(t, n, r, ext);
    return r;
  }

  // decode == reference stencil (distances bit-exact, nonzero-weight corners equal)
  void verify(const Tile &t, int64_t n, const TileRec &r, const std::vector<uint16_t> &ext) const {
    double dd[6];
    for (int f = 0; f < F; ++f) {
      const int lv = (r.geo >> (4 * f)) & 15;
      dd[f] = base[f >> 1] * std::ldexp(1.0, -lv);
    }
    for (int f = 0; f < F; ++f) {
      if ((r.geo >> (24 + f)) & 1) {
        const int lv = (r.geo >> (4 * f)) & 15;
        dd[f] = -dd[f ^ 1] - base[f >> 1] * std::ldexp(1.0, -lv);
      }
      if (dd[f] != d(f, n)) throw std::runtime_error("tile record: face distance mismatch");
    }
    if (r.geo & GEO_IRREGULAR) {
      const size_t id = size_t(r.idx[0]) | (size_t(r.idx[1]) << 16);
      const uint16_t *e = &ext[id * F * (D == 2 ? 3 : 6)];
      for (int f = 0; f < F; ++f) {
        int t1, t2;
        transverse(f >> 1, t1, t2);
        const uint16_t *fe = e + f * (D == 2 ? 3 : 6);
        const double qh1 = base[t1] * std::ldexp(1.0, -int((r.geo >> (4 * f)) & 15));
        const double m0 = fe[C] * unit[t1];
        const double p0 = qh1 - m0;
        if (m0 != dt(f, 0, n) || p0 != dt(f, 1, n)) throw std::runtime_error("tile record: offset mismatch");
        if (D == 3) {
          const double qh2 = base[t2] * std::ldexp(1.0, -int((r.geo >> (4 * f)) & 15));
          const double zm = fe[C + 1] * unit[t2];
          if (zm != dt(f, 2, n) || qh2 - zm != dt(f, 3, n)) throw std::runtime_error("tile record: offset mismatch");
        }
      }
    } else {
      for (int f = 0; f < F; ++f) {
        double w[4];
        weights(f, n, w);
        for (int c = 0; c < C; ++c)
          if (w[c] != 0.0 && local(t, nbr(f, c, n)) != r.idx[f])
            throw std::runtime_error("tile record: corner mismatch");
      }
    }
  }

  void make(Tile &t) const {
    const int64_t nT = t.e - t.s;
    t.rec.clear();
    t.slot.clear();
    t.ext.clear();
    std::vector<TileRec> all;
    all.reserve(nT + t.h1.size());
    for (int64_t n = t.s; n < t.e; ++n) all.push_back(encode(t, n, t.ext));
    for (int32_t h : t.h1) all.push_back(encode(t, h, t.ext));
    // group: T regular, T hanging, H1 regular, H1 hanging
    auto emit = [&](size_t lo, size_t hi, bool irregular) {
      for (size_t j = lo; j < hi; ++j)
        if (bool(all[j].geo & GEO_IRREGULAR) == irregular) {
          t.rec.push_back(all[j]);
          t.slot.push_back(uint16_t(j));
        }
    };
    emit(0, size_t(nT), false);
    emit(0, size_t(nT), true);
    emit(size_t(nT), all.size(), false);
    emit(size_t(nT), all.size(), true);
  }

  // splits [s, e) until every piece fits the LDS limits
  void cut(int32_t s, int32_t e, std::vector<Tile> &out) const {
    Tile t;
    t.s = s;
    t.e = e;
    if (halo(t)) {
      make(t);
      out.push_back(std::move(t));
      return;
    }
    if (e - s <= 1) throw std::runtime_error("a single node exceeds the tile limits");
    const int32_t m = s + (e - s) / 2;
    cut(s, m, out);
    cut(m, e, out);
  }

  void run(TileHost &H) {
    scan_scales();
    const int64_t nCand = (N + TileLimits<D>::TMAX - 1) / TileLimits<D>::TMAX;
    std::vector<std::vector<Tile>> parts(nCand);
    std::atomic<int64_t> cursor{0};
    std::string err;
    std::mutex emu;
    std::vector<std::thread> pool;
    for (int th = 0; th < num_threads(); ++th)
      pool.emplace_back([&]() {
        for (;;) {
          const int64_t i = cursor.fetch_add(1);
          if (i >= nCand) break;
          try {
            const int32_t s = int32_t(i * TileLimits<D>::TMAX);
            const int32_t e = int32_t(std::min<int64_t>(N, (i + 1) * TileLimits<D>::TMAX));
            cut(s, e, parts[i]);
          } catch (const std::exception &x) {
            std::lock_guard<std::mutex> g(emu);
            err = x.what();
            cursor.store(nCand);
          }
        }
      });
    for (auto &th : pool) th.join();
    if (!err.empty()) throw std::runtime_error(err);
    // concatenate
    H.dim = D;
    H.lmax = Lrel;
    for (int a = 0; a < D; ++a) H.size[a] = base[a];
    H.tile_start.assign(1, 0);
    H.halo_start.assign(1, 0);
    H.rec_start.assign(1, 0);
    H.h1count.clear();
    H.halo_idx.clear();
    H.rec.clear();
    H.ext.clear();
    const int ext_words = F * (D == 2 ? 3 : 6);
    size_t n_ext = 0;
    for (auto &p : parts)
      for (auto &t : p) {
        H.tile_start.push_back(t.e);
        H.h1count.push_back(int32_t(t.h1.size()));
        H.halo_idx.insert(H.halo_idx.end(), t.h1.begin(), t.h1.end());
        H.halo_idx.insert(H.halo_idx.end(), t.h2.begin(), t.h2.end());
        H.halo_start.push_back(int32_t(H.halo_idx.size()));
        // rebase extension ids of this tile
        const size_t base_id = n_ext;
        for (auto &r : t.rec) {
          if (r.geo & GEO_IRREGULAR) {
            const size_t id = (size_t(r.idx[0]) | (size_t(r.idx[1]) << 16)) + base_id;
            if (id > 0xFFFFFFFFull) throw std::runtime_error("too many irregular records");
            r.idx[0] = uint16_t(id & 0xFFFF);
            r.idx[1] = uint16_t(id >> 16);
          }
          H.rec.push_back(r);
        }
        H.rec_slot.insert(H.rec_slot.end(), t.slot.begin(), t.slot.end());
        H.rec_start.push_back(int32_t(H.rec.size()));
        H.ext.insert(H.ext.end(), t.ext.begin(), t.ext.end());
        n_ext += t.ext.size() / ext_words;
      }
    H.n_irregular = int64_t(n_ext);
    if (H.halo_idx.size() >= size_t(INT32_MAX) || H.rec.size() >= size_t(INT32_MAX))
      throw std::runtime_error("tile tables too large for int32 offsets");
  }
};

}  // namespace

void build_tiles(const casl_flat_stencil &tab, TileHost &out) {
  if (tab.dim == 2) {
    Builder<2> b(tab);
    b.run(out);
  } else {
    Builder<3> b(tab);
    b.run(out);
  }
}

}  // namespace casl
