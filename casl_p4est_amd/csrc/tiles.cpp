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
    // one scale per axis: a cell of level l spans base[a] * 2^-l along axis
    // a, the same l on every axis (the box edges may differ, cells need not
    // be cubes).  The largest arm cell along each axis is a cell of the
    // coarsest level present; a grid where that fails on some axis gives
    // inconsistent levels, which verify() rejects.
    for (int a = 0; a < D; ++a) base[a] = bmax[a];
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
    std::vector<uint16_t> ext;  // extension records of this tile's irregular records, in record order
    std::vector<uint32_t> urec;   // uniform records (UniRec<D>), T then H1
    int n_treg = 0, n_h1reg = 0;  // regular (non-uniform) records of T and of H1
    int n_tuni = 0, n_h1uni = 0;  // uniform records of T and of H1
  };
  static constexpr int UW = UniRec<D>::WORDS;

  // a regular record off the walls whose arms all sit at one level
  static bool uniform(const TileRec &r) {
    if (r.geo & (GEO_IRREGULAR | (63u << 24))) return false;
    for (int f = 1; f < F; ++f)
      if (((r.geo >> (4 * f)) & 15u) != (r.geo & 15u)) return false;
    return true;
  }
  static constexpr int NH = HangExt<D>::NH, WORDS = HangExt<D>::WORDS, SW = HangExt<D>::SW;

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

  // The record of node n; for a node with hanging faces also its extension
  // record (WORDS words at e) and, in *nh, how many faces hang (0: regular).
  TileRec encode(const Tile &t, int64_t n, uint16_t *e, int *nh) const {
    TileRec r{};
    uint32_t geo = 0;
    int hf[2] = {HANG_NONE, HANG_NONE}, k = 0;
    for (int f = 0; f < F; ++f) {
      geo |= uint32_t(level_of(f >> 1, arm_cell(f, n))) << (4 * f);
      if (wall(f, n)) geo |= 1u << (24 + f);
      double w[4];
      weights(f, n, w);
      int nz = 0, first = -1;
      for (int c = 0; c < C; ++c)
        if (w[c] != 0.0) {
          ++nz;
          if (first < 0) first = c;
        }
      if (first < 0) throw std::runtime_error("face without a nonzero interpolation weight");
      const int32_t l = local(t, nbr(f, first, n));
      if (l < 0 || l > 65535) throw std::runtime_error("tile-local index out of range");
      r.idx[f] = uint16_t(l);
      if (nz != 1) {
        if (k == NH) throw std::runtime_error("more hanging faces than the extension record holds");
        hf[k++] = f;
      }
    }
    *nh = k;
    if (k > 0) {
      geo |= GEO_IRREGULAR;
      std::fill(e, e + WORDS, uint16_t(0));
      e[0] = uint16_t(D == 2 ? hf[0] : (hf[0] | (hf[1] << 4)));
      for (int q = 0; q < k; ++q) {
        const int f = hf[q];
        int t1, t2;
        transverse(f >> 1, t1, t2);
        double w[4];
        weights(f, n, w);
        int first = -1;
        for (int c = 0; c < C; ++c)
          if (w[c] != 0.0 && first < 0) first = c;
        uint16_t *fe = e + 1 + q * SW;
        for (int c = 0; c < C; ++c) {
          const int32_t l = local(t, nbr(f, w[c] != 0.0 ? c : first, n));
          if (l < 0 || l > 65535) throw std::runtime_error("tile-local index out of range");
          fe[c] = uint16_t(l);
        }
        fe[C] = units(t1, dt(f, 0, n));
        if (D == 3) fe[C + 1] = units(t2, dt(f, 2, n));
      }
    }
    r.geo = geo;
    verify(t, n, r, e);
    return r;
  }

  // decode == reference stencil (distances bit-exact, nonzero-weight corners
  // equal); e is the extension record of an irregular record
  void verify(const Tile &t, int64_t n, const TileRec &r, const uint16_t *e) const {
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
    int hf[2] = {HANG_NONE, HANG_NONE};
    if (r.geo & GEO_IRREGULAR) {
      hf[0] = D == 2 ? e[0] : (e[0] & 15);
      if (D == 3) hf[1] = (e[0] >> 4) & 15;
      if (hf[0] >= F || (hf[1] != HANG_NONE && (hf[1] >= F || hf[1] <= hf[0])))
        throw std::runtime_error("tile record: bad hanging face");
    }
    for (int f = 0; f < F; ++f) {
      double w[4];
      weights(f, n, w);
      const int q = (f == hf[0]) ? 0 : (f == hf[1]) ? 1 : -1;
      if (q < 0) {
        // one corner carries the face value: the record's
        int nz = 0;
        for (int c = 0; c < C; ++c)
          if (w[c] != 0.0) {
            ++nz;
            if (local(t, nbr(f, c, n)) != r.idx[f]) throw std::runtime_error("tile record: corner mismatch");
          }
        if (nz != 1) throw std::runtime_error("tile record: hanging face not in the extension");
        continue;
      }
      int t1, t2;
      transverse(f >> 1, t1, t2);
      const uint16_t *fe = e + 1 + q * SW;
      for (int c = 0; c < C; ++c)
        if (w[c] != 0.0 && local(t, nbr(f, c, n)) != fe[c]) throw std::runtime_error("tile record: corner mismatch");
      const double qh1 = base[t1] * std::ldexp(1.0, -int((r.geo >> (4 * f)) & 15));
      const double m0 = fe[C] * unit[t1];
      if (m0 != dt(f, 0, n) || qh1 - m0 != dt(f, 1, n)) throw std::runtime_error("tile record: offset mismatch");
      if (D == 3) {
        const double qh2 = base[t2] * std::ldexp(1.0, -int((r.geo >> (4 * f)) & 15));
        const double zm = fe[C + 1] * unit[t2];
        if (zm != dt(f, 2, n) || qh2 - zm != dt(f, 3, n)) throw std::runtime_error("tile record: offset mismatch");
      }
    }
  }

  void make(Tile &t) const {
    const int64_t nT = t.e - t.s;
    t.rec.clear();
    t.slot.clear();
    t.ext.clear();
    t.urec.clear();
    const size_t nall = size_t(nT) + t.h1.size();
    std::vector<TileRec> all(nall);
    std::vector<uint16_t> exts(nall * WORDS);
    std::vector<int> kind(nall);  // number of hanging faces
    for (int64_t n = t.s; n < t.e; ++n) {
      const size_t j = size_t(n - t.s);
      all[j] = encode(t, n, &exts[j * WORDS], &kind[j]);
    }
    for (size_t i = 0; i < t.h1.size(); ++i) {
      const size_t j = size_t(nT) + i;
      all[j] = encode(t, t.h1[i], &exts[j * WORDS], &kind[j]);
    }
    // kind -1: uniform (compact record)
    for (size_t j = 0; j < nall; ++j)
      if (kind[j] == 0 && uniform(all[j])) kind[j] = -1;
    // group: T uniform, T regular, T hanging on 1 .. NH faces, then the same
    // for H1; extension records follow the irregular records in that order
    auto emit = [&](size_t lo, size_t hi, int k) {
      int cnt = 0;
      for (size_t j = lo; j < hi; ++j)
        if (kind[j] == k) {
          if (k < 0) {
            uint32_t w[UW];
            uni_pack<D>(all[j].idx, uint32_t(j), all[j].geo & 15u, w);
            t.urec.insert(t.urec.end(), w, w + UW);
          } else {
            t.rec.push_back(all[j]);
            t.slot.push_back(uint16_t(j));
            if (k > 0) t.ext.insert(t.ext.end(), exts.begin() + j * WORDS, exts.begin() + (j + 1) * WORDS);
          }
          ++cnt;
        }
      return cnt;
    };
    t.n_tuni = emit(0, size_t(nT), -1);
    t.n_treg = emit(0, size_t(nT), 0);
    for (int k = 1; k <= NH; ++k) emit(0, size_t(nT), k);
    t.n_h1uni = emit(size_t(nT), nall, -1);
    t.n_h1reg = emit(size_t(nT), nall, 0);
    for (int k = 1; k <= NH; ++k) emit(size_t(nT), nall, k);
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
    if (!tile_scales_supported(base, D))
      throw std::runtime_error("cells that are not cubes on a box whose edges are not dyadic");
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
    H.urec_start.assign(1, 0);
    H.urec.clear();
    H.nuni.clear();
    H.h1count.clear();
    H.halo_idx.clear();
    H.rec.clear();
    H.ext.clear();
    H.ext_start.assign(1, 0);
    H.nreg.clear();
    size_t n_ext = 0;
    for (auto &p : parts)
      for (auto &t : p) {
        H.tile_start.push_back(t.e);
        H.h1count.push_back(int32_t(t.h1.size()));
        H.halo_idx.insert(H.halo_idx.end(), t.h1.begin(), t.h1.end());
        H.halo_idx.insert(H.halo_idx.end(), t.h2.begin(), t.h2.end());
        H.halo_start.push_back(int32_t(H.halo_idx.size()));
        H.rec.insert(H.rec.end(), t.rec.begin(), t.rec.end());
        H.rec_slot.insert(H.rec_slot.end(), t.slot.begin(), t.slot.end());
        H.rec_start.push_back(int32_t(H.rec.size()));
        H.ext.insert(H.ext.end(), t.ext.begin(), t.ext.end());
        n_ext += t.ext.size() / WORDS;
        if (n_ext >= size_t(INT32_MAX)) throw std::runtime_error("too many irregular records");
        H.ext_start.push_back(int32_t(n_ext));
        H.nreg.push_back(uint32_t(t.n_treg) | (uint32_t(t.n_h1reg) << 16));
        H.urec.insert(H.urec.end(), t.urec.begin(), t.urec.end());
        H.urec_start.push_back(int32_t(H.urec.size() / UW));
        H.nuni.push_back(uint32_t(t.n_tuni) | (uint32_t(t.n_h1uni) << 16));
      }
    H.n_irregular = int64_t(n_ext);
    if (H.halo_idx.size() >= size_t(INT32_MAX) || H.rec.size() >= size_t(INT32_MAX) ||
        H.urec.size() / UW >= size_t(INT32_MAX))
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
