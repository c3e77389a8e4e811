// A caller of the library written against include/casl_reinit.h alone, as a
// reference-side shim would be: it builds a forest and its stencils through
// the C ABI, reinitializes phi held in host memory (casl_reinit_host: plain,
// masked and mixed precision) and writes
//   [N][4] doubles: phi_0, FP64 (2,2), FP64 masked (2,2), mixed (2,2)
// plus the mask bytes [N] to the file named on the command line, for the test
// to compare with the oracle.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "casl_reinit.h"

#define CHECK(call)                                                                        \
  do {                                                                                     \
    const int rc_ = (call);                                                                \
    if (rc_ != CASL_OK) {                                                                  \
      std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, casl_last_error());         \
      return 1;                                                                            \
    }                                                                                      \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <out file> <iterations>\n", argv[0]);
    return 2;
  }
  const int iters = std::atoi(argv[2]);
  const double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0};
  // the 2D star (|x-c| - (0.3 + 0.05 cos 5 theta)) (1 + 0.8 |x-c|^2)
  const double star[6] = {0.5, 0.5, 0.3, 0.05, 5.0, 0.8};
  casl_levelset ls{CASL_LS_STAR, 0, star, nullptr, nullptr};
  casl_forest_t *f = nullptr;
  CHECK(casl_forest_new(2, lo, hi, &f));
  CHECK(casl_forest_refine_levelset(f, 3, 8, 2.4, 0.0, &ls));
  CHECK(casl_forest_build_nodes(f));
  int64_t n_leaves = 0, n = 0;
  int32_t max_level = 0;
  CHECK(casl_forest_sizes(f, &n_leaves, &n, &max_level));
  const size_t nn = static_cast<size_t>(n);
  std::vector<double> phi0(nn), xyz(2 * nn);
  std::vector<int32_t> ixyz(2 * nn);
  CHECK(casl_forest_sample(f, &ls, phi0.data()));
  CHECK(casl_forest_get_nodes(f, ixyz.data(), xyz.data()));
  casl_flat_stencil tab;
  CHECK(casl_forest_init_neighbors(f, &tab));
  casl_stencil_t *st = nullptr;
  CHECK(casl_stencil_create(&tab, &st));

  std::vector<uint8_t> mask(nn);
  for (int64_t i = 0; i < n; ++i) mask[size_t(i)] = xyz[size_t(2 * i + 1)] > 0.45 ? 1 : 0;
  const double big = 1.7976931348623157e308;
  std::vector<double> fp64 = phi0, masked = phi0, mixed = phi0;
  CHECK(casl_reinit_host(st, fp64.data(), nullptr, 2, 2, iters, big, -big, CASL_PRECISION_FP64, nullptr));
  CHECK(casl_reinit_host(st, masked.data(), mask.data(), 2, 2, iters, big, -big, CASL_PRECISION_FP64, nullptr));
  CHECK(casl_reinit_host(st, mixed.data(), nullptr, 2, 2, iters, big, -big, CASL_PRECISION_MIXED, nullptr));
  // a mask with other orders is refused, as the masked reference call has none
  if (casl_reinit_host(st, mixed.data(), mask.data(), 1, 2, iters, big, -big, 0, nullptr) !=
      CASL_ERR_INVALID_ARGUMENT) {
    std::fprintf(stderr, "a mask with orders (1, 2) was not refused\n");
    return 1;
  }

  FILE *out = std::fopen(argv[1], "wb");
  if (!out) return 1;
  for (int64_t i = 0; i < n; ++i) {
    const double row[4] = {phi0[size_t(i)], fp64[size_t(i)], masked[size_t(i)], mixed[size_t(i)]};
    std::fwrite(row, sizeof(double), 4, out);
  }
  std::fwrite(mask.data(), 1, mask.size(), out);
  std::fclose(out);
  CHECK(casl_stencil_destroy(st));
  CHECK(casl_forest_destroy(f));
  std::printf("%lld nodes, %d iterations\n", static_cast<long long>(n), iters);
  return 0;
}
