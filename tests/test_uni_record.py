"""The compact uniform record (UniRec in csrc/tiles.hpp) round-trips every
field at its full width: corner slots, the record's own LDS slot and the arm
level, in 2D (two words, 12-bit fields) and 3D (three words, 13-bit corner
fields).  A small host program built with g++ against the header packs and
unpacks pseudo-random records."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdint>
#include <cstdio>
#include "tiles.hpp"
using namespace casl;

template <int D>
int check(uint64_t &state) {
  auto next = [&]() {
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return uint32_t(state >> 33);
  };
  constexpr int F = 2 * D;
  const uint32_t imax = 1u << UniRec<D>::IDX_BITS, smax = 1u << UniRec<D>::SLOT_BITS;
  for (int it = 0; it < 200000; ++it) {
    uint16_t idx[6] = {0, 0, 0, 0, 0, 0}, out[6] = {0, 0, 0, 0, 0, 0};
    for (int f = 0; f < F; ++f) idx[f] = uint16_t(it < 4 ? (it & 1 ? imax - 1 : 0) : next() % imax);
    const uint32_t slot = it < 4 ? (it & 2 ? smax - 1 : 0) : next() % smax;
    const uint32_t level = it < 4 ? (it & 1 ? 15 : 0) : next() % 16;
    uint32_t w[3] = {0, 0, 0};
    uni_pack<D>(idx, slot, level, w);
    int s = -1, l = -1;
    uni_unpack<D>(w, out, s, l);
    if (s != int(slot) || l != int(level) || uni_slot<D>(w) != int(slot)) return 1;
    for (int f = 0; f < F; ++f)
      if (out[f] != idx[f]) return 1;
  }
  return 0;
}

int main() {
  uint64_t state = 20261202;
  const int e2 = check<2>(state), e3 = check<3>(state);
  std::printf("%d %d\n", e2, e3);
  return e2 | e3;
}
"""


def test_uniform_record_round_trip(tmp_path):
    src = tmp_path / "uni.cpp"
    exe = tmp_path / "uni"
    src.write_text(PROGRAM)
    cc = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "casl_p4est_amd", "csrc"),
                         "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                        capture_output=True, text=True)
    if cc.returncode != 0:
        pytest.fail(cc.stderr[-2000:])
    run = subprocess.run([str(exe)], capture_output=True, text=True)
    assert run.returncode == 0, run.stdout + run.stderr
