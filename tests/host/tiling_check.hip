// Host-side check of the tile placement the fused kernels and the reverse TV
// kernel use (col_place / tiles_across in fused.h): for every image width
// (or height) up to 5000, every column is owned by exactly one tile, a tile
// owns nothing closer than its halo to a side where it does not meet the
// image edge, and owned columns lie in the tile's load range.  Runs on the
// CPU; prints "ok" or the first failure.
#include <cstdio>

#include "fused.h"

namespace {

using scigap::fused::ColPlace;
using scigap::fused::col_place;
using scigap::fused::tiles_across;

// One placement geometry: load width LW, halo HC, interior width OX.
bool check(const char* what, int LW, int HC, int OX, int n_max) {
  static int owners[6000];
  for (int W = 1; W <= n_max; ++W) {
    for (int c = 0; c < W; ++c) owners[c] = 0;
    const int n = tiles_across(W, LW, HC, OX);
    for (int t = 0; t < n; ++t) {
      const ColPlace p = col_place(t, n, W, LW, HC, OX);
      // owned columns inside the image (the kernels also test column < W)
      const int lo = p.own_lo < 0 ? 0 : p.own_lo, hi = p.own_hi > W ? W : p.own_hi;
      if (lo >= hi) continue;
      for (int c = lo; c < hi; ++c) ++owners[c];
      if (lo < p.c0 || hi > p.c0 + LW) {
        std::printf("%s W=%d tile %d: owns [%d,%d) outside its load range [%d,%d)\n", what, W, t,
                    lo, hi, p.c0, p.c0 + LW);
        return false;
      }
      // a side that is not the image edge needs HC loaded columns of halo
      if (lo > 0 && lo - p.c0 < HC) {
        std::printf("%s W=%d tile %d: left halo %d < %d\n", what, W, t, lo - p.c0, HC);
        return false;
      }
      if (hi < W && p.c0 + LW - hi < HC) {
        std::printf("%s W=%d tile %d: right halo %d < %d\n", what, W, t, p.c0 + LW - hi, HC);
        return false;
      }
      // lanes hold column pairs: boundaries stay even when the columns are aligned
      if (W % 2 == 0 && W >= LW && (lo % 2 != 0 || hi % 2 != 0)) {
        std::printf("%s W=%d tile %d: odd ownership boundary [%d,%d)\n", what, W, t, lo, hi);
        return false;
      }
    }
    for (int c = 0; c < W; ++c)
      if (owners[c] != 1) {
        std::printf("%s W=%d: column %d has %d owners\n", what, W, c, owners[c]);
        return false;
      }
  }
  return true;
}

}  // namespace

int main() {
  constexpr int kMax = 5000;
  bool ok = true;
  // Fused outer iteration (columns): 128-wide load tiles, halo D rounded up to even.
  for (int D = 1; D <= 6 && ok; ++D) {
    const int HC = (D + 1) & ~1;
    ok = check("fused columns", 128, HC, 128 - 2 * HC, kMax);
  }
  // Reverse TV (columns and rows): 128 x 64 load tiles, halo 2D on each side.
  for (int D = 1; D <= 6 && ok; ++D) {
    ok = check("reverse columns", 128, 2 * D, 128 - 4 * D, kMax) &&
         check("reverse rows", 64, 2 * D, 64 - 4 * D, kMax);
  }
  std::printf(ok ? "ok\n" : "failed\n");
  return ok ? 0 : 1;
}
