// Host check of the tiled-stencil launch shapes (nfgs_amd/csrc/launch_shape.h):
// for every grid and CU count below, every z-chunking a plan admits (the cost
// model and each tuned count zchunks_fit accepts) covers the planes exactly
// once, keeps its last chunk at least two planes long, and launches no more
// blocks than partial_capacity -- the per-slot partials a plan holds.  Prints
// "ok <cases>" or the first violation.
#include <cstdio>
#include <vector>

#include "launch_shape.h"

using namespace nfgs::shape;

static int fail(const char* what, int nx, int ny, int nz, int count, long cus) {
  std::printf("FAIL %s nx=%d ny=%d nz=%d count=%d cus=%ld\n", what, nx, ny, nz, count, cus);
  return 1;
}

static int check(int nx, int ny, int nz, long cus, long& cases) {
  const long cap = partial_capacity(nx, ny, nz, cus);
  for (int count = 0; count <= kMaxTunedChunks + 2; ++count) {
    if (count > 0 && !zchunks_fit(nz, count)) continue;
    const int len = chunk_len(nx, ny, nz, count, cus);
    const int nch = chunk_count(nz, len);
    if (len < 1) return fail("empty chunk length", nx, ny, nz, count, cus);
    if (count > 0 && nch != count) return fail("chunk count differs from the tuned count", nx, ny, nz, count, cus);
    // chunk c covers [c*len, min((c+1)*len, nz)): every plane once, none empty
    const int last = nz - (nch - 1) * len;
    if (last < 1 || last > len) return fail("chunks do not tile the planes", nx, ny, nz, count, cus);
    if (nch > 1 && last < 2) return fail("last chunk shorter than two planes", nx, ny, nz, count, cus);
    const long b = blocks(nx, ny, nz, count, cus);
    if (b != tiles_xy(nx, ny) * nch) return fail("block count", nx, ny, nz, count, cus);
    if (b > cap) return fail("launch exceeds the partials capacity", nx, ny, nz, count, cus);
    ++cases;
  }
  if (zchunks_fit(nz, kMaxTunedChunks + 1)) return fail("count above the maximum admitted", nx, ny, nz, 0, cus);
  return 0;
}

int main() {
  long cases = 0;
  const std::vector<int> lateral = {3, 4, 5, 9, 14, 17, 18, 33, 34, 42, 46, 50, 58, 66, 98, 130,
                                    194, 258, 386, 514};
  const std::vector<long> cus = {1, 80, 256, 304};
  for (long c : cus)
    for (int nx : lateral)
      for (int ny : lateral)
        for (int nz = 3; nz <= 700; nz += (nz < 80 ? 1 : 7))
          if (check(nx, ny, nz, c, cases)) return 1;
  // the grids of the batched scan (mudpack calculation grids of 128, 256 and
  // 320 columns over 128 planes) and of the headline solve
  const int grids[][3] = {{130, 130, 129}, {258, 258, 129}, {386, 386, 129}, {258, 258, 257}};
  for (auto& g : grids)
    if (check(g[0], g[1], g[2], 256, cases)) return 1;
  std::printf("ok %ld\n", cases);
  return 0;
}
