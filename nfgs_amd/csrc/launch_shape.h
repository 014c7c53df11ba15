// Launch shape of the z-marching tiled stencils (state pass, face pass,
// reconstruction, sources): the tile grid over (y, x), the planes per
// z-chunk and the number of workgroups of a launch.  Plain host arithmetic
// with no HIP dependency: the launchers, the plan's sizing of the monitor
// partials and the tuning checks all call these functions, so the partials a
// plan holds always cover the blocks any admissible launch writes.
#pragma once

namespace nfgs {
namespace shape {

constexpr int TX = 32, TY = 8;       // tile columns (x) and rows (y)
constexpr int kMaxTunedChunks = 16;  // largest z-chunk count a plan may be tuned to

// Tiles along an axis of n nodes with tiles of t: tiles start one node early
// when the last one would be a single node wide (the fused state pass needs
// two nodes of a tile for its one-sided end stencils); the count does not
// change otherwise.
inline int tiles_along(int n, int t) { return (n + (n % t == 1 ? 1 : 0) + t - 1) / t; }
inline long tiles_xy(int nx, int ny) { return (long)tiles_along(nx, TX) * tiles_along(ny, TY); }

// Chunks of length ceil(nz / count) that a launch holds.
inline int chunk_count(int nz, int len) { return (nz + len - 1) / len; }

// A tuned count: 1..kMaxTunedChunks chunks, every one non-empty, the last
// one at least two planes (the top node's one-sided dBz/dz takes Bz of
// planes nz-1, nz-2 and nz-3 from the chunk's own z-march).
inline bool zchunks_fit(int nz, int count) {
  if (count < 1 || count > kMaxTunedChunks || count > nz) return false;
  const int c = (nz + count - 1) / count;
  const int last = nz - (count - 1) * c;
  return count == 1 || last >= 2;
}

// Workgroups of the fused state pass resident on one CU at a time (36 KB of
// LDS and about 120 VGPRs per lane allow four), the residency the cost model
// plans for.
constexpr int kResidentBlocks = 4;

// Planes per chunk of the cost model.  A chunk re-stages about four planes
// (its Phi ring and the face below it), and a CU runs kResidentBlocks
// workgroups side by side, so the launch takes about
// ceil(blocks / (CUs * kResidentBlocks)) rounds of (chunk + 4) plane-steps;
// the chunk count minimises that, keeping the last chunk at least two planes
// long.  Of equal costs the larger count wins (its last round is the fuller
// one).  Measured against every count on MI355X: the model's choice is the
// fastest at 130x130x129, 386x386x129 and 514x514x257 and within 2% of it at
// 258x258x257.
inline int cost_model_chunk(int nx, int ny, int nz, long cus) {
  const long xy = tiles_xy(nx, ny);
  const long slots = cus * kResidentBlocks;
  long best = -1;
  int kc = nz;
  for (int nch = 1; nch <= (nz + 3) / 4; ++nch) {
    const int c = (nz + nch - 1) / nch;
    if (nch > 1 && nz - (nch - 1) * c < 2) continue;
    const long cost = ((xy * nch + slots - 1) / slots) * (c + 4);
    if (best < 0 || cost <= best) {
      best = cost;
      kc = c;
    }
  }
  return kc;
}

// Planes per chunk for a tuned count (zchunks > 0) or the cost model (0).
inline int chunk_len(int nx, int ny, int nz, int zchunks, long cus) {
  return zchunks > 0 ? (nz + zchunks - 1) / zchunks : cost_model_chunk(nx, ny, nz, cus);
}

// Workgroups of one tiled launch.
inline long blocks(int nx, int ny, int nz, int zchunks, long cus) {
  return tiles_xy(nx, ny) * chunk_count(nz, chunk_len(nx, ny, nz, zchunks, cus));
}

// Partials per monitor slot a plan holds: the largest block count of the
// cost model and of every admissible tuned count.
inline long partial_capacity(int nx, int ny, int nz, long cus) {
  long m = blocks(nx, ny, nz, 0, cus);
  for (int c = 1; c <= kMaxTunedChunks; ++c)
    if (zchunks_fit(nz, c)) {
      const long b = blocks(nx, ny, nz, c, cus);
      if (b > m) m = b;
    }
  return m;
}

}  // namespace shape
}  // namespace nfgs
