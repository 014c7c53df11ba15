// Point-wise stencil kernels of one NFGS iteration and the metric reduction.
//
// Layout: (component, z, y, x), x fastest; one lane per node, consecutive
// lanes on consecutive x so that every load is a contiguous run.  Neighbour
// samples are re-read through L1/L2.  Reductions keep FP64 per-block
// partials (fixed grid, fixed order) so every run sums identically.

#include "plan.h"

namespace nfgs {
namespace {

constexpr int TX = shape::TX, TY = shape::TY, HH = 2;
constexpr int PW = TX + 2 * HH, PH = TY + 2 * HH, PS = PW * PH;

// J.B / |B|^2 of the alpha projection: FP32 takes the hardware reciprocal
// (within 2 ulp of the quotient), FP64 the IEEE division.
__device__ __forceinline__ float alpha_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double alpha_div(double a, double b) { return a / b; }

template <typename T>
struct Grid {
  int nx, ny, nz;
  long plane, N;
  T dx, dy, dz;
  T idx2, idy2;
  T i2dx, i2dy, i2dz, idz;  // 1/(2 dx), 1/(2 dy), 1/(2 dz), 1/dz
  int tgx, tgy;  // tiles along x and y (the launch is 1-D over tgx*tgy*z-chunks)
  int sx, sy;    // tiling shift: 1 when a last tile would be one node wide
  int kc;        // planes per z-chunk
};

// Planes per z-chunk of the tiled stencils: the plan's tuned count or the
// cost model (launch_shape.h).
int z_chunk(const nfgs_plan* p) {
  return shape::chunk_len(p->d.nx, p->d.ny, p->d.nz, p->zchunks, cu_count_stencil());
}

template <typename T>
Grid<T> make_grid(const nfgs_plan* p) {
  Grid<T> g;
  g.nx = p->d.nx;
  g.ny = p->d.ny;
  g.nz = p->d.nz;
  g.plane = p->plane;
  g.N = p->N;
  g.dx = T(p->d.dx);
  g.dy = T(p->d.dy);
  g.dz = T(p->d.dz);
  g.idx2 = T(1.0 / (p->d.dx * p->d.dx));
  g.idy2 = T(1.0 / (p->d.dy * p->d.dy));
  g.i2dx = T(1.0 / (2.0 * p->d.dx));
  g.i2dy = T(1.0 / (2.0 * p->d.dy));
  g.i2dz = T(1.0 / (2.0 * p->d.dz));
  g.idz = T(1.0 / p->d.dz);
  // Tiles start at -s so that no tile is a single node wide (the fused
  // state pass needs two nodes of a tile for its one-sided end stencils);
  // the tile count does not change.
  g.sx = g.nx % TX == 1 ? 1 : 0;
  g.sy = g.ny % TY == 1 ? 1 : 0;
  g.tgx = (g.nx + g.sx + TX - 1) / TX;
  g.tgy = (g.ny + g.sy + TY - 1) / TY;
  g.kc = z_chunk(p);
  return g;
}

// ---------------------------------------------------------------------------
// z-marching tiles.  A block owns a TX x TY column tile and a chunk of z
// planes; the planes a stencil needs are staged in LDS with a 2-node halo
// (enough for the one-sided end stencils), and the block walks up in z so
// every array is read once per plane.  Halo points outside the domain hold a
// clamped in-domain value and are never used by a valid stencil (end stencils
// are one-sided).
// ---------------------------------------------------------------------------

struct Tile {
  int id;  // logical tile id (x fastest, then y, then z-chunk)
  int x0, y0, k0, tx, ty, x, y;
  bool valid;
  bool narrow;  // a last-column tile holding at most TX/4 grid columns
  int c;        // LDS index of this thread's point
};

template <typename T>
__device__ __forceinline__ Tile make_tile(const Grid<T>& g) {
  Tile t;
  t.id = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = t.id % g.tgx, r = t.id / g.tgx;
  t.x0 = bx * TX - g.sx;
  t.y0 = (r % g.tgy) * TY - g.sy;
  t.k0 = (r / g.tgy) * g.kc;
  // A narrow tile (the grid's last few columns, e.g. 2 of 32 on a 258-wide
  // grid) numbers its points column-major, so its valid points fill the first
  // wave and the other waves, whose lanes all lie outside the grid, skip the
  // per-point work.
  t.narrow = g.nx - t.x0 <= TX / 4;
  t.tx = t.narrow ? (int)threadIdx.x / TY : (int)threadIdx.x % TX;
  t.ty = t.narrow ? (int)threadIdx.x % TY : (int)threadIdx.x / TX;
  t.x = t.x0 + t.tx;
  t.y = t.y0 + t.ty;
  t.valid = t.x >= 0 && t.y >= 0 && t.x < g.nx && t.y < g.ny;
  t.c = (t.ty + HH) * PW + t.tx + HH;
  return t;
}

// Two-step staging of a plane region (tile + halo): fetch() issues the
// global loads into registers, commit() writes them to an LDS slot later, so
// the loads of plane p+1 are in flight while plane p is being computed.
constexpr int PPT = (PS + 255) / 256;  // staged elements per thread (256 threads)
constexpr int PSA = PPT * 256;         // LDS slot size: every thread stores PPT words

template <typename T>
struct Fetch {
  T v[PPT];
};

// In-plane byte offsets of the elements this thread stages, computed once per
// block.  Halo points outside the domain take a clamped in-domain offset
// (never used by a valid stencil); slots past PS repeat the last element.
struct Halo {
  unsigned off[PPT];
  unsigned own;  // byte offset of the thread's own point (clamped)
};

template <typename T>
__device__ __forceinline__ Halo make_halo(const Tile& t, const Grid<T>& g) {
  Halo h;
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    int e = threadIdx.x + 256 * i;
    e = e < PS ? e : PS - 1;
    const int ly = e / PW, lx = e - ly * PW;
    int y = t.y0 + ly - HH, x = t.x0 + lx - HH;
    y = y < 0 ? 0 : (y >= g.ny ? g.ny - 1 : y);
    x = x < 0 ? 0 : (x >= g.nx ? g.nx - 1 : x);
    h.off[i] = (unsigned)((y * g.nx + x) * (int)sizeof(T));
  }
  int x = t.x, y = t.y;
  x = x < 0 ? 0 : (x >= g.nx ? g.nx - 1 : x);
  y = y < 0 ? 0 : (y >= g.ny ? g.ny - 1 : y);
  h.own = (unsigned)((y * g.nx + x) * (int)sizeof(T));
  return h;
}

template <typename T>
__device__ __forceinline__ void fetch(Fetch<T>& r, const T* plane, const Halo& h) {
#pragma unroll
  for (int i = 0; i < PPT; ++i) r.v[i] = ld(plane, h.off[i]);
}

template <typename T>
__device__ __forceinline__ void commit(T* slot, const Fetch<T>& r) {
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    const int e = threadIdx.x + 256 * i;
    slot[e] = r.v[i];
  }
}

template <typename T>
__device__ __forceinline__ void commit_diff(T* slot, const Fetch<T>& a, const Fetch<T>& b) {
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    const int e = threadIdx.x + 256 * i;
    slot[e] = a.v[i] - b.v[i];
  }
}

// Derivative helpers of the state pass.  INNER tiles (every point the pass
// touches is at least one node from the grid edge) take the centred forms.
// Edge tiles evaluate the centred and the one-sided second-order end form
// for every lane and select, so their lanes do not diverge: the end form at
// the last node is the negated first-node form with the inward step reversed
// (the same value, rounding included).  The tiling keeps two inward nodes of
// every end node in the staged region, so every read is in bounds; reads a
// lane does not select are discarded by the select.
template <typename G, typename T>
__device__ __forceinline__ T fd_sel(G g, int i, int n, T r) {
  const bool lo = i == 0, hi = i == n - 1;
  const T fm = g(-1), fp = g(1), f0 = g(0), f2 = g(hi ? -2 : 2);
  const T one = (T(-3) * f0 + T(4) * (hi ? fm : fp)) - f2;
  const T d = (lo || hi) ? (hi ? -one : one) : (fp - fm);
  return d * r;
}
template <bool INNER, typename T>
__device__ __forceinline__ T fdr(const T* f, long s, int i, int n, T r) {
  if (INNER) return (f[s] - f[-s]) * r;
  return fd_sel([&](int o) { return f[o * s]; }, i, n, r);
}
template <typename T, bool INNER, typename G>
__device__ __forceinline__ T fdgr(G g, int i, int n, T r) {
  if (INNER) return (g(1) - g(-1)) * r;
  return fd_sel(g, i, n, r);
}
// Mirror-Neumann 5-point lateral Laplacian; at an end node the outward
// neighbour's weight moves to the inward one (weights 2, -2, 0).
template <bool INNER, typename T>
__device__ __forceinline__ T lapr(const T* p, int y, int x, int ny, int nx, long ld, T idx2,
                                  T idy2) {
  const T* c = p + (long)y * ld + x;
  if (INNER) {
    T acc = T(0);
    acc += (c[1] - T(2) * c[0] + c[-1]) * idx2;
    acc += (c[ld] - T(2) * c[0] + c[-ld]) * idy2;
    return acc;
  }
  const T wxp = x == 0 ? T(2) : (x == nx - 1 ? T(0) : T(1));
  const T wxm = x == nx - 1 ? T(2) : (x == 0 ? T(0) : T(1));
  const T wyp = y == 0 ? T(2) : (y == ny - 1 ? T(0) : T(1));
  const T wym = y == ny - 1 ? T(2) : (y == 0 ? T(0) : T(1));
  const T ax = (wxp * c[1] - T(2) * c[0] + wxm * c[-1]) * idx2;
  const T ay = (wyp * c[ld] - T(2) * c[0] + wym * c[-ld]) * idy2;
  return ax + ay;
}

// lapr with the centred form chosen per axis: IX (IY) when no point the pass
// touches lies on an x (y) end of the grid.
template <bool IX, bool IY, typename T>
__device__ __forceinline__ T lapr2(const T* p, int y, int x, int ny, int nx, long ld, T idx2,
                                   T idy2) {
  if (IX && IY) return lapr<true>(p, y, x, ny, nx, ld, idx2, idy2);
  if (!IX && !IY) return lapr<false>(p, y, x, ny, nx, ld, idx2, idy2);
  const T* c = p + (long)y * ld + x;
  T ax, ay;
  if (IX) {
    ax = (c[1] - T(2) * c[0] + c[-1]) * idx2;
  } else {
    const T wxp = x == 0 ? T(2) : (x == nx - 1 ? T(0) : T(1));
    const T wxm = x == nx - 1 ? T(2) : (x == 0 ? T(0) : T(1));
    ax = (wxp * c[1] - T(2) * c[0] + wxm * c[-1]) * idx2;
  }
  if (IY) {
    ay = (c[ld] - T(2) * c[0] + c[-ld]) * idy2;
  } else {
    const T wyp = y == 0 ? T(2) : (y == ny - 1 ? T(0) : T(1));
    const T wym = y == ny - 1 ? T(2) : (y == 0 ? T(0) : T(1));
    ay = (wyp * c[ld] - T(2) * c[0] + wym * c[-ld]) * idy2;
  }
  return ax + ay;
}

template <typename T>
__device__ __forceinline__ T lds_dx(const T* s, const Tile& t, const Grid<T>& g) {
  return fdr<false>(s + t.c, 1, t.x, g.nx, g.i2dx);
}
template <typename T>
__device__ __forceinline__ T lds_dy(const T* s, const Tile& t, const Grid<T>& g) {
  return fdr<false>(s + t.c, (long)PW, t.y, g.ny, g.i2dy);
}
template <typename T>
__device__ __forceinline__ T lds_lap(const T* s, const Tile& t, const Grid<T>& g) {
  return lapr<false>(s + t.c - ((long)t.y * PW + t.x), t.y, t.x, g.ny, g.nx, (long)PW, g.idx2,
                     g.idy2);
}

// Register set of the loads one node of the reconstruction needs.
template <typename T>
struct ReconSet {
  Fetch<T> phi, az;  // top Phi plane of the node's ring, Az plane
  T gx, gy, gz;      // B_bg at the thread's own point
};

// Reconstruction of B from the scalars (nfgs_solver.py:1711-1715, :3311-3341,
// :5901-5916, :6069-6078).
template <typename T>
__global__ __launch_bounds__(256) void reconstruct_b_kernel(
    Grid<T> g, const T* __restrict__ phi, const T* __restrict__ az, const T* __restrict__ bg,
    const T* __restrict__ bz_obs, const T* __restrict__ hprof, T* __restrict__ b) {
  __shared__ T sphi[3][PSA];
  __shared__ T saz[PSA];
  __shared__ T sdz[PSA];
  __shared__ T sown[3][256];  // B_bg at the own points of the current node
  const Tile t = make_tile(g);
  const Halo h = make_halo(t, g);
  const int k0 = t.k0, k1 = min(k0 + g.kc, g.nz);
  auto pl = [&](auto* base, int k) { return base + (long)k * g.plane; };
  // Phi planes lo..lo+2 are resident while node k is built
  auto lo_of = [&](int k) { return max(0, min(k - 1, g.nz - 3)); };
  {
    const int lo = lo_of(k0);
    Fetch<T> f0, f1, f2;
    fetch(f0, pl(phi, lo), h);
    fetch(f1, pl(phi, lo + 1), h);
    fetch(f2, pl(phi, lo + 2), h);
    commit(sphi[lo % 3], f0);
    commit(sphi[(lo + 1) % 3], f1);
    commit(sphi[(lo + 2) % 3], f2);
  }
  // Loads of node k+2 are issued while node k is computed (two register
  // sets, alternating), so two planes of loads per block are in flight.
  // Every step loads (node indices clamped to the grid) so that the wait
  // counts before each commit are exact.
  auto fetch_set = [&](ReconSet<T>& q, int k) {
    const int kc = min(k, g.nz - 1);
    fetch(q.phi, pl(phi, lo_of(kc) + 2), h);
    fetch(q.az, pl(az, kc), h);
    q.gx = ld(pl(bg, kc), h.own);
    q.gy = ld(pl(bg + g.N, kc), h.own);
    q.gz = ld(pl(bg + 2 * g.N, kc), h.own);
  };
  auto step = [&](ReconSet<T>& q, int k) {
    const bool act = k < k1;
    if (act) {
      if (k > k0 && lo_of(k) > lo_of(k - 1)) commit(sphi[(lo_of(k) + 2) % 3], q.phi);
      commit(saz, q.az);
      // own-point values also pass through LDS, so their registers are free
      // for the next loads as soon as they are committed
      sown[0][threadIdx.x] = q.gx;
      sown[1][threadIdx.x] = q.gy;
      sown[2][threadIdx.x] = q.gz;
    }
    lds_barrier();
    fetch_set(q, k + 2);
    if (act) {
      // dPhi/dz on the tile with halo
#pragma unroll
      for (int i = 0; i < PPT; ++i) {
        const int e = threadIdx.x + 256 * i;  // padding slots hold unused values
        T d;
        if (k == 0)
          d = (T(-3) * sphi[0][e] + T(4) * sphi[1][e] - sphi[2][e]) * g.i2dz;
        else if (k == g.nz - 1)
          d = (T(3) * sphi[k % 3][e] - T(4) * sphi[(k - 1) % 3][e] + sphi[(k - 2) % 3][e]) *
              g.i2dz;
        else
          d = (sphi[(k + 1) % 3][e] - sphi[(k - 1) % 3][e]) * g.i2dz;
        sdz[e] = d;
      }
    }
    lds_barrier();
    if (act && t.valid) {
      const T cgx = sown[0][threadIdx.x], cgy = sown[1][threadIdx.x], cgz = sown[2][threadIdx.x];
      const T bcx = -lds_dx(sdz, t, g) + lds_dy(saz, t, g);
      const T bcy = -lds_dy(sdz, t, g) - lds_dx(saz, t, g);
      const T bcz = lds_lap(sphi[k % 3], t, g);
      T bx = cgx + bcx;
      T by = cgy + bcy;
      T bz = cgz + bcz;
      if (k == 0) bz = ld(bz_obs, h.own);
      st(pl(b, k), h.own, bx + hprof[k]);
      st(pl(b + g.N, k), h.own, by + hprof[g.nz + k]);
      st(pl(b + 2 * g.N, k), h.own, bz);
    }
    lds_barrier();
  };
  ReconSet<T> sa, sb;
  fetch_set(sa, k0);
  fetch_set(sb, k0 + 1);
  for (int k = k0; k < k1; k += 2) {
    step(sa, k);
    step(sb, k + 1);
  }
}

// Vertical-face pass (nfgs_solver.py:5738-5839, :5882-5895, :2689-2697,
// :5917-5941, :6702-6710).
template <typename T>
struct FaceVal {
  T b[3], j[3], p[3];  // face B, face current, alpha-projected current
};

struct Window {
  int z0, z1, y0, y1, x0, x1;
};

// With jc (a caller's current J, 3 volumes) the centre Jz is J[2] and the
// bottom row of J_par is J[:, 0], as the reference's iteration_step takes
// them from its J argument (:6690-6711); az is then not read.  Without jc,
// Jz is -lap_xy(Az) with its boundary values (the state's own current).
template <typename T>
__global__ __launch_bounds__(256) void face_pass_kernel(
    Grid<T> g, const T* __restrict__ B, const T* __restrict__ bg, const T* __restrict__ az,
    const T* __restrict__ jc, const T* __restrict__ jzt, T* __restrict__ jp,
    T* __restrict__ jout, double* partials, long pstride, T floor_b2, T amax, Window w) {
  __shared__ T scz[2][PSA];
  __shared__ T sbx[2][PSA];
  __shared__ T sby[2][PSA];
  __shared__ T saz[PSA];
  __shared__ T sown[3][256];  // Bz, B_bg,x, B_bg,y at the own points of plane p
  __shared__ double red[NFGS_NQ * 16];
  const Tile t = make_tile(g);
  const Halo h = make_halo(t, g);
  const int k0 = t.k0, k1 = min(k0 + g.kc, g.nz);
  const int ps = max(k0 - 1, 0), pe = min(k1, g.nz - 1);
  auto pl = [&](auto* base, int k) { return base + (long)k * g.plane; };
  const bool edge = t.y == 0 || t.y == g.ny - 1 || t.x == 0 || t.x == g.nx - 1;
  const bool in_win = t.y >= w.y0 && t.y < w.y1 && t.x >= w.x0 && t.x < w.x1;
  const T* Bx = B;
  const T* By = B + g.N;
  const T* Bz = B + 2 * g.N;
  // bottom current at the own point: the toroidal one, or the caller's J[2, 0]
  const T jz0 = jc ? ld(jc + 2 * g.N, h.own) : ld(jzt, h.own);
  // per-thread sums over the z-chunk in the solver dtype; blocks combine
  // them in FP64
  T acc[7] = {0, 0, 0, 0, 0, 0, 0};
  // previous plane at the own point; pz2 = Bz two planes below
  T pb[3] = {0, 0, 0}, pg[2] = {0, 0}, pjz = 0, pz2 = 0;
  FaceVal<T> fprev;
  Fetch<T> fbz, fgz, fbx, fby, faz;
  T obz, ogx, ogy, ojz = 0;  // own Bz, bg_x, bg_y (and caller Jz) of the fetched plane
  // Every iteration loads a plane (clamped to pe) so the waits before each
  // commit are exact.
  auto fetch_plane = [&](int q) {
    q = min(q, pe);
    fetch(fbz, pl(Bz, q), h);
    fetch(fgz, pl(bg + 2 * g.N, q), h);
    fetch(fbx, pl(Bx, q), h);
    fetch(fby, pl(By, q), h);
    if (jc)
      ojz = ld(pl(jc + 2 * g.N, q), h.own);
    else
      fetch(faz, pl(az, q), h);
    obz = ld(pl(Bz, q), h.own);
    ogx = ld(pl(bg, q), h.own);
    ogy = ld(pl(bg + g.N, q), h.own);
  };
  fetch_plane(ps);
  for (int p = ps; p <= pe; ++p) {
    commit_diff(scz[p & 1], fbz, fgz);
    commit(sbx[p & 1], fbx);
    commit(sby[p & 1], fby);
    if (!jc) commit(saz, faz);
    sown[0][threadIdx.x] = obz;
    sown[1][threadIdx.x] = ogx;
    sown[2][threadIdx.x] = ogy;
    const T jz_given = ojz;
    lds_barrier();
    fetch_plane(p + 1);
    const T cbz = sown[0][threadIdx.x], cgx = sown[1][threadIdx.x], cgy = sown[2][threadIdx.x];
    T cb[3] = {0, 0, 0}, cg[2] = {0, 0}, cjz = 0;
    if (t.valid) {
      cb[0] = sbx[p & 1][t.c];
      cb[1] = sby[p & 1][t.c];
      cb[2] = cbz;
      cg[0] = cgx;
      cg[1] = cgy;
      if (jc)
        cjz = jz_given;
      else if (p == 0)
        cjz = jz0;
      else if (p == g.nz - 1 || edge)
        cjz = T(0);
      else
        cjz = -lds_lap(saz, t, g);
    }
    if (p > ps && t.valid) {
      const int f = p - 1;  // face between planes f and p
      FaceVal<T> F;
      F.b[0] = T(0.5) * (cb[0] + pb[0]);
      F.b[1] = T(0.5) * (cb[1] + pb[1]);
      F.b[2] = T(0.5) * (cb[2] + pb[2]);
      const T dbx = ((cb[0] - cg[0]) - (pb[0] - pg[0])) * g.idz;
      const T dby = ((cb[1] - cg[1]) - (pb[1] - pg[1])) * g.idz;
      const T* s1 = scz[p & 1];
      const T* s0 = scz[f & 1];
      auto cfz = [&](int o) { return T(0.5) * (s1[t.c + o] + s0[t.c + o]); };
      const T dyc = fdgr<T, false>([&](int o) { return cfz(o * PW); }, t.y, g.ny, g.i2dy);
      const T dxc = fdgr<T, false>([&](int o) { return cfz(o); }, t.x, g.nx, g.i2dx);
      F.j[0] = dyc - dby;
      F.j[1] = dbx - dxc;
      F.j[2] = T(0.5) * (cjz + pjz);
      {
        T b2 = F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2];
        b2 = b2 < floor_b2 ? floor_b2 : b2;
        const T jb = F.j[0] * F.b[0] + F.j[1] * F.b[1] + F.j[2] * F.b[2];
        T a = alpha_div(jb, b2);
        a = a < -amax ? -amax : (a > amax ? amax : a);
        F.p[0] = a * F.b[0];
        F.p[1] = a * F.b[1];
        F.p[2] = a * F.b[2];
      }
      const int k = f;  // node finalised now
      if (k >= k0 && k < k1) {
        if (jp) {
          if (k == 0 && jc) {
            st(pl(jp, k), h.own, ld(jc, h.own));
            st(pl(jp + g.N, k), h.own, ld(jc + g.N, h.own));
            st(pl(jp + 2 * g.N, k), h.own, jz0);
          } else if (k == 0) {
            st(pl(jp, k), h.own, edge ? T(0) : F.j[0]);
            st(pl(jp + g.N, k), h.own, edge ? T(0) : F.j[1]);
            st(pl(jp + 2 * g.N, k), h.own, jz0);
          } else {
            st(pl(jp, k), h.own, T(0.5) * (fprev.p[0] + F.p[0]));
            st(pl(jp + g.N, k), h.own, T(0.5) * (fprev.p[1] + F.p[1]));
            st(pl(jp + 2 * g.N, k), h.own, T(0.5) * (fprev.p[2] + F.p[2]));
          }
        }
        if (jout) {
          T c[3];
          if (k == 0) {
            c[0] = F.j[0]; c[1] = F.j[1]; c[2] = F.j[2];
          } else {
            c[0] = T(0.5) * (fprev.j[0] + F.j[0]);
            c[1] = T(0.5) * (fprev.j[1] + F.j[1]);
            c[2] = T(0.5) * (fprev.j[2] + F.j[2]);
          }
          acc[6] += c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
          acc[5] += F.j[0] * F.j[0] + F.j[1] * F.j[1] + F.j[2] * F.j[2];
          st(pl(jout, k), h.own, edge ? T(0) : c[0]);
          st(pl(jout + g.N, k), h.own, edge ? T(0) : c[1]);
          st(pl(jout + 2 * g.N, k), h.own, pjz);
        }
        if (partials) {
          if (k >= w.z0 && k < w.z1 && in_win) {
            T bn = msqrt(F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2]);
            bn = bn < T(1e-10) ? T(1e-10) : bn;
            const T jn = msqrt(F.j[0] * F.j[0] + F.j[1] * F.j[1] + F.j[2] * F.j[2]);
            const T c0 = F.j[1] * F.b[2] - F.j[2] * F.b[1];
            const T c1 = F.j[2] * F.b[0] - F.j[0] * F.b[2];
            const T c2 = F.j[0] * F.b[1] - F.j[1] * F.b[0];
            acc[0] += mdiv(msqrt(c0 * c0 + c1 * c1 + c2 * c2), bn);
            acc[1] += jn;
          }
          const T dvx = lds_dx(sbx[k & 1], t, g);
          const T dvy = lds_dy(sby[k & 1], t, g);
          // d(Bz)/dz at node k from the planes held in registers (the
          // bottom node also needs plane 2)
          T dvz;
          if (k == 0)
            dvz = (T(-3) * pb[2] + T(4) * cb[2] - ld(pl(Bz, 2), h.own)) * g.i2dz;
          else
            dvz = (cb[2] - pz2) * g.i2dz;
          const T div = dvx + dvy + dvz;
          const T b2 = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
          acc[2] += div < T(0) ? -div : div;
          acc[3] += msqrt(b2);
          acc[4] += b2;
        }
      }
      fprev = F;
    }
    pz2 = pb[2];
    pb[0] = cb[0]; pb[1] = cb[1]; pb[2] = cb[2];
    pg[0] = cg[0]; pg[1] = cg[1];
    pjz = cjz;
    lds_barrier();
  }
  // top node: no face above it
  const int kt = g.nz - 1;
  if (k1 == g.nz && t.valid) {
    if (jp) {
      st(pl(jp, kt), h.own, fprev.p[0]);
      st(pl(jp + g.N, kt), h.own, fprev.p[1]);
      st(pl(jp + 2 * g.N, kt), h.own, fprev.p[2]);
    }
    if (jout) {
      acc[6] += fprev.j[0] * fprev.j[0] + fprev.j[1] * fprev.j[1] + fprev.j[2] * fprev.j[2];
      st(pl(jout, kt), h.own, T(0));
      st(pl(jout + g.N, kt), h.own, T(0));
      st(pl(jout + 2 * g.N, kt), h.own, T(0));
    }
    if (partials) {
      const T dvx = lds_dx(sbx[kt & 1], t, g);
      const T dvy = lds_dy(sby[kt & 1], t, g);
      // one-sided d(Bz)/dz: planes kt (pb), kt-1 (pz2) and kt-2
      const T dvz = (T(3) * pb[2] - T(4) * pz2 + ld(pl(Bz, kt - 2), h.own)) * g.i2dz;
      const T div = dvx + dvy + dvz;
      const T b2 = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
      acc[2] += div < T(0) ? -div : div;
      acc[3] += msqrt(b2);
      acc[4] += b2;
    }
  }
  if (!partials) return;
  double accd[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) accd[q] = (double)acc[q];
  block_sum<7>(accd, red);
  if (threadIdx.x == 0) {
    const long bid = t.id;
    partials[NFGS_Q_SIGMA_NUM * pstride + bid] = accd[0];
    partials[NFGS_Q_SIGMA_DEN * pstride + bid] = accd[1];
    partials[NFGS_Q_DIV_ABS * pstride + bid] = accd[2];
    partials[NFGS_Q_BMAG * pstride + bid] = accd[3];
    partials[NFGS_Q_B2 * pstride + bid] = accd[4];
    partials[NFGS_Q_FACE_J2 * pstride + bid] = accd[5];
    partials[NFGS_Q_CELL_J2 * pstride + bid] = accd[6];
  }
}

// ---------------------------------------------------------------------------
// Fused state pass: reconstruction of B (nfgs_solver.py:1711-1715,
// :3311-3341, :5901-5916, :6069-6078) and the vertical-face pass
// (:5738-5839, :5882-5895, :2689-2697, :5917-5941, :6702-6710) in one z-march.
// B of plane p is built in LDS on the tile plus a one-node ring (R1) from
// Phi_c and Az staged with a two-node ring, then consumed by the face pass of
// the same block; B reaches HBM only when b_out is given.  Arithmetic is the
// same, operation by operation, as reconstruct_b_kernel followed by
// face_pass_kernel.
// ---------------------------------------------------------------------------
constexpr int R1W = PW - 2, R1H = PH - 2, R1N = R1W * R1H;  // ring-1 region
constexpr int R1PT = (R1N + 255) / 256;

template <typename T>
struct StateSet {
  Fetch<T> phi, az, gx, gy;  // top Phi plane of the ring, Az, B_bg x and y
  T gz;                      // B_bg z at the own point
};

template <typename T>
struct StateLds {
  T sphi[3][PSA];
  T saz[2][PSA];
  T sgx[2][PSA];
  T sgy[2][PSA];
  T sdz[PSA];
  T sbx[2][PSA];
  T sby[2][PSA];
  T scz[2][PSA];  // Bz - B_bg,z on R1
  T sgz[2][256];
  double red[NFGS_NQ * 16];
};

template <typename T, bool IX, bool IY, bool ITER>
__device__ __forceinline__ void state_pass_body(
    StateLds<T>& L, const Tile& t, const Grid<T>& g, const T* __restrict__ phi,
    const T* __restrict__ az, const T* __restrict__ bg, const T* __restrict__ bz_obs,
    const T* __restrict__ hprof, const T* __restrict__ jzt, const T* __restrict__ jzc,
    const T* __restrict__ jh, T* __restrict__ bout, T* __restrict__ jp, T* __restrict__ jout,
    double* partials, long pstride, T floor_b2, T amax, const Window& w) {
  const Halo h = make_halo(t, g);
  const int k0 = t.k0, k1 = min(k0 + g.kc, g.nz);
  const int ps = max(k0 - 1, 0), pe = min(k1, g.nz - 1);
  // ITER: the iteration's outputs (J_par and the monitor partials, no J)
  if (ITER) jout = nullptr;
  const bool has_jp = ITER || jp != nullptr;
  const bool has_pa = ITER || partials != nullptr;
  auto pl = [&](auto* base, int k) { return base + (long)k * g.plane; };
  auto lo_of = [&](int k) { return max(0, min(k - 1, g.nz - 3)); };
  const bool edge = (!IY && (t.y == 0 || t.y == g.ny - 1)) || (!IX && (t.x == 0 || t.x == g.nx - 1));
  const bool in_win = t.y >= w.y0 && t.y < w.y1 && t.x >= w.x0 && t.x < w.x1;
  const T jz0 = ld(jzt, h.own);

  // ring of Phi planes lo..lo+2 for the first plane
  {
    const int lo = lo_of(ps);
    Fetch<T> f0, f1, f2;
    fetch(f0, pl(phi, lo), h);
    fetch(f1, pl(phi, lo + 1), h);
    fetch(f2, pl(phi, lo + 2), h);
    commit(L.sphi[lo % 3], f0);
    commit(L.sphi[(lo + 1) % 3], f1);
    commit(L.sphi[(lo + 2) % 3], f2);
  }
  auto fetch_set = [&](StateSet<T>& q, int p) {
    const int pc = min(p, pe);
    fetch(q.phi, pl(phi, lo_of(pc) + 2), h);
    fetch(q.az, pl(az, pc), h);
    fetch(q.gx, pl(bg, pc), h);
    fetch(q.gy, pl(bg + g.N, pc), h);
    q.gz = ld(pl(bg + 2 * g.N, pc), h.own);
  };

  // per-thread sums over the z-chunk in the solver dtype; blocks combine
  // them in FP64
  T acc[7] = {0, 0, 0, 0, 0, 0, 0};
  T pb[3] = {0, 0, 0}, pg[2] = {0, 0}, pjz = 0, pz2 = 0, pz3 = 0;
  FaceVal<T> fprev;

  auto step = [&](StateSet<T>& q, int p) {
    const bool act = p <= pe;
    const int sl = p & 1;
    // ring slots of planes p, p+1 (= p-2) and p-1
    const int r0 = p % 3, rp = r0 == 2 ? 0 : r0 + 1, rm = r0 == 0 ? 2 : r0 - 1;
    if (act) {
      // a new top plane (lo_of(p) + 2 = p + 1) enters the ring at interior p
      if (p > ps && lo_of(p) > lo_of(p - 1)) commit(L.sphi[rp], q.phi);
      commit(L.saz[sl], q.az);
      commit(L.sgx[sl], q.gx);
      commit(L.sgy[sl], q.gy);
      L.sgz[sl][threadIdx.x] = q.gz;
    }
    lds_barrier();
    fetch_set(q, p + 1);
    if (act) {
      // (A) dPhi/dz on every staged point
#pragma unroll
      for (int i = 0; i < PPT; ++i) {
        const int e = threadIdx.x + 256 * i;
        T d;
        if (p == 0)
          d = (T(-3) * L.sphi[0][e] + T(4) * L.sphi[1][e] - L.sphi[2][e]) * g.i2dz;
        else if (p == g.nz - 1)
          d = (T(3) * L.sphi[r0][e] - T(4) * L.sphi[rm][e] + L.sphi[rp][e]) * g.i2dz;
        else
          d = (L.sphi[rp][e] - L.sphi[rm][e]) * g.i2dz;
        L.sdz[e] = d;
      }
    }
    lds_barrier();
    if (act) {
      // (B) B on R1; points outside the grid hold 0 and are never read
      const T hx = hprof[p], hy = hprof[g.nz + p];
#pragma unroll
      for (int i = 0; i < R1PT; ++i) {
        const int r = threadIdx.x + 256 * i;
        int ly, lx;
        bool on;
        if (!IX && t.narrow) {
          // narrow tiles: column-major, so the in-grid ring points share a wave
          on = r < R1N;
          ly = 1 + r % R1H;
          lx = 1 + r / R1H;
        } else if (i == 0) {
          // the tile's own columns, one row per half-wave (conflict-free LDS
          // banks); the first TY ring rows
          on = true;
          ly = 1 + (int)threadIdx.x / TX;
          lx = HH + (int)threadIdx.x % TX;
        } else if ((int)threadIdx.x < (R1H - TY) * TX) {
          // the ring's last two rows, own columns
          on = true;
          ly = 1 + TY + (int)threadIdx.x / TX;
          lx = HH + (int)threadIdx.x % TX;
        } else {
          // the ring's two side columns
          const int j = (int)threadIdx.x - (R1H - TY) * TX;
          on = j < 2 * R1H;
          ly = 1 + (j >> 1);
          lx = (j & 1) ? R1W : 1;
        }
        if (on) {
          const int e = ly * PW + lx;
          const int y = t.y0 + ly - HH, x = t.x0 + lx - HH;
          T vbx = T(0), vby = T(0), vcz = T(0);
          if ((IX || (x >= 0 && x < g.nx)) && (IY || (y >= 0 && y < g.ny))) {
            const T dxd = fdr<IX>(L.sdz + e, 1, x, g.nx, g.i2dx);
            const T dyd = fdr<IY>(L.sdz + e, (long)PW, y, g.ny, g.i2dy);
            const T dxa = fdr<IX>(L.saz[sl] + e, 1, x, g.nx, g.i2dx);
            const T dya = fdr<IY>(L.saz[sl] + e, (long)PW, y, g.ny, g.i2dy);
            const T bcx = -dxd + dya;
            const T bcy = -dyd - dxa;
            vbx = (L.sgx[sl][e] + bcx) + hx;
            vby = (L.sgy[sl][e] + bcy) + hy;
            if (p == 0) {
              const unsigned o = (unsigned)((y * g.nx + x) * (int)sizeof(T));
              vcz = ld(bz_obs, o) - ld(bg + 2 * g.N, o);
            } else {
              vcz = lapr2<IX, IY>(L.sphi[r0] + e - (y * PW + x), y, x, g.ny, g.nx, (long)PW, g.idx2,
                              g.idy2);
            }
          }
          L.sbx[sl][e] = vbx;
          L.sby[sl][e] = vby;
          L.scz[sl][e] = vcz;
        }
      }
    }
    lds_barrier();
    if (!act) return;
    // (C) face pass at the own point
    T cb[3] = {0, 0, 0}, cg[2] = {0, 0}, cjz = 0;
    if (((IX && IY) || t.valid)) {
      cb[0] = L.sbx[sl][t.c];
      cb[1] = L.sby[sl][t.c];
      cb[2] = p == 0 ? ld(bz_obs, h.own) : L.sgz[sl][threadIdx.x] + L.scz[sl][t.c];
      cg[0] = L.sgx[sl][t.c];
      cg[1] = L.sgy[sl][t.c];
      if (jzc)
        cjz = ld(pl(jzc, p), h.own);  // a closure's cell J_z, boundary values included
      else if (p == 0)
        cjz = jz0;
      else if (p == g.nz - 1 || edge)
        cjz = T(0);
      else
        cjz = -lapr2<IX, IY>(L.saz[sl] + t.c - (t.y * PW + t.x), t.y, t.x, g.ny, g.nx, (long)PW,
                         g.idx2, g.idy2);
      if (bout && p >= k0 && p < k1) {
        st(pl(bout, p), h.own, cb[0]);
        st(pl(bout + g.N, p), h.own, cb[1]);
        st(pl(bout + 2 * g.N, p), h.own, cb[2]);
      }
    }
    if (p > ps && ((IX && IY) || t.valid)) {
      const int f = p - 1;  // face between planes f and p
      FaceVal<T> F;
      F.b[0] = T(0.5) * (cb[0] + pb[0]);
      F.b[1] = T(0.5) * (cb[1] + pb[1]);
      F.b[2] = T(0.5) * (cb[2] + pb[2]);
      const T dbx = ((cb[0] - cg[0]) - (pb[0] - pg[0])) * g.idz;
      const T dby = ((cb[1] - cg[1]) - (pb[1] - pg[1])) * g.idz;
      const T* s1 = L.scz[sl];
      const T* s0 = L.scz[f & 1];
      auto cfz = [&](int o) { return T(0.5) * (s1[t.c + o] + s0[t.c + o]); };
      const T dyc = fdgr<T, IY>([&](int o) { return cfz(o * PW); }, t.y, g.ny, g.i2dy);
      const T dxc = fdgr<T, IX>([&](int o) { return cfz(o); }, t.x, g.nx, g.i2dx);
      F.j[0] = dyc - dby;
      F.j[1] = dbx - dxc;
      F.j[2] = T(0.5) * (cjz + pjz);
      {
        T b2 = F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2];
        b2 = b2 < floor_b2 ? floor_b2 : b2;
        const T jb = F.j[0] * F.b[0] + F.j[1] * F.b[1] + F.j[2] * F.b[2];
        T a = alpha_div(jb, b2);
        a = a < -amax ? -amax : (a > amax ? amax : a);
        F.p[0] = a * F.b[0];
        F.p[1] = a * F.b[1];
        F.p[2] = a * F.b[2];
      }
      const int k = f;  // node finalised now
      if (k >= k0 && k < k1) {
        if (has_jp) {
          if (k == 0) {
            // the side walls carry the closure's wall current (none: zero)
            st(pl(jp, k), h.own, edge ? (jh ? jh[0] : T(0)) : F.j[0]);
            st(pl(jp + g.N, k), h.own, edge ? (jh ? jh[g.nz] : T(0)) : F.j[1]);
            st(pl(jp + 2 * g.N, k), h.own, jz0);
          } else {
            st(pl(jp, k), h.own, T(0.5) * (fprev.p[0] + F.p[0]));
            st(pl(jp + g.N, k), h.own, T(0.5) * (fprev.p[1] + F.p[1]));
            st(pl(jp + 2 * g.N, k), h.own, T(0.5) * (fprev.p[2] + F.p[2]));
          }
        }
        if (!ITER && jout) {
          T c[3];
          if (k == 0) {
            c[0] = F.j[0]; c[1] = F.j[1]; c[2] = F.j[2];
          } else {
            c[0] = T(0.5) * (fprev.j[0] + F.j[0]);
            c[1] = T(0.5) * (fprev.j[1] + F.j[1]);
            c[2] = T(0.5) * (fprev.j[2] + F.j[2]);
          }
          acc[6] += c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
          acc[5] += F.j[0] * F.j[0] + F.j[1] * F.j[1] + F.j[2] * F.j[2];
          st(pl(jout, k), h.own, edge ? (jh ? jh[k] : T(0)) : c[0]);
          st(pl(jout + g.N, k), h.own, edge ? (jh ? jh[g.nz + k] : T(0)) : c[1]);
          st(pl(jout + 2 * g.N, k), h.own, pjz);
        }
        if (has_pa) {
          if (k >= w.z0 && k < w.z1 && in_win) {
            T bn = msqrt(F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2]);
            bn = bn < T(1e-10) ? T(1e-10) : bn;
            const T jn = msqrt(F.j[0] * F.j[0] + F.j[1] * F.j[1] + F.j[2] * F.j[2]);
            const T c0 = F.j[1] * F.b[2] - F.j[2] * F.b[1];
            const T c1 = F.j[2] * F.b[0] - F.j[0] * F.b[2];
            const T c2 = F.j[0] * F.b[1] - F.j[1] * F.b[0];
            acc[0] += mdiv(msqrt(c0 * c0 + c1 * c1 + c2 * c2), bn);
            acc[1] += jn;
          }
          const T dvx = fdr<IX>(L.sbx[k & 1] + t.c, 1, t.x, g.nx, g.i2dx);
          const T dvy = fdr<IY>(L.sby[k & 1] + t.c, (long)PW, t.y, g.ny, g.i2dy);
          T dvz;
          if (k == 0) {
            // Bz of plane 2 = B_bg,z + lap_xy(Phi_c) there (Phi ring holds 0..2)
            const T bz2 = ld(pl(bg + 2 * g.N, 2), h.own) +
                          lapr2<IX, IY>(L.sphi[2] + t.c - (t.y * PW + t.x), t.y, t.x, g.ny, g.nx,
                                    (long)PW, g.idx2, g.idy2);
            dvz = (T(-3) * pb[2] + T(4) * cb[2] - bz2) * g.i2dz;
          } else {
            dvz = (cb[2] - pz2) * g.i2dz;
          }
          const T div = dvx + dvy + dvz;
          const T b2 = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
          acc[2] += div < T(0) ? -div : div;
          acc[3] += msqrt(b2);
          acc[4] += b2;
        }
      }
      fprev = F;
    }
    pz3 = pz2;
    pz2 = pb[2];
    pb[0] = cb[0]; pb[1] = cb[1]; pb[2] = cb[2];
    pg[0] = cg[0]; pg[1] = cg[1];
    pjz = cjz;
  };
  StateSet<T> sa;
  fetch_set(sa, ps);
  for (int p = ps; p <= pe; ++p) step(sa, p);
  // top node: no face above it (planes kt, kt-1, kt-2 of Bz are pb, pz2, pz3)
  const int kt = g.nz - 1;
  if (k1 == g.nz && ((IX && IY) || t.valid)) {
    if (has_jp) {
      st(pl(jp, kt), h.own, fprev.p[0]);
      st(pl(jp + g.N, kt), h.own, fprev.p[1]);
      st(pl(jp + 2 * g.N, kt), h.own, fprev.p[2]);
    }
    if (!ITER && jout) {
      acc[6] += fprev.j[0] * fprev.j[0] + fprev.j[1] * fprev.j[1] + fprev.j[2] * fprev.j[2];
      st(pl(jout, kt), h.own, T(0));
      st(pl(jout + g.N, kt), h.own, T(0));
      st(pl(jout + 2 * g.N, kt), h.own, T(0));
    }
    if (has_pa) {
      const T dvx = fdr<IX>(L.sbx[kt & 1] + t.c, 1, t.x, g.nx, g.i2dx);
      const T dvy = fdr<IY>(L.sby[kt & 1] + t.c, (long)PW, t.y, g.ny, g.i2dy);
      const T dvz = (T(3) * pb[2] - T(4) * pz2 + pz3) * g.i2dz;
      const T div = dvx + dvy + dvz;
      const T b2 = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
      acc[2] += div < T(0) ? -div : div;
      acc[3] += msqrt(b2);
      acc[4] += b2;
    }
  }
  if (!has_pa) return;
  double accd[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) accd[q] = (double)acc[q];
  block_sum<7>(accd, L.red);
  if (threadIdx.x == 0) {
    const long bid = t.id;
    partials[NFGS_Q_SIGMA_NUM * pstride + bid] = accd[0];
    partials[NFGS_Q_SIGMA_DEN * pstride + bid] = accd[1];
    partials[NFGS_Q_DIV_ABS * pstride + bid] = accd[2];
    partials[NFGS_Q_BMAG * pstride + bid] = accd[3];
    partials[NFGS_Q_B2 * pstride + bid] = accd[4];
    partials[NFGS_Q_FACE_J2 * pstride + bid] = accd[5];
    partials[NFGS_Q_CELL_J2 * pstride + bid] = accd[6];
  }
}

template <typename T, bool ITER>
__global__ __launch_bounds__(256) void state_pass_kernel(
    Grid<T> g, const T* __restrict__ phi, const T* __restrict__ az, const T* __restrict__ bg,
    const T* __restrict__ bz_obs, const T* __restrict__ hprof, const T* __restrict__ jzt,
    const T* __restrict__ jzc, const T* __restrict__ jh, T* __restrict__ bout,
    T* __restrict__ jp, T* __restrict__ jout, double* partials, long pstride, T floor_b2,
    T amax, Window w) {
  __shared__ StateLds<T> L;
  const Tile t = make_tile(g);
  // centred stencils along an axis when no point the pass touches lies on
  // that axis' ends; the one-sided selects only where a tile meets an end
  const bool ix = t.x0 >= 2 && t.x0 + TX <= g.nx - 2;
  const bool iy = t.y0 >= 2 && t.y0 + TY <= g.ny - 2;
#define NFGS_STATE_BODY(IX_, IY_)                                                            \
  state_pass_body<T, IX_, IY_, ITER>(L, t, g, phi, az, bg, bz_obs, hprof, jzt, jzc, jh, bout,  \
                                     jp, jout, partials, pstride, floor_b2, amax, w)
  if (ix && iy)
    NFGS_STATE_BODY(true, true);
  else if (ix)
    NFGS_STATE_BODY(true, false);
  else if (iy)
    NFGS_STATE_BODY(false, true);
  else
    NFGS_STATE_BODY(false, false);
#undef NFGS_STATE_BODY
}

// Register set of the loads one node of the sources needs.
template <typename T>
struct SourceSet {
  Fetch<T> x, y, z;  // J_x, J_y of plane k+1; J_z of plane k
};

// Sources of the two Poisson problems (nfgs_solver.py:6726-6745, :1729-1749).
template <typename T>
__global__ __launch_bounds__(256) void sources_kernel(Grid<T> g, const T* __restrict__ jp,
                                                      T* __restrict__ sb, T* __restrict__ sj,
                                                      double* partials, long pstride, T i4dxdz,
                                                      T i4dydz) {
  __shared__ T sjx[3][PSA];
  __shared__ T sjy[3][PSA];
  __shared__ T sjz[PSA];
  __shared__ double red[NFGS_NQ * 16];
  const Tile t = make_tile(g);
  const Halo h = make_halo(t, g);
  const int k0 = t.k0, k1 = min(k0 + g.kc, g.nz);
  auto pl = [&](auto* base, int k) { return base + (long)k * g.plane; };
  const T* jx = jp;
  const T* jy = jp + g.N;
  const T* jzp = jp + 2 * g.N;
  const int nxi = g.nx - 2;
  const long ip = (long)(g.ny - 2) * nxi;
  const bool interior = t.y >= 1 && t.y <= g.ny - 2 && t.x >= 1 && t.x <= g.nx - 2;
  // byte offset of the own point in the compact interior planes of S_Jz
  const unsigned sjo = interior ? (unsigned)(((t.y - 1) * nxi + (t.x - 1)) * (int)sizeof(T)) : 0u;
  T acc1 = T(0);  // sum of S_Bcz^2 over the thread's nodes (FP64 per block)
  for (int q = max(k0 - 1, 0); q <= k0; ++q) {
    Fetch<T> fx, fy;
    fetch(fx, pl(jx, q), h);
    fetch(fy, pl(jy, q), h);
    commit(sjx[q % 3], fx);
    commit(sjy[q % 3], fy);
  }
  // Node k needs J_x, J_y of plane k+1 and J_z of plane k; the loads of node
  // k+2 are issued while node k is computed (two register sets).  Every step
  // loads (planes clamped to the grid) so the waits before each commit are
  // exact.
  auto fetch_set = [&](SourceSet<T>& q, int k) {
    const int kx = min(k + 1, g.nz - 1), kz = min(k, g.nz - 1);
    fetch(q.x, pl(jx, kx), h);
    fetch(q.y, pl(jy, kx), h);
    fetch(q.z, pl(jzp, kz), h);
  };
  auto step = [&](SourceSet<T>& q, int k) {
    const bool act = k < k1;
    // ring slots of planes k, k+1 and k-1
    const int r0 = k % 3, rp = r0 == 2 ? 0 : r0 + 1, rm = r0 == 0 ? 2 : r0 - 1;
    if (act) {
      if (k + 1 <= g.nz - 1) {
        commit(sjx[rp], q.x);
        commit(sjy[rp], q.y);
      }
      commit(sjz, q.z);
    }
    lds_barrier();
    fetch_set(q, k + 2);
    if (act && t.valid) {
      const T s = lds_dy(sjx[r0], t, g) - lds_dx(sjy[r0], t, g);
      st(pl(sb, k), h.own, s);
      acc1 += s * s;
      if (interior) {
        T v = lds_lap(sjz, t, g);
        if (k >= 1 && k <= g.nz - 2) {
          const T* xu = sjx[rp] + t.c;
          const T* xd = sjx[rm] + t.c;
          const T dxz = (xu[1] - xd[1] - xu[-1] + xd[-1]) * i4dxdz;
          const T* yu = sjy[rp] + t.c;
          const T* yd = sjy[rm] + t.c;
          const T dyz = (yu[PW] - yd[PW] - yu[-PW] + yd[-PW]) * i4dydz;
          v = v - dxz - dyz;
        }
        st(sj + (long)k * ip, sjo, v);
      }
    }
    lds_barrier();
  };
  SourceSet<T> sa, sb2;
  fetch_set(sa, k0);
  fetch_set(sb2, k0 + 1);
  for (int k = k0; k < k1; k += 2) {
    step(sa, k);
    step(sb2, k + 1);
  }
  if (!partials) return;
  double acc[1] = {(double)acc1};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) {
    const long bid = t.id;
    partials[NFGS_Q_SRC2 * pstride + bid] = acc[0];
  }
}

// Monitor sums of the per-block partials.  The first kRedBatch partials of a
// thread per quantity are loaded together (one memory round trip for grids up
// to kRedBatch * 1024 blocks); the sum order is fixed.
constexpr int kRedBatch = 4;
static_assert(kRedBatch == 4, "the batch is summed pairwise below");
// Sums slot q's partials of the `active` blocks of the last launch (stride
// pstride between slots; the zero-mode slot holds one value).
__global__ __launch_bounds__(1024) void reduce_kernel(const double* __restrict__ partials,
                                                      long pstride, long active,
                                                      double* __restrict__ metrics) {
  __shared__ double red[NFGS_NQ * 16];
  double acc[NFGS_NQ];
  const long nt = blockDim.x;
#pragma unroll
  for (int q = 0; q < NFGS_NQ; ++q) {
    const long cnt = q == NFGS_Q_ZERO_MODE ? 1 : active;
    const double* src = partials + q * pstride;
    double v[kRedBatch];
#pragma unroll
    for (int i = 0; i < kRedBatch; ++i) {
      const long b = threadIdx.x + i * nt;
      v[i] = b < cnt ? src[b] : 0.0;
    }
    double s = (v[0] + v[1]) + (v[2] + v[3]);
    for (long b = threadIdx.x + kRedBatch * nt; b < cnt; b += nt) s += src[b];
    acc[q] = s;
  }
  block_sum<NFGS_NQ>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NFGS_NQ; ++q) metrics[q] = acc[q];
  }
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, long n, T v) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

dim3 tiled_grid(const nfgs_plan* p) {
  return dim3((unsigned)shape::blocks(p->d.nx, p->d.ny, p->d.nz, p->zchunks, cu_count_stencil()));
}

// A launch that writes monitor partials must fit the plan's partials (one
// value per block and slot); nfgs_plan_set_tuning admits only chunkings that
// do, so this fails only on a plan whose layout was changed behind it.
int check_partials(const nfgs_plan* p, const double* partials) {
  if (partials && (long)tiled_grid(p).x > p->pblocks) {
    set_error("tiled launch has more blocks than the plan's partials hold");
    return NFGS_ERR_STATE;
  }
  return NFGS_OK;
}

template <typename T>
int reconstruct_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi, const void* az,
                  void* b, hipStream_t s) {
  const Grid<T> g = make_grid<T>(p);
  hipLaunchKernelGGL((reconstruct_b_kernel<T>), tiled_grid(p), dim3(TX * TY), 0, s, g, static_cast<const T*>(phi), static_cast<const T*>(az),
                     static_cast<const T*>(bnd->b_bg), static_cast<const T*>(bnd->bz_obs),
                     static_cast<const T*>(bnd->h_profile), static_cast<T*>(b));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

template <typename T>
int face_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b, const void* az,
           const void* jc, void* jp, void* jout, double* partials, hipStream_t s) {
  if (const int rc = check_partials(p, partials)) return rc;
  const Grid<T> g = make_grid<T>(p);
  const Window w{p->d.win_z0, p->d.win_z1, p->d.win_y0, p->d.win_y1, p->d.win_x0, p->d.win_x1};
  hipLaunchKernelGGL((face_pass_kernel<T>), tiled_grid(p), dim3(TX * TY), 0, s, g,
                     static_cast<const T*>(b), static_cast<const T*>(bnd->b_bg),
                     static_cast<const T*>(az), static_cast<const T*>(jc),
                     static_cast<const T*>(bnd->jz_t),
                     static_cast<T*>(jp), static_cast<T*>(jout), partials, (long)p->pblocks,
                     T(p->d.weak_field_floor_b2), T(p->d.alpha_max), w);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

template <typename T>
int state_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi, const void* az,
            void* b, void* jp, void* jout, double* partials, const void* jzc, const void* jh,
            hipStream_t s) {
  if (const int rc = check_partials(p, partials)) return rc;
  const Grid<T> g = make_grid<T>(p);
  const Window w{p->d.win_z0, p->d.win_z1, p->d.win_y0, p->d.win_y1, p->d.win_x0, p->d.win_x1};
  auto kern = (jp && !jout && partials) ? state_pass_kernel<T, true> : state_pass_kernel<T, false>;
  hipLaunchKernelGGL(kern, tiled_grid(p), dim3(TX * TY), 0, s, g,
                     static_cast<const T*>(phi), static_cast<const T*>(az),
                     static_cast<const T*>(bnd->b_bg), static_cast<const T*>(bnd->bz_obs),
                     static_cast<const T*>(bnd->h_profile), static_cast<const T*>(bnd->jz_t),
                     static_cast<const T*>(jzc), static_cast<const T*>(jh),
                     static_cast<T*>(b), static_cast<T*>(jp), static_cast<T*>(jout), partials,
                     (long)p->pblocks, T(p->d.weak_field_floor_b2), T(p->d.alpha_max), w);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

template <typename T>
int sources_t(const nfgs_plan* p, const void* jp, void* sb, void* sj, double* partials,
              hipStream_t s) {
  if (const int rc = check_partials(p, partials)) return rc;
  const Grid<T> g = make_grid<T>(p);
  hipLaunchKernelGGL((sources_kernel<T>), tiled_grid(p), dim3(TX * TY), 0, s, g,
                     static_cast<const T*>(jp), static_cast<T*>(sb), static_cast<T*>(sj),
                     partials, (long)p->pblocks, T(1.0 / (4.0 * p->d.dx * p->d.dz)),
                     T(1.0 / (4.0 * p->d.dy * p->d.dz)));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace

int cu_count_stencil() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

long tiled_blocks(const nfgs_plan* p) {
  const dim3 d = tiled_grid(p);
  return (long)d.x * d.y * d.z;
}

int fill_zero(const nfgs_plan* p, void* ptr, long n, hipStream_t s) {
  if (n <= 0) return NFGS_OK;
  const long nb = (n + 255) / 256;
  const unsigned grid = (unsigned)(nb < 4096 ? nb : 4096);
  if (p->d.dtype == NFGS_F32)
    hipLaunchKernelGGL((fill_kernel<float>), dim3(grid), dim3(256), 0, s,
                       static_cast<float*>(ptr), n, 0.0f);
  else
    hipLaunchKernelGGL((fill_kernel<double>), dim3(grid), dim3(256), 0, s,
                       static_cast<double*>(ptr), n, 0.0);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int fill_zero_f64(double* ptr, long n, hipStream_t s) {
  if (n <= 0) return NFGS_OK;
  const long nb = (n + 255) / 256;
  hipLaunchKernelGGL((fill_kernel<double>), dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0,
                     s, ptr, n, 0.0);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int state_pass_closure(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi,
                       const void* az, void* b, void* jp, void* jout, double* partials,
                       const void* jzc, const void* jh, hipStream_t s) {
  if (p->d.dtype == NFGS_F32) return state_t<float>(p, bnd, phi, az, b, jp, jout, partials, jzc, jh, s);
  return state_t<double>(p, bnd, phi, az, b, jp, jout, partials, jzc, jh, s);
}

}  // namespace nfgs

using namespace nfgs;

extern "C" int nfgs_reconstruct_b(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi,
                                  const void* az, void* b_out, void* stream) {
  NFGS_REQUIRE(p && bnd && phi && az && b_out, NFGS_ERR_ARG, "reconstruct_b: null argument");
  NFGS_REQUIRE(bnd->b_bg && bnd->bz_obs && bnd->h_profile, NFGS_ERR_STATE,
               "reconstruct_b: boundary is not set");
  if (p->d.dtype == NFGS_F32) return reconstruct_t<float>(p, bnd, phi, az, b_out, as_stream(stream));
  return reconstruct_t<double>(p, bnd, phi, az, b_out, as_stream(stream));
}

extern "C" int nfgs_face_pass(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b,
                              const void* az, void* jp_out, void* j_out, double* partials,
                              void* stream) {
  NFGS_REQUIRE(p && bnd && b && az, NFGS_ERR_ARG, "face_pass: null argument");
  NFGS_REQUIRE(bnd->b_bg && bnd->jz_t, NFGS_ERR_STATE, "face_pass: boundary is not set");
  if (p->d.dtype == NFGS_F32)
    return face_t<float>(p, bnd, b, az, nullptr, jp_out, j_out, partials, as_stream(stream));
  return face_t<double>(p, bnd, b, az, nullptr, jp_out, j_out, partials, as_stream(stream));
}

extern "C" int nfgs_face_pass_j(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b,
                                const void* j, void* jp_out, double* partials, void* stream) {
  NFGS_REQUIRE(p && bnd && b && j && (jp_out || partials), NFGS_ERR_ARG,
               "face_pass_j: null argument");
  NFGS_REQUIRE(bnd->b_bg && bnd->jz_t, NFGS_ERR_STATE, "face_pass_j: boundary is not set");
  if (p->d.dtype == NFGS_F32)
    return face_t<float>(p, bnd, b, nullptr, j, jp_out, nullptr, partials, as_stream(stream));
  return face_t<double>(p, bnd, b, nullptr, j, jp_out, nullptr, partials, as_stream(stream));
}

extern "C" int nfgs_state_pass(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi,
                               const void* az, void* b_out, void* jp_out, void* j_out,
                               double* partials, void* stream) {
  NFGS_REQUIRE(p && bnd && phi && az, NFGS_ERR_ARG, "state_pass: null argument");
  NFGS_REQUIRE(p->d.nz >= 3, NFGS_ERR_ARG, "state_pass: nz must be >= 3");
  NFGS_REQUIRE(bnd->b_bg && bnd->bz_obs && bnd->h_profile && bnd->jz_t, NFGS_ERR_STATE,
               "state_pass: boundary is not set");
  return state_pass_closure(p, bnd, phi, az, b_out, jp_out, j_out, partials, nullptr, nullptr,
                            as_stream(stream));
}

extern "C" int nfgs_sources(const nfgs_plan* p, const void* jp, void* s_bcz, void* s_jz,
                            double* partials, void* stream) {
  NFGS_REQUIRE(p && jp && s_bcz && s_jz, NFGS_ERR_ARG, "sources: null argument");
  if (p->d.dtype == NFGS_F32) return sources_t<float>(p, jp, s_bcz, s_jz, partials, as_stream(stream));
  return sources_t<double>(p, jp, s_bcz, s_jz, partials, as_stream(stream));
}

extern "C" int nfgs_reduce(const nfgs_plan* p, const double* partials, double* metrics,
                           void* stream) {
  NFGS_REQUIRE(p && partials && metrics, NFGS_ERR_ARG, "reduce: null argument");
  if (const int rc = check_partials(p, partials)) return rc;
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, as_stream(stream), partials,
                     (long)p->pblocks, tiled_blocks(p), metrics);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}
