// Point-wise stencil kernels of one NFGS iteration and the metric reduction.
//
// Layout: (component, z, y, x), x fastest; one lane per node, consecutive
// lanes on consecutive x so that every load is a contiguous run.  Neighbour
// samples are re-read through L1/L2.  Reductions keep FP64 per-block
// partials (fixed grid, fixed order) so every run sums identically.

#include "plan.h"

namespace nfgs {
namespace {

constexpr int TX = 32, TY = 8, HH = 2;
constexpr int PW = TX + 2 * HH, PH = TY + 2 * HH, PS = PW * PH;
constexpr int KCHUNK = 32;

template <typename T>
struct Grid {
  int nx, ny, nz;
  long plane, N;
  T dx, dy, dz;
  T idx2, idy2;
  int tgx, tgy;  // tiles along x and y (the launch is 1-D over tgx*tgy*z-chunks)
};

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
  g.tgx = (g.nx + TX - 1) / TX;
  g.tgy = (g.ny + TY - 1) / TY;
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
  int c;  // LDS index of this thread's point
};

template <typename T>
__device__ __forceinline__ Tile make_tile(const Grid<T>& g) {
  Tile t;
  t.id = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = t.id % g.tgx, r = t.id / g.tgx;
  t.x0 = bx * TX;
  t.y0 = (r % g.tgy) * TY;
  t.k0 = (r / g.tgy) * KCHUNK;
  t.tx = threadIdx.x % TX;
  t.ty = threadIdx.x / TX;
  t.x = t.x0 + t.tx;
  t.y = t.y0 + t.ty;
  t.valid = t.x < g.nx && t.y < g.ny;
  t.c = (t.ty + HH) * PW + t.tx + HH;
  return t;
}

// in-domain offset of a thread's own point (clamped for threads past the edge)
template <typename T>
__device__ __forceinline__ long own_offset(const Tile& t, const Grid<T>& g) {
  const int x = t.x < g.nx ? t.x : g.nx - 1, y = t.y < g.ny ? t.y : g.ny - 1;
  return (long)y * g.nx + x;
}

// Two-step staging of a plane region (tile + halo): fetch() issues the
// global loads into registers, commit() writes them to an LDS slot later, so
// the loads of plane p+1 are in flight while plane p is being computed.
constexpr int PPT = (PS + 255) / 256;  // staged elements per thread (256 threads)

template <typename T>
struct Fetch {
  T v[PPT];
};

template <typename T>
__device__ __forceinline__ void fetch(Fetch<T>& r, const T* plane, const Tile& t,
                                      const Grid<T>& g) {
  // Unconditional loads from clamped coordinates: halo points outside the
  // domain get a neighbour's value, which no valid stencil reads.
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    int e = threadIdx.x + 256 * i;
    e = e < PS ? e : PS - 1;
    const int ly = e / PW, lx = e - ly * PW;
    int y = t.y0 + ly - HH, x = t.x0 + lx - HH;
    y = y < 0 ? 0 : (y >= g.ny ? g.ny - 1 : y);
    x = x < 0 ? 0 : (x >= g.nx ? g.nx - 1 : x);
    r.v[i] = plane[(long)y * g.nx + x];
  }
}

template <typename T>
__device__ __forceinline__ void commit(T* slot, const Fetch<T>& r) {
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    const int e = threadIdx.x + 256 * i;
    if (e < PS) slot[e] = r.v[i];
  }
}

template <typename T>
__device__ __forceinline__ void commit_diff(T* slot, const Fetch<T>& a, const Fetch<T>& b) {
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    const int e = threadIdx.x + 256 * i;
    if (e < PS) slot[e] = a.v[i] - b.v[i];
  }
}

template <typename T>
__device__ __forceinline__ T lds_dx(const T* s, const Tile& t, const Grid<T>& g) {
  return fd1(s + t.c, 1, t.x, g.nx, g.dx);
}
template <typename T>
__device__ __forceinline__ T lds_dy(const T* s, const Tile& t, const Grid<T>& g) {
  return fd1(s + t.c, (long)PW, t.y, g.ny, g.dy);
}
template <typename T>
__device__ __forceinline__ T lds_lap(const T* s, const Tile& t, const Grid<T>& g) {
  return lap_xy_at(s + t.c - ((long)t.y * PW + t.x), t.y, t.x, g.ny, g.nx, (long)PW, g.idx2,
                   g.idy2);
}

// Reconstruction of B from the scalars (nfgs_solver.py:1711-1715, :3311-3341,
// :5901-5916, :6069-6078).
template <typename T>
__global__ __launch_bounds__(256) void reconstruct_b_kernel(
    Grid<T> g, const T* __restrict__ phi, const T* __restrict__ az, const T* __restrict__ bg,
    const T* __restrict__ bz_obs, const T* __restrict__ hprof, T* __restrict__ b) {
  __shared__ T sphi[3][PS];
  __shared__ T saz[PS];
  __shared__ T sdz[PS];
  const Tile t = make_tile(g);
  const int k0 = t.k0, k1 = min(k0 + KCHUNK, g.nz);
  // Phi planes lo..lo+2 are resident while node k is built
  auto lo_of = [&](int k) { return max(0, min(k - 1, g.nz - 3)); };
  const long own = own_offset(t, g);
  int lo = lo_of(k0);
  {
    Fetch<T> f0, f1, f2;
    fetch(f0, phi + (long)lo * g.plane, t, g);
    fetch(f1, phi + (long)(lo + 1) * g.plane, t, g);
    fetch(f2, phi + (long)(lo + 2) * g.plane, t, g);
    commit(sphi[lo % 3], f0);
    commit(sphi[(lo + 1) % 3], f1);
    commit(sphi[(lo + 2) % 3], f2);
  }
  Fetch<T> fphi, faz;
  fetch(faz, az + (long)k0 * g.plane, t, g);
  bool have_phi = false;
  T gx, gy, gz;
  {
    const long i = (long)k0 * g.plane + own;
    gx = bg[i];
    gy = bg[g.N + i];
    gz = bg[2 * g.N + i];
  }
  for (int k = k0; k < k1; ++k) {
    if (have_phi) {
      commit(sphi[(lo + 2) % 3], fphi);
      have_phi = false;
    }
    commit(saz, faz);
    const T cgx = gx, cgy = gy, cgz = gz;
    __syncthreads();
    // prefetch what node k+1 needs
    if (k + 1 < k1) {
      const int nlo = lo_of(k + 1);
      if (nlo > lo) {
        fetch(fphi, phi + (long)(nlo + 2) * g.plane, t, g);
        have_phi = true;
        lo = nlo;
      }
      fetch(faz, az + (long)(k + 1) * g.plane, t, g);
      const long i = (long)(k + 1) * g.plane + own;
      gx = bg[i];
      gy = bg[g.N + i];
      gz = bg[2 * g.N + i];
    }
    // dPhi/dz on the tile with halo
    for (int e = threadIdx.x; e < PS; e += blockDim.x) {
      T d;
      if (k == 0)
        d = (T(-3) * sphi[0][e] + T(4) * sphi[1][e] - sphi[2][e]) / (T(2) * g.dz);
      else if (k == g.nz - 1)
        d = (T(3) * sphi[k % 3][e] - T(4) * sphi[(k - 1) % 3][e] + sphi[(k - 2) % 3][e]) /
            (T(2) * g.dz);
      else
        d = (sphi[(k + 1) % 3][e] - sphi[(k - 1) % 3][e]) / (T(2) * g.dz);
      sdz[e] = d;
    }
    __syncthreads();
    if (t.valid) {
      const long i = (long)k * g.plane + own;
      const T bcx = -lds_dx(sdz, t, g) + lds_dy(saz, t, g);
      const T bcy = -lds_dy(sdz, t, g) - lds_dx(saz, t, g);
      const T bcz = lds_lap(sphi[k % 3], t, g);
      T bx = cgx + bcx;
      T by = cgy + bcy;
      T bz = cgz + bcz;
      if (k == 0) bz = bz_obs[own];
      b[i] = bx + hprof[k];
      b[g.N + i] = by + hprof[g.nz + k];
      b[2 * g.N + i] = bz;
    }
    __syncthreads();
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

template <typename T>
__global__ __launch_bounds__(256) void face_pass_kernel(
    Grid<T> g, const T* __restrict__ B, const T* __restrict__ bg, const T* __restrict__ az,
    const T* __restrict__ jzt, T* __restrict__ jp, T* __restrict__ jout, double* partials,
    long pstride, T floor_b2, T amax, Window w) {
  __shared__ T scz[2][PS];
  __shared__ T sbx[2][PS];
  __shared__ T sby[2][PS];
  __shared__ T saz[PS];
  __shared__ double red[NFGS_NQ * 16];
  const Tile t = make_tile(g);
  const int k0 = t.k0, k1 = min(k0 + KCHUNK, g.nz);
  const int ps = max(k0 - 1, 0), pe = min(k1, g.nz - 1);
  const long off = own_offset(t, g);
  const bool edge = t.y == 0 || t.y == g.ny - 1 || t.x == 0 || t.x == g.nx - 1;
  const bool in_win = t.y >= w.y0 && t.y < w.y1 && t.x >= w.x0 && t.x < w.x1;
  const T* Bx = B;
  const T* By = B + g.N;
  const T* Bz = B + 2 * g.N;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  T pb[3] = {0, 0, 0}, pg[2] = {0, 0}, pjz = 0;  // previous plane, own point
  FaceVal<T> fprev;
  Fetch<T> fbz, fgz, fbx, fby, faz;
  T obz = 0, ogx = 0, ogy = 0;  // own Bz, bg_x, bg_y of the fetched plane
  auto fetch_plane = [&](int q) {
    const long qo = (long)q * g.plane;
    fetch(fbz, Bz + qo, t, g);
    fetch(fgz, bg + 2 * g.N + qo, t, g);
    fetch(fbx, Bx + qo, t, g);
    fetch(fby, By + qo, t, g);
    fetch(faz, az + qo, t, g);
    obz = Bz[qo + off];
    ogx = bg[qo + off];
    ogy = bg[g.N + qo + off];
  };
  fetch_plane(ps);
  for (int p = ps; p <= pe; ++p) {
    commit_diff(scz[p & 1], fbz, fgz);
    commit(sbx[p & 1], fbx);
    commit(sby[p & 1], fby);
    commit(saz, faz);
    const T cbz = obz, cgx = ogx, cgy = ogy;
    __syncthreads();
    if (p + 1 <= pe) fetch_plane(p + 1);
    T cb[3] = {0, 0, 0}, cg[2] = {0, 0}, cjz = 0;
    if (t.valid) {
      cb[0] = sbx[p & 1][t.c];
      cb[1] = sby[p & 1][t.c];
      cb[2] = cbz;
      cg[0] = cgx;
      cg[1] = cgy;
      if (p == 0)
        cjz = jzt[off];
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
      const T dbx = ((cb[0] - cg[0]) - (pb[0] - pg[0])) / g.dz;
      const T dby = ((cb[1] - cg[1]) - (pb[1] - pg[1])) / g.dz;
      const T* s1 = scz[p & 1];
      const T* s0 = scz[f & 1];
      auto cfz = [&](int o) { return T(0.5) * (s1[t.c + o] + s0[t.c + o]); };
      const T dyc = fd1g<T>([&](int o) { return cfz(o * PW); }, t.y, g.ny, g.dy);
      const T dxc = fd1g<T>([&](int o) { return cfz(o); }, t.x, g.nx, g.dx);
      F.j[0] = dyc - dby;
      F.j[1] = dbx - dxc;
      F.j[2] = T(0.5) * (cjz + pjz);
      {
        T b2 = F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2];
        b2 = b2 < floor_b2 ? floor_b2 : b2;
        const T jb = F.j[0] * F.b[0] + F.j[1] * F.b[1] + F.j[2] * F.b[2];
        T a = jb / b2;
        a = a < -amax ? -amax : (a > amax ? amax : a);
        F.p[0] = a * F.b[0];
        F.p[1] = a * F.b[1];
        F.p[2] = a * F.b[2];
      }
      const int k = f;  // node finalised now
      if (k >= k0 && k < k1) {
        const long i = (long)k * g.plane + off;
        if (jp) {
          if (k == 0) {
            jp[i] = edge ? T(0) : F.j[0];
            jp[g.N + i] = edge ? T(0) : F.j[1];
            jp[2 * g.N + i] = jzt[off];
          } else {
            jp[i] = T(0.5) * (fprev.p[0] + F.p[0]);
            jp[g.N + i] = T(0.5) * (fprev.p[1] + F.p[1]);
            jp[2 * g.N + i] = T(0.5) * (fprev.p[2] + F.p[2]);
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
          acc[6] += (double)c[0] * c[0] + (double)c[1] * c[1] + (double)c[2] * c[2];
          acc[5] += (double)F.j[0] * F.j[0] + (double)F.j[1] * F.j[1] + (double)F.j[2] * F.j[2];
          jout[i] = edge ? T(0) : c[0];
          jout[g.N + i] = edge ? T(0) : c[1];
          jout[2 * g.N + i] = pjz;
        }
        if (partials) {
          if (k >= w.z0 && k < w.z1 && in_win) {
            T bn = sqrt(F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2]);
            bn = bn < T(1e-10) ? T(1e-10) : bn;
            const T jn = sqrt(F.j[0] * F.j[0] + F.j[1] * F.j[1] + F.j[2] * F.j[2]);
            const T c0 = F.j[1] * F.b[2] - F.j[2] * F.b[1];
            const T c1 = F.j[2] * F.b[0] - F.j[0] * F.b[2];
            const T c2 = F.j[0] * F.b[1] - F.j[1] * F.b[0];
            acc[0] += (double)(sqrt(c0 * c0 + c1 * c1 + c2 * c2) / bn);
            acc[1] += (double)jn;
          }
          const T dvx = lds_dx(sbx[k & 1], t, g);
          const T dvy = lds_dy(sby[k & 1], t, g);
          const T dvz = fd1(Bz + i, g.plane, k, g.nz, g.dz);
          const T div = dvx + dvy + dvz;
          const T b2 = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
          acc[2] += (double)(div < T(0) ? -div : div);
          acc[3] += (double)sqrt(b2);
          acc[4] += (double)b2;
        }
      }
      fprev = F;
    }
    pb[0] = cb[0]; pb[1] = cb[1]; pb[2] = cb[2];
    pg[0] = cg[0]; pg[1] = cg[1];
    pjz = cjz;
    __syncthreads();
  }
  // top node: no face above it
  const int kt = g.nz - 1;
  if (k1 == g.nz && t.valid) {
    const long i = (long)kt * g.plane + off;
    if (jp) {
      jp[i] = fprev.p[0];
      jp[g.N + i] = fprev.p[1];
      jp[2 * g.N + i] = fprev.p[2];
    }
    if (jout) {
      acc[6] += (double)fprev.j[0] * fprev.j[0] + (double)fprev.j[1] * fprev.j[1] +
                (double)fprev.j[2] * fprev.j[2];
      jout[i] = T(0);
      jout[g.N + i] = T(0);
      jout[2 * g.N + i] = T(0);
    }
    if (partials) {
      const T dvx = lds_dx(sbx[kt & 1], t, g);
      const T dvy = lds_dy(sby[kt & 1], t, g);
      const T dvz = fd1(Bz + i, g.plane, kt, g.nz, g.dz);
      const T div = dvx + dvy + dvz;
      const T b2 = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
      acc[2] += (double)(div < T(0) ? -div : div);
      acc[3] += (double)sqrt(b2);
      acc[4] += (double)b2;
    }
  }
  if (!partials) return;
  block_sum<7>(acc, red);
  if (threadIdx.x == 0) {
    const long bid = t.id;
    partials[NFGS_Q_SIGMA_NUM * pstride + bid] = acc[0];
    partials[NFGS_Q_SIGMA_DEN * pstride + bid] = acc[1];
    partials[NFGS_Q_DIV_ABS * pstride + bid] = acc[2];
    partials[NFGS_Q_BMAG * pstride + bid] = acc[3];
    partials[NFGS_Q_B2 * pstride + bid] = acc[4];
    partials[NFGS_Q_FACE_J2 * pstride + bid] = acc[5];
    partials[NFGS_Q_CELL_J2 * pstride + bid] = acc[6];
  }
}

// Sources of the two Poisson problems (nfgs_solver.py:6726-6745, :1729-1749).
template <typename T>
__global__ __launch_bounds__(256) void sources_kernel(Grid<T> g, const T* __restrict__ jp,
                                                      T* __restrict__ sb, T* __restrict__ sj,
                                                      double* partials, long pstride, T i4dxdz,
                                                      T i4dydz) {
  __shared__ T sjx[3][PS];
  __shared__ T sjy[3][PS];
  __shared__ T sjz[PS];
  __shared__ double red[NFGS_NQ * 16];
  const Tile t = make_tile(g);
  const int k0 = t.k0, k1 = min(k0 + KCHUNK, g.nz);
  const T* jx = jp;
  const T* jy = jp + g.N;
  const T* jzp = jp + 2 * g.N;
  const int nxi = g.nx - 2;
  const long ip = (long)(g.ny - 2) * nxi;
  const bool interior = t.y >= 1 && t.y <= g.ny - 2 && t.x >= 1 && t.x <= g.nx - 2;
  double acc[1] = {0.0};
  Fetch<T> fx, fy, fz;
  for (int q = max(k0 - 1, 0); q <= k0; ++q) {
    fetch(fx, jx + (long)q * g.plane, t, g);
    fetch(fy, jy + (long)q * g.plane, t, g);
    commit(sjx[q % 3], fx);
    commit(sjy[q % 3], fy);
  }
  if (k0 + 1 <= g.nz - 1) {
    fetch(fx, jx + (long)(k0 + 1) * g.plane, t, g);
    fetch(fy, jy + (long)(k0 + 1) * g.plane, t, g);
  }
  fetch(fz, jzp + (long)k0 * g.plane, t, g);
  for (int k = k0; k < k1; ++k) {
    if (k + 1 <= g.nz - 1) {
      commit(sjx[(k + 1) % 3], fx);
      commit(sjy[(k + 1) % 3], fy);
    }
    commit(sjz, fz);
    __syncthreads();
    if (k + 1 < k1) {
      if (k + 2 <= g.nz - 1) {
        fetch(fx, jx + (long)(k + 2) * g.plane, t, g);
        fetch(fy, jy + (long)(k + 2) * g.plane, t, g);
      }
      fetch(fz, jzp + (long)(k + 1) * g.plane, t, g);
    }
    if (t.valid) {
      const long i = (long)k * g.plane + (long)t.y * g.nx + t.x;
      const T s = lds_dy(sjx[k % 3], t, g) - lds_dx(sjy[k % 3], t, g);
      sb[i] = s;
      acc[0] += (double)s * s;
      if (interior) {
        T v = lds_lap(sjz, t, g);
        if (k >= 1 && k <= g.nz - 2) {
          const T* xu = sjx[(k + 1) % 3] + t.c;
          const T* xd = sjx[(k - 1) % 3] + t.c;
          const T dxz = (xu[1] - xd[1] - xu[-1] + xd[-1]) * i4dxdz;
          const T* yu = sjy[(k + 1) % 3] + t.c;
          const T* yd = sjy[(k - 1) % 3] + t.c;
          const T dyz = (yu[PW] - yd[PW] - yu[-PW] + yd[-PW]) * i4dydz;
          v = v - dxz - dyz;
        }
        sj[(long)k * ip + (long)(t.y - 1) * nxi + (t.x - 1)] = v;
      }
    }
    __syncthreads();
  }
  if (!partials) return;
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) {
    const long bid = t.id;
    partials[NFGS_Q_SRC2 * pstride + bid] = acc[0];
  }
}

__global__ void reduce_kernel(const double* __restrict__ partials, long pstride,
                              double* __restrict__ metrics) {
  __shared__ double red[NFGS_NQ * 16];
  double acc[NFGS_NQ];
#pragma unroll
  for (int q = 0; q < NFGS_NQ; ++q) {
    const long cnt = q == NFGS_Q_ZERO_MODE ? 1 : pstride;
    double s = 0.0;
    for (long b = threadIdx.x; b < cnt; b += blockDim.x) s += partials[q * pstride + b];
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
  return dim3((unsigned)(((p->d.nx + TX - 1) / TX) * ((p->d.ny + TY - 1) / TY) *
                         ((p->d.nz + KCHUNK - 1) / KCHUNK)));
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
int face_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b, const void* az, void* jp,
           void* jout, double* partials, hipStream_t s) {
  const Grid<T> g = make_grid<T>(p);
  const Window w{p->d.win_z0, p->d.win_z1, p->d.win_y0, p->d.win_y1, p->d.win_x0, p->d.win_x1};
  hipLaunchKernelGGL((face_pass_kernel<T>), tiled_grid(p), dim3(TX * TY), 0, s, g,
                     static_cast<const T*>(b), static_cast<const T*>(bnd->b_bg),
                     static_cast<const T*>(az), static_cast<const T*>(bnd->jz_t),
                     static_cast<T*>(jp), static_cast<T*>(jout), partials, (long)p->pblocks,
                     T(p->d.weak_field_floor_b2), T(p->d.alpha_max), w);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

template <typename T>
int sources_t(const nfgs_plan* p, const void* jp, void* sb, void* sj, double* partials,
              hipStream_t s) {
  const Grid<T> g = make_grid<T>(p);
  hipLaunchKernelGGL((sources_kernel<T>), tiled_grid(p), dim3(TX * TY), 0, s, g,
                     static_cast<const T*>(jp), static_cast<T*>(sb), static_cast<T*>(sj),
                     partials, (long)p->pblocks, T(1.0 / (4.0 * p->d.dx * p->d.dz)),
                     T(1.0 / (4.0 * p->d.dy * p->d.dz)));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace

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
    return face_t<float>(p, bnd, b, az, jp_out, j_out, partials, as_stream(stream));
  return face_t<double>(p, bnd, b, az, jp_out, j_out, partials, as_stream(stream));
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
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, as_stream(stream), partials,
                     (long)p->pblocks, metrics);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}
