// Point-wise stencil kernels of one NFGS iteration and the metric reduction.
//
// Layout: (component, z, y, x), x fastest; one lane per node, consecutive
// lanes on consecutive x so that every load is a contiguous run.  Neighbour
// samples are re-read through L1/L2.  Reductions keep FP64 per-block
// partials (fixed grid, fixed order) so every run sums identically.

#include "plan.h"

namespace nfgs {
namespace {

template <typename T>
struct Grid {
  int nx, ny, nz;
  long plane, N;
  T dx, dy, dz;
  T idx2, idy2;
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
  return g;
}

// ---------------------------------------------------------------------------
// Reconstruction of B from the scalars (nfgs_solver.py:1711-1715, :3311-3341,
// :5901-5916, :6069-6078).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T dz_at(const T* f, const Grid<T>& g, int k, long off) {
  return fd1(f + (long)k * g.plane + off, g.plane, k, g.nz, g.dz);
}

template <typename T>
__global__ __launch_bounds__(256) void reconstruct_b_kernel(
    Grid<T> g, const T* __restrict__ phi, const T* __restrict__ az, const T* __restrict__ bg,
    const T* __restrict__ bz_obs, const T* __restrict__ hprof, T* __restrict__ b) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += stride) {
    const int k = (int)(i / g.plane);
    const long off = i - (long)k * g.plane;
    const int y = (int)(off / g.nx), x = (int)(off - (long)y * g.nx);
    const T* azk = az + (long)k * g.plane;
    // d/dx and d/dy of dPhi/dz, each built from the z-derivative at neighbours
    auto dzx = [&](int o) { return dz_at(phi, g, k, off + o); };
    auto dzy = [&](int o) { return dz_at(phi, g, k, off + (long)o * g.nx); };
    const T ddx = fd1g<T>(dzx, x, g.nx, g.dx);
    const T ddy = fd1g<T>(dzy, y, g.ny, g.dy);
    const T azx = fd1(azk + off, 1, x, g.nx, g.dx);
    const T azy = fd1(azk + off, g.nx, y, g.ny, g.dy);
    const T bcx = -ddx + azy;
    const T bcy = -ddy - azx;
    const T bcz = lap_xy_at(phi + (long)k * g.plane, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
    T bx = bg[i] + bcx;
    T by = bg[g.N + i] + bcy;
    T bz = bg[2 * g.N + i] + bcz;
    if (k == 0) bz = bz_obs[off];
    bx = bx + hprof[k];
    by = by + hprof[g.nz + k];
    b[i] = bx;
    b[g.N + i] = by;
    b[2 * g.N + i] = bz;
  }
}

// ---------------------------------------------------------------------------
// Vertical-face state (nfgs_solver.py:5738-5778) and everything built on it.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T jz_at(const T* az, const T* jzt, const Grid<T>& g, int k, int y,
                                   int x) {
  if (k == 0) return jzt[(long)y * g.nx + x];
  if (k == g.nz - 1 || y == 0 || y == g.ny - 1 || x == 0 || x == g.nx - 1) return T(0);
  return -lap_xy_at(az + (long)k * g.plane, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
}

template <typename T>
struct Face {
  T b[3];
  T j[3];
};

template <typename T>
__device__ __forceinline__ Face<T> face_at(const T* B, const T* bg, const T* az, const T* jzt,
                                           const Grid<T>& g, int f, int y, int x) {
  Face<T> F;
  const long off = (long)y * g.nx + x;
  const long i0 = (long)f * g.plane + off, i1 = i0 + g.plane;
  const T bx0 = B[i0], bx1 = B[i1];
  const T by0 = B[g.N + i0], by1 = B[g.N + i1];
  const T bz0 = B[2 * g.N + i0], bz1 = B[2 * g.N + i1];
  F.b[0] = T(0.5) * (bx1 + bx0);
  F.b[1] = T(0.5) * (by1 + by0);
  F.b[2] = T(0.5) * (bz1 + bz0);
  const T cx0 = bx0 - bg[i0], cx1 = bx1 - bg[i1];
  const T cy0 = by0 - bg[g.N + i0], cy1 = by1 - bg[g.N + i1];
  const T dbx = (cx1 - cx0) / g.dz;
  const T dby = (cy1 - cy0) / g.dz;
  const T* Bz = B + 2 * g.N;
  const T* Gz = bg + 2 * g.N;
  // carrier z on the face at lateral offset o (elements)
  auto cfz = [&](long o) {
    const T c1 = Bz[i1 + o] - Gz[i1 + o];
    const T c0 = Bz[i0 + o] - Gz[i0 + o];
    return T(0.5) * (c1 + c0);
  };
  const T dyc = fd1g<T>([&](int o) { return cfz((long)o * g.nx); }, y, g.ny, g.dy);
  const T dxc = fd1g<T>([&](int o) { return cfz((long)o); }, x, g.nx, g.dx);
  F.j[0] = dyc - dby;
  F.j[1] = dbx - dxc;
  F.j[2] = T(0.5) * (jz_at(az, jzt, g, f + 1, y, x) + jz_at(az, jzt, g, f, y, x));
  return F;
}

template <typename T>
__device__ __forceinline__ void par_current(const Face<T>& F, T floor_b2, T amax, T (&jp)[3]) {
  T b2 = F.b[0] * F.b[0] + F.b[1] * F.b[1] + F.b[2] * F.b[2];
  b2 = b2 < floor_b2 ? floor_b2 : b2;
  const T jb = F.j[0] * F.b[0] + F.j[1] * F.b[1] + F.j[2] * F.b[2];
  T a = jb / b2;
  a = a < -amax ? -amax : (a > amax ? amax : a);
  jp[0] = a * F.b[0];
  jp[1] = a * F.b[1];
  jp[2] = a * F.b[2];
}

struct Window {
  int z0, z1, y0, y1, x0, x1;
};

template <typename T>
__global__ __launch_bounds__(256) void face_pass_kernel(
    Grid<T> g, const T* __restrict__ B, const T* __restrict__ bg, const T* __restrict__ az,
    const T* __restrict__ jzt, T* __restrict__ jp, T* __restrict__ jout, double* partials,
    long pstride, T floor_b2, T amax, Window w) {
  __shared__ double red[NFGS_NQ * 16];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const bool want_j = jout != nullptr;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += stride) {
    const int k = (int)(i / g.plane);
    const long off = i - (long)k * g.plane;
    const int y = (int)(off / g.nx), x = (int)(off - (long)y * g.nx);
    const bool edge = (y == 0 || y == g.ny - 1 || x == 0 || x == g.nx - 1);
    Face<T> lo, hi;
    const bool has_lo = k >= 1, has_hi = k <= g.nz - 2;
    if (has_lo) lo = face_at(B, bg, az, jzt, g, k - 1, y, x);
    if (has_hi) hi = face_at(B, bg, az, jzt, g, k, y, x);
    if (jp) {
      T out[3];
      if (k == 0) {
        // bottom row of J_par is the state current J[:, 0] (nfgs_solver.py:6702-6710)
        out[0] = edge ? T(0) : hi.j[0];
        out[1] = edge ? T(0) : hi.j[1];
        out[2] = jzt[off];
      } else {
        T pl[3];
        par_current(lo, floor_b2, amax, pl);
        if (has_hi) {
          T ph[3];
          par_current(hi, floor_b2, amax, ph);
          out[0] = T(0.5) * (pl[0] + ph[0]);
          out[1] = T(0.5) * (pl[1] + ph[1]);
          out[2] = T(0.5) * (pl[2] + ph[2]);
        } else {
          out[0] = pl[0];
          out[1] = pl[1];
          out[2] = pl[2];
        }
      }
      jp[i] = out[0];
      jp[g.N + i] = out[1];
      jp[2 * g.N + i] = out[2];
    }
    if (want_j) {
      // staggered centre current with the NFGS boundary values (:5917-5941)
      T c[3];
      if (k == 0) {
        c[0] = hi.j[0]; c[1] = hi.j[1]; c[2] = hi.j[2];
      } else if (!has_hi) {
        c[0] = lo.j[0]; c[1] = lo.j[1]; c[2] = lo.j[2];
      } else {
        c[0] = T(0.5) * (lo.j[0] + hi.j[0]);
        c[1] = T(0.5) * (lo.j[1] + hi.j[1]);
        c[2] = T(0.5) * (lo.j[2] + hi.j[2]);
      }
      acc[6] += (double)c[0] * c[0] + (double)c[1] * c[1] + (double)c[2] * c[2];
      if (has_hi)
        acc[5] += (double)hi.j[0] * hi.j[0] + (double)hi.j[1] * hi.j[1] +
                  (double)hi.j[2] * hi.j[2];
      const bool zero_xy = edge || k == g.nz - 1;
      jout[i] = zero_xy ? T(0) : c[0];
      jout[g.N + i] = zero_xy ? T(0) : c[1];
      jout[2 * g.N + i] = jz_at(az, jzt, g, k, y, x);
    }
    if (partials) {
      // sigma on the face window (:5882-5895)
      if (has_hi && k >= w.z0 && k < w.z1 && y >= w.y0 && y < w.y1 && x >= w.x0 && x < w.x1) {
        const T* jb = hi.j;
        const T* bb = hi.b;
        T bn = sqrt(bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2]);
        bn = bn < T(1e-10) ? T(1e-10) : bn;
        const T jn = sqrt(jb[0] * jb[0] + jb[1] * jb[1] + jb[2] * jb[2]);
        const T c0 = jb[1] * bb[2] - jb[2] * bb[1];
        const T c1 = jb[2] * bb[0] - jb[0] * bb[2];
        const T c2 = jb[0] * bb[1] - jb[1] * bb[0];
        const T cn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        acc[0] += (double)(cn / bn);
        acc[1] += (double)jn;
      }
      // divB, |B| and |B|^2 on every node (:2689-2697)
      const T* Bx = B;
      const T* By = B + g.N;
      const T* Bz = B + 2 * g.N;
      const T dvx = fd1(Bx + i, 1, x, g.nx, g.dx);
      const T dvy = fd1(By + i, (long)g.nx, y, g.ny, g.dy);
      const T dvz = fd1(Bz + i, g.plane, k, g.nz, g.dz);
      const T div = dvx + dvy + dvz;
      const T bx = Bx[i], by = By[i], bz = Bz[i];
      const T b2 = bx * bx + by * by + bz * bz;
      acc[2] += (double)(div < T(0) ? -div : div);
      acc[3] += (double)sqrt(b2);
      acc[4] += (double)b2;
    }
  }
  if (!partials) return;
  block_sum<7>(acc, red);
  if (threadIdx.x == 0) {
    const long b = blockIdx.x;
    partials[NFGS_Q_SIGMA_NUM * pstride + b] = acc[0];
    partials[NFGS_Q_SIGMA_DEN * pstride + b] = acc[1];
    partials[NFGS_Q_DIV_ABS * pstride + b] = acc[2];
    partials[NFGS_Q_BMAG * pstride + b] = acc[3];
    partials[NFGS_Q_B2 * pstride + b] = acc[4];
    partials[NFGS_Q_FACE_J2 * pstride + b] = acc[5];
    partials[NFGS_Q_CELL_J2 * pstride + b] = acc[6];
  }
}

// ---------------------------------------------------------------------------
// Sources of the two Poisson problems (nfgs_solver.py:6726-6745, :1729-1749).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void sources_kernel(Grid<T> g, const T* __restrict__ jp,
                                                      T* __restrict__ sb, T* __restrict__ sj,
                                                      double* partials, long pstride, T i4dxdz,
                                                      T i4dydz) {
  __shared__ double red[NFGS_NQ * 16];
  double acc[1] = {0.0};
  const T* jx = jp;
  const T* jy = jp + g.N;
  const T* jzp = jp + 2 * g.N;
  const long stride = (long)gridDim.x * blockDim.x;
  const int nxi = g.nx - 2;
  const long ip = (long)(g.ny - 2) * nxi;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += stride) {
    const int k = (int)(i / g.plane);
    const long off = i - (long)k * g.plane;
    const int y = (int)(off / g.nx), x = (int)(off - (long)y * g.nx);
    const T s = fd1(jx + i, (long)g.nx, y, g.ny, g.dy) - fd1(jy + i, 1, x, g.nx, g.dx);
    sb[i] = s;
    acc[0] += (double)s * s;
    if (y >= 1 && y <= g.ny - 2 && x >= 1 && x <= g.nx - 2) {
      T v = lap_xy_at(jzp + (long)k * g.plane, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
      if (k >= 1 && k <= g.nz - 2) {
        const long up = g.plane, nx = g.nx;
        const T* fx = jx + i;
        const T dxz = (fx[up + 1] - fx[-up + 1] - fx[up - 1] + fx[-up - 1]) * i4dxdz;
        const T* fy = jy + i;
        const T dyz = (fy[up + nx] - fy[-up + nx] - fy[up - nx] + fy[-up - nx]) * i4dydz;
        v = v - dxz - dyz;
      }
      sj[(long)k * ip + (long)(y - 1) * nxi + (x - 1)] = v;
    }
  }
  if (!partials) return;
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partials[NFGS_Q_SRC2 * pstride + blockIdx.x] = acc[0];
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

template <typename T>
int reconstruct_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi, const void* az,
                  void* b, hipStream_t s) {
  const Grid<T> g = make_grid<T>(p);
  const long nb = (g.N + 255) / 256;
  hipLaunchKernelGGL((reconstruct_b_kernel<T>), dim3((unsigned)(nb < 65536 ? nb : 65536)),
                     dim3(256), 0, s, g, static_cast<const T*>(phi), static_cast<const T*>(az),
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
  hipLaunchKernelGGL((face_pass_kernel<T>), dim3(p->pblocks), dim3(256), 0, s, g,
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
  hipLaunchKernelGGL((sources_kernel<T>), dim3(p->pblocks), dim3(256), 0, s, g,
                     static_cast<const T*>(jp), static_cast<T*>(sb), static_cast<T*>(sj),
                     partials, (long)p->pblocks, T(1.0 / (4.0 * p->d.dx * p->d.dz)),
                     T(1.0 / (4.0 * p->d.dy * p->d.dz)));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace

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
