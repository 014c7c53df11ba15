// Per-iteration diagnostics (record_iteration_diagnostics=True,
// nfgs_solver.py:6660-6680 and :6806-6931) at mu = 1: FP64 sums from which
// the host forms the reference's relative RMS values.
//
// Three passes, none of them on the iteration's own path:
// * source edges: S_Jz on the lateral walls from the parallel current the
//   step starts from (the iteration's sources kernel keeps interior S_Jz
//   only, which is all the Az solve reads);
// * residuals: the map residuals of the step and the PDE residuals of its
//   new state against the step's sources, with source_phi = IDCT2(DCT2(S_Bcz)
//   / lambda) and the Az right-hand side IDST2(-DST2(S_Jz) / lambda_D) formed
//   by the plan's transforms;
// * face fractions: weak-field and alpha-clip fractions on the vertical faces
//   of a state (the next step's input) and the bottom imprint of its B.
// Every value is formed in the solver dtype, as the reference forms it, and
// squared and summed in FP64 in a fixed order.

#include "plan.h"
#include "util.h"

namespace nfgs {
namespace {

constexpr int kDiagBlocks = 1024;

template <typename T>
struct DiagGrid {
  int nx, ny, nz;
  long plane, N;
  T idx2, idy2, idz2, i4dxdz, i4dydz, dz;
};

template <typename T>
DiagGrid<T> diag_grid(const nfgs_plan* p) {
  DiagGrid<T> g;
  g.nx = p->d.nx;
  g.ny = p->d.ny;
  g.nz = p->d.nz;
  g.plane = p->plane;
  g.N = p->N;
  g.idx2 = T(1.0 / (p->d.dx * p->d.dx));
  g.idy2 = T(1.0 / (p->d.dy * p->d.dy));
  g.idz2 = T(1.0 / (p->d.dz * p->d.dz));
  g.i4dxdz = T(1.0 / (4.0 * p->d.dx * p->d.dz));
  g.i4dydz = T(1.0 / (4.0 * p->d.dy * p->d.dz));
  g.dz = T(p->d.dz);
  return g;
}

// S_Jz = lap_xy(Jp_z) - d2_dxdz(Jp_x) - d2_dydz(Jp_y) (nfgs_solver.py:6726-6745,
// :1729-1749) on the lateral walls of every plane.
template <typename T>
__global__ __launch_bounds__(256) void source_edges_kernel(DiagGrid<T> g, const T* __restrict__ jp,
                                                           T* __restrict__ sje) {
  const long per = 2L * g.nx + 2L * (g.ny - 2);
  const long n = per * g.nz;
  const T* jx = jp;
  const T* jy = jp + g.N;
  const T* jz = jp + 2 * g.N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i / per);
    const long e = i - (long)z * per;
    int y, x;
    if (e < g.nx) {
      y = 0;
      x = (int)e;
    } else if (e < 2L * g.nx) {
      y = g.ny - 1;
      x = (int)(e - g.nx);
    } else {
      const long k = e - 2L * g.nx;
      y = 1 + (int)(k >> 1);
      x = (k & 1) ? g.nx - 1 : 0;
    }
    const long o = (long)z * g.plane + (long)y * g.nx + x;
    T v = lap_xy_at(jz + (long)z * g.plane, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
    if (z >= 1 && z <= g.nz - 2) {
      const long up = g.plane, r = g.nx;
      if (x >= 1 && x <= g.nx - 2)
        v = v - (jx[o + up + 1] - jx[o - up + 1] - jx[o + up - 1] + jx[o - up - 1]) * g.i4dxdz;
      if (y >= 1 && y <= g.ny - 2)
        v = v - (jy[o + up + r] - jy[o - up + r] - jy[o + up - r] + jy[o - up - r]) * g.i4dydz;
    }
    sje[o] = v;
  }
}

// hat[q][m] /= lambda[m] with the zero mode set to 0 (source_phi), or
// hat[q][m] *= -1 / lambda_D[m] (the Az right-hand side)
template <typename T>
__global__ __launch_bounds__(256) void modal_scale_kernel(T* __restrict__ hat, const T* __restrict__ lam,
                                                          long plane, long n, int dirichlet) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long m = i % plane;
    if (dirichlet)
      hat[i] = -hat[i] / lam[m];
    else
      hat[i] = m == 0 ? T(0) : hat[i] / lam[m];
  }
}

template <typename T>
struct ResidualIn {
  const T *phi0, *az0, *phi, *az;  // old and new state
  const T *lp, *la;                // lap_xy of the new Phi_c and Az
  const T *sb, *sj, *sje;          // S_Bcz (full), S_Jz (interior, compact), S_Jz on the walls
  const T *sp, *sa;                // source_phi (full), Az right-hand side (interior, compact)
  const T* target;                 // bottom dPhi_c/dz target (plane)
  T rho;
};

template <typename T>
__device__ __forceinline__ double sq(T v) {
  const double d = (double)v;
  return d * d;
}

template <typename T>
__global__ __launch_bounds__(256) void residual_kernel(DiagGrid<T> g, ResidualIn<T> a,
                                                       double* __restrict__ partials) {
  __shared__ double red[NFGS_DR_COUNT * 16];
  double acc[NFGS_DR_COUNT];
#pragma unroll
  for (int q = 0; q < NFGS_DR_COUNT; ++q) acc[q] = 0.0;
  const long pl = g.plane;
  const int nxi = g.nx - 2;
  const long ip = (long)(g.ny - 2) * nxi;
  const int zb0 = g.nz > 4 ? 2 : 1;  // B_c,z and poloidal rows: planes zb0 .. nz-2
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < g.N;
       i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i / pl);
    const long r = i - (long)z * pl;
    const int y = (int)(r / g.nx), x = (int)(r - (long)y * g.nx);
    const bool wall = x == 0 || y == 0 || x == g.nx - 1 || y == g.ny - 1;
    const long ic = (long)z * ip + (long)(y - 1) * nxi + (x - 1);  // compact interior index
    {
      const T dp = a.phi[i] - a.phi0[i], da = a.az[i] - a.az0[i];
      acc[NFGS_DR_PHI_MAP] += sq(dp);
      acc[NFGS_DR_PHI_REF] += sq(a.phi0[i]);
      acc[NFGS_DR_AZ_MAP] += sq(da);
      acc[NFGS_DR_AZ_REF] += sq(a.az0[i]);
    }
    if (z < 1 || z > g.nz - 2) continue;
    // d2/dz2 at an interior plane (nfgs_solver.py:1717-1727)
    auto d2z = [&](const T* f) { return (f[i + pl] - T(2) * f[i] + f[i - pl]) * g.idz2; };
    const T d2p = d2z(a.phi), lpz = a.lp[i];
    const T res_phi = (d2p + lpz) - a.sp[i];
    if (z >= zb0) {
      // B_c,z = lap_xy(Phi_c): (d_zz + lap_xy) B_c,z - S_Bcz
      const T lbc = lap_xy_at(a.lp + (long)z * pl, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
      const T res_b = (d2z(a.lp) + lbc) - a.sb[i];
      acc[NFGS_DR_BCZ_RES] += sq(res_b);
      acc[NFGS_DR_BCZ_SRC] += sq(a.sb[i]);
      acc[NFGS_DR_POL_RES] += sq(res_phi);
      acc[NFGS_DR_POL_D2Z] += sq(d2p);
      acc[NFGS_DR_POL_LAP] += sq(lpz);
      acc[NFGS_DR_POL_SRC] += sq(a.sp[i]);
    }
    {
      // J_z* = -lap_xy(Az): (d_zz + lap_xy) J_z* - S_Jz
      const T djz = -d2z(a.la);
      const T ljz = -lap_xy_at(a.la + (long)z * pl, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
      const T sjv = wall ? a.sje[i] : a.sj[ic];
      acc[NFGS_DR_JZ_RES] += sq((djz + ljz) - sjv);
      acc[NFGS_DR_JZ_SRC] += sq(sjv);
    }
    if (!wall) {
      const T d2a = d2z(a.az), lza = a.la[i], sav = a.sa[ic];
      acc[NFGS_DR_TOR_RES] += sq((d2a + lza) - sav);
      acc[NFGS_DR_TOR_D2Z] += sq(d2a);
      acc[NFGS_DR_TOR_LAP] += sq(lza);
      acc[NFGS_DR_TOR_SRC] += sq(sav);
    }
    if (z == 1) {
      // the rho-mixed first interior row
      const T mpde = (T(1) - a.rho) * res_phi;
      const T dphi = (T(-3) * a.phi[r] + T(4) * a.phi[pl + r] - a.phi[2 * pl + r]) / (T(2) * g.dz);
      const T mder = -a.rho * (dphi - a.target[r]) / g.dz;
      acc[NFGS_DR_MIX_RES] += sq(mpde + mder);
      acc[NFGS_DR_MIX_PDE] += sq(mpde);
      acc[NFGS_DR_MIX_DER] += sq(mder);
      acc[NFGS_DR_MIX_SRC] += sq(a.sp[i]);
    }
  }
  block_sum<NFGS_DR_COUNT>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NFGS_DR_COUNT; ++q) partials[(long)q * kDiagBlocks + blockIdx.x] = acc[q];
  }
}

// Vertical-face state of (B, J_z) (nfgs_solver.py:5738-5779) and the two
// fractions of :6660-6680; the bottom imprint of B (:6897-6901).
template <typename T>
__global__ __launch_bounds__(256) void face_fraction_kernel(
    DiagGrid<T> g, const T* __restrict__ b, const T* __restrict__ bg, const T* __restrict__ az,
    const T* __restrict__ jzg, const T* __restrict__ jzt, const T* __restrict__ bxo,
    const T* __restrict__ byo,
    const T* __restrict__ bzo, T floor_b2, T amax, T i2dx, T i2dy, T idz,
    double* __restrict__ partials) {
  __shared__ double red[NFGS_DF_COUNT * 16];
  double acc[NFGS_DF_COUNT] = {0.0, 0.0, 0.0, 0.0};
  const long pl = g.plane, nf = (long)(g.nz - 1) * pl;
  const T* bx = b;
  const T* by = b + g.N;
  const T* bz = b + 2 * g.N;
  const T* gx = bg;
  const T* gy = bg + g.N;
  const T* gz = bg + 2 * g.N;
  // centre J_z: the caller's (jzg), or the state's own: the bottom current,
  // zero on the top plane and the walls, -lap_xy(Az) inside
  auto jzc = [&](int k, int y, int x) -> T {
    if (jzg) return jzg[(long)k * pl + (long)y * g.nx + x];
    if (k == 0) return jzt[(long)y * g.nx + x];
    if (k == g.nz - 1 || x == 0 || y == 0 || x == g.nx - 1 || y == g.ny - 1) return T(0);
    return -lap_xy_at(az + (long)k * pl, y, x, g.ny, g.nx, (long)g.nx, g.idx2, g.idy2);
  };
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
       i += (long)gridDim.x * blockDim.x) {
    const int f = (int)(i / pl);
    const long r = i - (long)f * pl;
    const int y = (int)(r / g.nx), x = (int)(r - (long)y * g.nx);
    const long lo = i, hi = i + pl;
    const T fb0 = T(0.5) * (bx[hi] + bx[lo]);
    const T fb1 = T(0.5) * (by[hi] + by[lo]);
    const T fb2 = T(0.5) * (bz[hi] + bz[lo]);
    // face value of the carrier's z component at an offset o along the plane
    auto cfz = [&](long o) {
      return T(0.5) * ((bz[hi + o] - gz[hi + o]) + (bz[lo + o] - gz[lo + o]));
    };
    T dyc, dxc;
    if (y > 0 && y < g.ny - 1)
      dyc = (cfz(g.nx) - cfz(-g.nx)) * i2dy;
    else if (y == 0)
      dyc = (T(-3) * cfz(0) + T(4) * cfz(g.nx) - cfz(2L * g.nx)) * i2dy;
    else
      dyc = (T(3) * cfz(0) - T(4) * cfz(-g.nx) + cfz(-2L * g.nx)) * i2dy;
    if (x > 0 && x < g.nx - 1)
      dxc = (cfz(1) - cfz(-1)) * i2dx;
    else if (x == 0)
      dxc = (T(-3) * cfz(0) + T(4) * cfz(1) - cfz(2)) * i2dx;
    else
      dxc = (T(3) * cfz(0) - T(4) * cfz(-1) + cfz(-2)) * i2dx;
    const T dbx = ((bx[hi] - gx[hi]) - (bx[lo] - gx[lo])) * idz;
    const T dby = ((by[hi] - gy[hi]) - (by[lo] - gy[lo])) * idz;
    const T fj0 = dyc - dby, fj1 = dbx - dxc;
    const T fj2 = T(0.5) * (jzc(f + 1, y, x) + jzc(f, y, x));
    const T b2 = fb0 * fb0 + fb1 * fb1 + fb2 * fb2;
    acc[NFGS_DF_WEAK] += b2 <= floor_b2 ? 1.0 : 0.0;
    const T alpha = (fj0 * fb0 + fj1 * fb1 + fj2 * fb2) / (b2 < floor_b2 ? floor_b2 : b2);
    acc[NFGS_DF_CLIP] += (alpha < T(0) ? -alpha : alpha) > amax ? 1.0 : 0.0;
    if (f == 0) {
      acc[NFGS_DF_IMPRINT] += sq(T(bx[r] - bxo[r])) + sq(T(by[r] - byo[r])) + sq(T(bz[r] - bzo[r]));
      acc[NFGS_DF_OBS] += sq(bxo[r]) + sq(byo[r]) + sq(bzo[r]);
    }
  }
  block_sum<NFGS_DF_COUNT>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NFGS_DF_COUNT; ++q) partials[(long)q * kDiagBlocks + blockIdx.x] = acc[q];
  }
}

// out[q] = sum of the kDiagBlocks partials of slot q, in block order
__global__ __launch_bounds__(64) void diag_reduce_kernel(const double* __restrict__ partials,
                                                         int nq, double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  for (int b = 0; b < kDiagBlocks; ++b) s += partials[(long)q * kDiagBlocks + b];
  out[q] = s;
}

unsigned grid_for(long n) {
  const long nb = (n + 255) / 256;
  return (unsigned)(nb < kDiagBlocks ? nb : kDiagBlocks);
}

// Workspace: lap_xy(Phi_c), lap_xy(Az), source_phi and two transform
// volumes, the compact Az right-hand side, then the FP64 partials.
struct DiagWs {
  char* base;
  long N, ip;
  int esize;
  void* lp() const { return base; }
  void* la() const { return base + N * esize; }
  void* sp() const { return base + 2 * N * esize; }
  void* t0() const { return base + 3 * N * esize; }
  void* t1() const { return base + 4 * N * esize; }
  void* sa() const { return base + 5 * N * esize; }
  double* partials() const {
    const long off = (5 * N + ip) * esize;
    return reinterpret_cast<double*>(base + ((off + 255) & ~255L));
  }
};

DiagWs diag_ws(const nfgs_plan* p, void* w) {
  return DiagWs{static_cast<char*>(w), p->N, (long)p->d.nz * p->iplane, p->esize};
}

template <typename T>
int residuals_t(const nfgs_plan* p, const void* phi0, const void* az0, const void* phi,
                const void* az, const void* sb, const void* sj, const void* sje,
                const void* target, void* workspace, double* out, hipStream_t s) {
  const DiagGrid<T> g = diag_grid<T>(p);
  const DiagWs w = diag_ws(p, workspace);
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz, nxi = nx - 2, nyi = ny - 2;
  const long pl = p->plane, ip = p->iplane;
  // lap_xy of the new state
  NFGS_TRY(nfgs_lap_xy(p, phi, w.lp(), nz, s));
  NFGS_TRY(nfgs_lap_xy(p, az, w.la(), nz, s));
  // source_phi = IDCT2(DCT2(S_Bcz) / lambda), zero mode 0 (:6833-6839)
  NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_FORWARD, sb, 1, nx, pl, w.t0(), 1, nx, pl, ny, nz, s));
  NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_FORWARD, w.t0(), nx, 1, pl, w.t1(), nx, 1, pl, nx, nz, s));
  hipLaunchKernelGGL((modal_scale_kernel<T>), dim3(grid_for(p->N)), dim3(256), 0, s,
                     static_cast<T*>(w.t1()), static_cast<const T*>(p->lam), pl, p->N, 0);
  NFGS_LAUNCH_CHECK();
  NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_INVERSE, w.t1(), nx, 1, pl, w.t0(), nx, 1, pl, nx, nz, s));
  NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_INVERSE, w.t0(), 1, nx, pl, w.sp(), 1, nx, pl, ny, nz, s));
  // Az right-hand side = IDST2(-DST2(S_Jz) / lambda_D) on the interior (:6841-6853)
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_FORWARD, sj, 1, nxi, ip, w.t0(), 1, nxi, ip, nyi, nz, s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_FORWARD, w.t0(), nxi, 1, ip, w.t1(), nxi, 1, ip, nxi, nz, s));
  hipLaunchKernelGGL((modal_scale_kernel<T>), dim3(grid_for((long)nz * ip)), dim3(256), 0, s,
                     static_cast<T*>(w.t1()), static_cast<const T*>(p->lamd), ip, (long)nz * ip, 1);
  NFGS_LAUNCH_CHECK();
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_INVERSE, w.t1(), nxi, 1, ip, w.t0(), nxi, 1, ip, nxi, nz, s));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_INVERSE, w.t0(), 1, nxi, ip, w.sa(), 1, nxi, ip, nyi, nz, s));
  ResidualIn<T> a;
  a.phi0 = static_cast<const T*>(phi0);
  a.az0 = static_cast<const T*>(az0);
  a.phi = static_cast<const T*>(phi);
  a.az = static_cast<const T*>(az);
  a.lp = static_cast<const T*>(w.lp());
  a.la = static_cast<const T*>(w.la());
  a.sb = static_cast<const T*>(sb);
  a.sj = static_cast<const T*>(sj);
  a.sje = static_cast<const T*>(sje);
  a.sp = static_cast<const T*>(w.sp());
  a.sa = static_cast<const T*>(w.sa());
  a.target = static_cast<const T*>(target);
  a.rho = T(p->d.rho);
  hipLaunchKernelGGL((residual_kernel<T>), dim3(kDiagBlocks), dim3(256), 0, s, g, a, w.partials());
  NFGS_LAUNCH_CHECK();
  hipLaunchKernelGGL(diag_reduce_kernel, dim3(1), dim3(64), 0, s, w.partials(), NFGS_DR_COUNT, out);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace
}  // namespace nfgs

using namespace nfgs;

extern "C" long nfgs_iteration_diagnostics_workspace_bytes(const nfgs_plan* p) {
  if (!p) return 0;
  const long off = (5 * p->N + (long)p->d.nz * p->iplane) * p->esize;
  return ((off + 255) & ~255L) + (long)NFGS_DR_COUNT * kDiagBlocks * (long)sizeof(double);
}

extern "C" int nfgs_diag_source_edges(const nfgs_plan* p, const void* jp, void* sj_walls,
                                      void* stream) {
  NFGS_REQUIRE(p && jp && sj_walls, NFGS_ERR_ARG, "diag_source_edges: null argument");
  hipStream_t s = as_stream(stream);
  const long n = (2L * p->d.nx + 2L * (p->d.ny - 2)) * p->d.nz;
  if (p->d.dtype == NFGS_F32)
    hipLaunchKernelGGL((source_edges_kernel<float>), dim3(grid_for(n)), dim3(256), 0, s,
                       diag_grid<float>(p), static_cast<const float*>(jp),
                       static_cast<float*>(sj_walls));
  else
    hipLaunchKernelGGL((source_edges_kernel<double>), dim3(grid_for(n)), dim3(256), 0, s,
                       diag_grid<double>(p), static_cast<const double*>(jp),
                       static_cast<double*>(sj_walls));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

extern "C" int nfgs_iteration_residuals(const nfgs_plan* p, const void* phi_old,
                                        const void* az_old, const void* phi_new,
                                        const void* az_new, const void* s_bcz, const void* s_jz,
                                        const void* sj_walls, const void* dphi_target,
                                        void* workspace, double* out, void* stream) {
  NFGS_REQUIRE(p && phi_old && az_old && phi_new && az_new && s_bcz && s_jz && sj_walls &&
                   dphi_target && workspace && out,
               NFGS_ERR_ARG, "iteration_residuals: null argument");
  NFGS_REQUIRE(p->d.nz >= 3, NFGS_ERR_ARG, "iteration_residuals: nz must be >= 3");
  if (p->d.dtype == NFGS_F32)
    return residuals_t<float>(p, phi_old, az_old, phi_new, az_new, s_bcz, s_jz, sj_walls,
                              dphi_target, workspace, out, as_stream(stream));
  return residuals_t<double>(p, phi_old, az_old, phi_new, az_new, s_bcz, s_jz, sj_walls,
                             dphi_target, workspace, out, as_stream(stream));
}

extern "C" int nfgs_face_fractions(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b,
                                   const void* az, const void* jz, const void* bx_obs,
                                   const void* by_obs, void* workspace, double* out, void* stream) {
  NFGS_REQUIRE(p && bnd && b && (az || jz) && bx_obs && by_obs && workspace && out, NFGS_ERR_ARG,
               "face_fractions: null argument");
  NFGS_REQUIRE(bnd->b_bg && bnd->bz_obs && bnd->jz_t, NFGS_ERR_STATE,
               "face_fractions: boundary is not set");
  hipStream_t s = as_stream(stream);
  double* part = diag_ws(p, workspace).partials();
  auto run = [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL((face_fraction_kernel<T>), dim3(kDiagBlocks), dim3(256), 0, s,
                       diag_grid<T>(p), static_cast<const T*>(b), static_cast<const T*>(bnd->b_bg),
                       static_cast<const T*>(az), static_cast<const T*>(jz),
                       static_cast<const T*>(bnd->jz_t),
                       static_cast<const T*>(bx_obs), static_cast<const T*>(by_obs),
                       static_cast<const T*>(bnd->bz_obs), T(p->d.weak_field_floor_b2),
                       T(p->d.alpha_max), T(1.0 / (2.0 * p->d.dx)), T(1.0 / (2.0 * p->d.dy)),
                       T(1.0 / p->d.dz), part);
  };
  if (p->d.dtype == NFGS_F32)
    run(0.0f);
  else
    run(0.0);
  NFGS_LAUNCH_CHECK();
  hipLaunchKernelGGL(diag_reduce_kernel, dim3(1), dim3(64), 0, s, part, NFGS_DF_COUNT, out);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}
