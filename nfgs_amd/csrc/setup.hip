// Boundary setup on the device: bottom Jz, Az_bottom, the toroidal bottom
// current and the Gaussian separation.  Runs once per boundary; it reuses the
// lateral transform kernel for every DCT-I/DST-I and small point kernels for
// the modal algebra, with deterministic two-stage reductions for diagnostics.

#include "util.h"

namespace nfgs {
namespace {

// ---------------------------------------------------------------------------
// Bottom Jz: centred curl on a helper grid, sign-separated balancing, edge
// zeroing, crop (nfgs_solver.py:1498-1567; unpadded form :4519-4546).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void jz_kernel(const T* __restrict__ bx, const T* __restrict__ by,
                                                  T* __restrict__ jz, int ny, int nx, int hny,
                                                  int hnx, int bfy, int bfx, T two_dx, T two_dy,
                                                  int mode, int legacy) {
  __shared__ double sh[3 * 1024];
  __shared__ T factor[2];
  auto val = [&](const T* a, int i, int j) -> T {
    const int y = i - bfy, x = j - bfx;
    return (y >= 0 && y < ny && x >= 0 && x < nx) ? a[(long)y * nx + x] : T(0);
  };
  auto curl = [&](int i, int j) -> T {
    if (i < 1 || i > hny - 2 || j < 1 || j > hnx - 2) return T(0);
    return (val(by, i, j + 1) - val(by, i, j - 1)) / two_dx -
           (val(bx, i + 1, j) - val(bx, i - 1, j)) / two_dy;
  };
  double tot = 0.0, pos = 0.0, neg = 0.0;
  const long hn = (long)hny * hnx;
  if (mode >= 2) {
    for (long e = threadIdx.x; e < hn; e += blockDim.x) {
      const int i = (int)(e / hnx), j = (int)(e - (long)i * hnx);
      const T c = curl(i, j);
      if (c == T(0)) continue;
      // cell quadrature weight: the number of the four shifted interior
      // blocks [1:-1,2:], [2:,1:-1], [1:-1,1:-1], [2:,2:] containing (i, j)
      const bool ia = i >= 1 && i <= hny - 2, ib = i >= 2 && i <= hny - 1;
      const bool ja = j >= 1 && j <= hnx - 2, jb = j >= 2 && j <= hnx - 1;
      const int w = (ia && jb) + (ib && ja) + (ia && ja) + (ib && jb);
      tot += w * (double)c;
      if (c >= T(0)) pos += w * (double)c;
      else neg += w * (double)c;
    }
  }
  sh[threadIdx.x] = tot;
  sh[1024 + threadIdx.x] = pos;
  sh[2048 + threadIdx.x] = neg;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[threadIdx.x] += sh[threadIdx.x + w];
      sh[1024 + threadIdx.x] += sh[1024 + threadIdx.x + w];
      sh[2048 + threadIdx.x] += sh[2048 + threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double t = 0.25 * sh[0], ps = 0.25 * sh[1024], ns = 0.25 * sh[2048];
    double fp = 1.0, fn = 1.0;
    bool apply = mode >= 2 && mode <= 4;
    if (apply) {
      double tgt;
      if (mode == 2) tgt = ps - 0.5 * t;
      else if (mode == 3) tgt = ps > -ns ? ps : -ns;
      else tgt = ps < -ns ? ps : -ns;
      if (ps != 0.0) fp = tgt / ps;
      if (ns != 0.0) fn = tgt / -ns;
    }
    factor[0] = T(fp);
    factor[1] = T(fn);
  }
  __syncthreads();
  const long n = (long)ny * nx;
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    const int y = (int)(e / nx), x = (int)(e - (long)y * nx);
    const int i = y + bfy, j = x + bfx;
    T c = curl(i, j);
    c = c >= T(0) ? factor[0] * c : factor[1] * c;
    if (legacy && (i == 0 || i == hny - 1 || j == 0 || j == hnx - 1)) c = T(0);
    jz[e] = c;
  }
}

// out = exp(-z k) * hat (kind 0) or (-k exp(-z k)) * hat (kind 1), per mode.
template <typename T>
int decay_fill(const nfgs_plan* p, T* out, const T* hat, int kind, hipStream_t s) {
  const long pl = p->plane;
  const T dz = T(p->d.dz);
  const T* kk = static_cast<const T*>(p->kk);
  return launch_map(p->N, [=] __device__(long i) {
    const int z = (int)(i / pl);
    const long m = i - (long)z * pl;
    const T zz = T(z) * dz;
    const T dec = exp(-(zz * kk[m]));
    out[i] = kind == 0 ? dec * hat[m] : ((-kk[m]) * dec) * hat[m];
  }, s);
}

template <typename T>
int gaussian_t(const nfgs_plan* p, const T* bx, const T* by, const T* bz, const T* azb, T* b_bg,
               T* bz_above, T* phi_bottom, T* dphi_target, T* d0, T* pah, T* h_target,
               double* diag, void* workspace, hipStream_t s) {
  Ctx<T> c{p, s, p->d.nx, p->d.ny, p->d.nz, p->plane};
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const long pl = c.pl, N = p->N;
  const T dx = T(p->d.dx), dy = T(p->d.dy), dz = T(p->d.dz);
  const double norm = 4.0 * (nx - 1) * (ny - 1);
  const T* lam = static_cast<const T*>(p->lam);
  const T* kk = static_cast<const T*>(p->kk);
  const T* ax = static_cast<const T*>(p->ax);
  const T* ay = static_cast<const T*>(p->ay);

  T* w = static_cast<T*>(workspace);
  T* P[20];
  for (int i = 0; i < 20; ++i) P[i] = w + (long)i * pl;
  T* V0 = w + 20 * pl;
  T* V1 = V0 + N;
  T* V2 = V1 + N;
  T* V3 = V2 + N;
  double* part = reinterpret_cast<double*>(
      (reinterpret_cast<uintptr_t>(V3 + N) + 255) & ~static_cast<uintptr_t>(255));

  T *bxh = P[0], *byh = P[1], *bzh = P[2], *t1 = P[3], *t2 = P[4], *A = P[5], *Bh = P[6];
  T *dhat = P[7], *dtot_hat = P[8], *bzb = P[9], *bza = P[10], *pbh = P[11], *dtot = P[12];

  // Constant modes of the observed maps (nfgs_solver.py:4675-4685).
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, bx, bxh, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, by, byh, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, bz, bzh, t1, 1));
  NFGS_TRY(launch_map(1, [=] __device__(long) {
    diag[NFGS_GD_BX_MEAN] = (double)bxh[0] / norm;
    diag[NFGS_GD_BY_MEAN] = (double)byh[0] / norm;
    diag[NFGS_GD_BZ_MEAN] = (double)bzh[0] / norm;
  }, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return (double)bz[i]; }, part,
                     diag + NFGS_GD_BZ_SUM, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { double v = bz[i]; return v < 0 ? -v : v; },
                     part, diag + NFGS_GD_BZ_ABS_SUM, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return (double)bx[i]; }, part,
                     diag + NFGS_GD_BX_SUM, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return (double)by[i]; }, part,
                     diag + NFGS_GD_BY_SUM, 0, s));

  // Spectral divergence D0 (nfgs_solver.py:4438-4446, :4788-4808).
  NFGS_TRY(dstx_dcty(c, NFGS_FORWARD, bx, A, t1, 1));
  NFGS_TRY(dctx_dsty(c, NFGS_FORWARD, by, Bh, t1, 1));
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    dhat[i] = ax[i % nx] * A[i] + ay[i / nx] * Bh[i];
  }, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, dhat, d0, t1, 1));
  NFGS_TRY(launch_map(1, [=] __device__(long) { diag[NFGS_GD_D0_MEAN] = (double)dhat[0] / norm; }, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { double v = d0[i]; return v * v; }, part,
                     diag + NFGS_GD_D0_SQ, 0, s));

  // Below/above split of the non-mean modes (nfgs_solver.py:4810-4841).
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    if (i == 0) {
      dtot_hat[0] = bzb[0] = bza[0] = pbh[0] = pah[0] = T(0);
      return;
    }
    const T dh = dhat[i], l = lam[i], k = kk[i], nm = bzh[i];
    dtot_hat[i] = -dh / l;
    const T lo = T(0.5) * (nm + dh / k);
    const T hi = T(0.5) * (nm - dh / k);
    bzb[i] = lo;
    bza[i] = hi;
    pbh[i] = lo / l;
    pah[i] = hi / l;
  }, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, bza, bz_above, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, pah, phi_bottom, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, dtot_hat, dtot, t1, 1));

  // Upward-decaying below field (nfgs_solver.py:4843-4861).
  T* phi_below = V2;
  T* dphi_dz = V3;
  NFGS_TRY(decay_fill(p, V0, pbh, 0, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, V0, phi_below, V1, nz));
  NFGS_TRY(decay_fill(p, V0, bzb, 0, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, V0, b_bg + 2 * N, V1, nz));
  NFGS_TRY(decay_fill(p, V0, pbh, 1, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, V0, dphi_dz, V1, nz));
  const long NN = N;
  // B_background[2] = Bz_below + Bz_open_mean (uniform open flux)
  T* bgz = b_bg + 2 * N;
  NFGS_TRY(launch_map(NN, [=] __device__(long i) { bgz[i] = bgz[i] + T(bzh[0] / T(norm)); }, s));

  // Bottom derivative target (nfgs_solver.py:4866-4868).
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    const T d = (T(-3) * phi_below[i] + T(4) * phi_below[pl + i] - phi_below[2 * pl + i]) /
                (T(2) * dz);
    dphi_target[i] = dtot[i] - d;
  }, s));

  // Invariant diagnostic lap_xy(phi_below) + open - B_bg,z (nfgs_solver.py:4873-4881).
  const T idx2 = T(1.0 / (p->d.dx * p->d.dx)), idy2 = T(1.0 / (p->d.dy * p->d.dy));
  NFGS_TRY(reduce_to(NN, [=] __device__(long i) {
    const int z = (int)(i / pl);
    const long m = i - (long)z * pl;
    const int y = (int)(m / nx), x = (int)(m - (long)y * nx);
    const T lz = lap_xy_at(phi_below + (long)z * pl, y, x, ny, nx, (long)nx, idx2, idy2);
    const double v = (double)((lz + T(bzh[0] / T(norm))) - bgz[i]);
    return v * v;
  }, part, diag + NFGS_GD_BG_INV_SQ, 0, s));

  // Horizontal below field B_h = -grad_h(dPhi/dz) (nfgs_solver.py:4852-4856).
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, dphi_dz, V0, V1, nz));
  NFGS_TRY(neg_spectral_grad(c, V0, b_bg, b_bg + N, V1, V2, nz));

  // Bottom harmonic residual and the frozen linear harmonic target
  // (nfgs_solver.py:4990-5056, :5069-5096).
  T *dph = P[13], *dphi = P[14], *gxh = P[15], *bpx = P[16], *bpy = P[17], *rx = P[18],
    *ry = P[19];
  NFGS_TRY(launch_map(pl, [=] __device__(long i) { dph[i] = kk[i] * (pah[i] - pbh[i]); }, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, dph, dphi, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, dphi, gxh, t1, 1));
  NFGS_TRY(neg_spectral_grad(c, gxh, bpx, bpy, t1, t2, 1));
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    const int y = (int)(i / nx), x = (int)(i - (long)y * nx);
    const T btx = d1_plane(azb, y, x, ny, nx, dy, false);
    const T bty = -d1_plane(azb, y, x, ny, nx, dx, true);
    rx[i] = bx[i] - (bpx[i] + btx);
    ry[i] = by[i] - (bpy[i] + bty);
  }, s));
  T *rxh = P[13], *ryh = P[14];
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, rx, rxh, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, ry, ryh, t1, 1));
  NFGS_TRY(launch_map(1, [=] __device__(long) {
    const T cx = rxh[0] / T(norm), cy = ryh[0] / T(norm);
    h_target[0] = cx;
    h_target[1] = cy;
    diag[NFGS_GD_HCONST_X] = (double)cx;
    diag[NFGS_GD_HCONST_Y] = (double)cy;
  }, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    return (double)rx[i] * rx[i] + (double)ry[i] * ry[i];
  }, part, diag + NFGS_GD_HRES_SQ, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    return (double)bx[i] * bx[i] + (double)by[i] * by[i];
  }, part, diag + NFGS_GD_HOBS_SQ, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const double a = (double)(rx[i] - h_target[0]), b = (double)(ry[i] - h_target[1]);
    return a * a + b * b;
  }, part, diag + NFGS_GD_HNONCONST_SQ, 0, s));
  // spectral divergence of the residual
  T *sa = P[15], *sb = P[16], *dv = P[17];
  NFGS_TRY(dstx_dcty(c, NFGS_FORWARD, rx, sa, t1, 1));
  NFGS_TRY(dctx_dsty(c, NFGS_FORWARD, ry, sb, t1, 1));
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    t2[i] = ax[i % nx] * sa[i] + ay[i / nx] * sb[i];
  }, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, t2, dv, t1, 1));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return (double)dv[i] * dv[i]; }, part,
                     diag + NFGS_GD_HDIV_SQ, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const int y = (int)(i / nx), x = (int)(i - (long)y * nx);
    const double v = (double)(d1_plane(ry, y, x, ny, nx, dx, true) -
                              d1_plane(rx, y, x, ny, nx, dy, false));
    return v * v;
  }, part, diag + NFGS_GD_HCURL_SQ, 0, s));

  // Bottom reconstruction B_bg,z(0) + Bz_above - Bz_obs (nfgs_solver.py:4873).
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const double v = (double)(bgz[i] + bz_above[i] - bz[i]);
    return v * v;
  }, part, diag + NFGS_GD_RECON_SQ, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const double v = (double)(bgz[i] + bz_above[i] - bz[i]);
    return v < 0 ? -v : v;
  }, part, diag + NFGS_GD_RECON_MAX, 1, s));
  return NFGS_OK;
}

template <typename T>
int poisson2d_dirichlet_t(const nfgs_plan* p, const T* src, double scale, T* out, T* work,
                          hipStream_t s) {
  const int nx = p->d.nx, ny = p->d.ny, nxi = nx - 2, nyi = ny - 2;
  const long ip = p->iplane;
  T* w0 = work;
  T* w1 = work + ip;
  const T* lamd = static_cast<const T*>(p->lamd);
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_FORWARD, src + nx + 1, 1, nx, 0, w0, 1, nxi, 0, nyi,
                            1, s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_FORWARD, w0, nxi, 1, 0, w1, nxi, 1, 0, nxi, 1, s));
  const T sc = T(scale);
  NFGS_TRY(launch_map(ip, [=] __device__(long i) { w1[i] = sc * (w1[i] / lamd[i]); }, s));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_INVERSE, w1, 1, nxi, 0, w0, 1, nxi, 0, nyi, 1, s));
  NFGS_TRY(fill_zero(p, out, p->plane, s));
  return launch_transform(p, &p->dst_y, NFGS_INVERSE, w0, nxi, 1, 0, out + nx + 1, nx, 1, 0, nxi,
                          1, s);
}

template <typename T>
int plane_lap_t(const nfgs_plan* p, const T* in, double scale, int zero_edges, T* out,
                hipStream_t s) {
  const int nx = p->d.nx, ny = p->d.ny;
  const T idx2 = T(1.0 / (p->d.dx * p->d.dx)), idy2 = T(1.0 / (p->d.dy * p->d.dy));
  const T sc = T(scale);
  return launch_map(p->plane, [=] __device__(long i) {
    const int y = (int)(i / nx), x = (int)(i - (long)y * nx);
    const bool edge = y == 0 || y == ny - 1 || x == 0 || x == nx - 1;
    out[i] = (zero_edges && edge) ? T(0) : sc * lap_xy_at(in, y, x, ny, nx, (long)nx, idx2, idy2);
  }, s);
}

// Projection of a caller's field into the NFGS scalars: the handoff of
// nfgs_solver.py:6080-6147 with (B, Az) supplied by the caller.  Phi_c[0] is
// Phi_above_bottom and Phi_c[1:] solves the endpoint-Neumann lap_xy Phi_c =
// B_z - B_bg,z plane by plane (:1120-1130, :1146-1180: the constant mode of
// every plane is set to zero; its value per height, :4466-4468, goes to
// mean_bcz).  Az keeps the solver's fixed planes: Az_bottom at the bottom,
// zero on the top plane and on the lateral walls (the planes the iteration
// never rewrites).  Workspace: 3 volumes.
template <typename T>
int project_field_t(const nfgs_plan* p, const nfgs_boundary* bnd, const T* b, const T* az_in,
                    T* phi, T* az, double* mean_bcz, void* workspace, hipStream_t s) {
  Ctx<T> c{p, s, p->d.nx, p->d.ny, p->d.nz, p->plane};
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const long pl = c.pl, N = p->N;
  const double norm = 4.0 * (nx - 1) * (ny - 1);
  const T* lam = static_cast<const T*>(p->lam);
  const T* bgz = static_cast<const T*>(bnd->b_bg) + 2 * N;
  const T* bz = b + 2 * N;
  const T* pb = static_cast<const T*>(bnd->phi_bottom);
  const T* ab = static_cast<const T*>(bnd->az_bottom);
  T* v0 = static_cast<T*>(workspace);
  T* v1 = v0 + N;
  T* v2 = v1 + N;
  NFGS_TRY(launch_map(N, [=] __device__(long i) { v0[i] = bz[i] - bgz[i]; }, s));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, v0, v1, v2, nz));
  NFGS_TRY(launch_map(nz, [=] __device__(long q) { mean_bcz[q] = (double)v1[q * pl] / norm; }, s));
  NFGS_TRY(launch_map(N - pl, [=] __device__(long i) {
    const long j = i % pl;
    v1[pl + i] = j == 0 ? T(0) : v1[pl + i] / lam[j];
  }, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, v1 + pl, phi + pl, v2, nz - 1));
  return launch_map(N, [=] __device__(long i) {
    const long q = i / pl, j = i - q * pl;
    if (q == 0) {
      phi[i] = pb[j];
      az[i] = ab[j];
      return;
    }
    const int y = (int)(j / nx), x = (int)(j - (long)y * nx);
    const bool wall = q == nz - 1 || y == 0 || y == ny - 1 || x == 0 || x == nx - 1;
    az[i] = wall ? T(0) : az_in[i];
  }, s);
}

// The above-bottom (outward) Gaussian field on z = 0, (3, ny, nx)
// component-first: B_h = -grad_h(dPhi/dz) with dPhi/dz = idct2(k Phi^>_hat)
// and the alpha-consistent spectral gradient, B_z = Bz_above
// (nfgs_solver.py:4921-4943, gradient :4385-4406).  Workspace: 5 planes.
template <typename T>
int above_field_t(const nfgs_plan* p, const T* pah, const T* bz_above, T* out, void* workspace,
                  hipStream_t s) {
  Ctx<T> c{p, s, p->d.nx, p->d.ny, p->d.nz, p->plane};
  const long pl = c.pl;
  const T* kk = static_cast<const T*>(p->kk);
  T* w = static_cast<T*>(workspace);
  T *dph = w, *dphi = w + pl, *gh = w + 2 * pl, *t1 = w + 3 * pl, *t2 = w + 4 * pl;
  T* oz = out + 2 * pl;
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    dph[i] = kk[i] * pah[i];
    oz[i] = bz_above[i];
  }, s));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, dph, dphi, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, dphi, gh, t1, 1));
  return neg_spectral_grad(c, gh, out, out + pl, t1, t2, 1);
}

template <typename T>
int two_stage_tables_t(const nfgs_plan* p, const nfgs_boundary* bnd, T* out, hipStream_t s) {
  const long pl = p->plane;
  const T* lam = static_cast<const T*>(p->lam);
  const T* bh = static_cast<const T*>(bnd->phi_bottom_hat);
  const T* qh = static_cast<const T*>(bnd->q_hat);
  return launch_map(pl, [=] __device__(long i) {
    out[i] = lam[i] * bh[i];
    out[pl + i] = qh ? lam[i] * qh[i] : T(0);
  }, s);
}

template <typename T>
int divide_by_lambda_t(const nfgs_plan* p, T* hat, int nplanes, hipStream_t s) {
  const long pl = p->plane;
  const T* lam = static_cast<const T*>(p->lam);
  return launch_map(pl * nplanes, [=] __device__(long i) {
    const long j = i % pl;
    hat[i] = j == 0 ? T(0) : hat[i] / lam[j];
  }, s);
}

}  // namespace

int two_stage_tables(const nfgs_plan* p, const nfgs_boundary* bnd, void* out, hipStream_t s) {
  if (p->d.dtype == NFGS_F32) return two_stage_tables_t(p, bnd, static_cast<float*>(out), s);
  return two_stage_tables_t(p, bnd, static_cast<double*>(out), s);
}

int divide_by_lambda(const nfgs_plan* p, void* hat, int nplanes, hipStream_t s) {
  if (p->d.dtype == NFGS_F32) return divide_by_lambda_t(p, static_cast<float*>(hat), nplanes, s);
  return divide_by_lambda_t(p, static_cast<double*>(hat), nplanes, s);
}

}  // namespace nfgs

using namespace nfgs;

extern "C" long nfgs_above_field_workspace_bytes(const nfgs_plan* p) {
  return p ? 5 * p->plane * (long)p->esize : 0;
}

extern "C" int nfgs_gaussian_above_field(const nfgs_plan* p, const void* phi_above_hat,
                                         const void* bz_above, void* out, void* workspace,
                                         void* stream) {
  NFGS_REQUIRE(p && phi_above_hat && bz_above && out && workspace, NFGS_ERR_ARG,
               "gaussian_above_field: null argument");
  NFGS_REQUIRE(out != phi_above_hat && out != bz_above, NFGS_ERR_ARG,
               "gaussian_above_field: the output must not alias an input");
  hipStream_t s = as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    return above_field_t<float>(p, static_cast<const float*>(phi_above_hat),
                                static_cast<const float*>(bz_above), static_cast<float*>(out),
                                workspace, s);
  return above_field_t<double>(p, static_cast<const double*>(phi_above_hat),
                               static_cast<const double*>(bz_above), static_cast<double*>(out),
                               workspace, s);
}

extern "C" long nfgs_project_workspace_bytes(const nfgs_plan* p) {
  return p ? 3 * p->N * (long)p->esize : 0;
}

extern "C" int nfgs_project_field(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b,
                                  const void* az_in, void* phi, void* az, double* mean_bcz,
                                  void* workspace, void* stream) {
  NFGS_REQUIRE(p && bnd && b && az_in && phi && az && mean_bcz && workspace, NFGS_ERR_ARG,
               "project_field: null argument");
  NFGS_REQUIRE(bnd->b_bg && bnd->phi_bottom && bnd->az_bottom, NFGS_ERR_ARG,
               "project_field: the boundary has no B_background / bottom planes");
  NFGS_REQUIRE(phi != b && az != b && phi != az, NFGS_ERR_ARG,
               "project_field: outputs must not alias the field or each other");
  hipStream_t s = as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    return project_field_t<float>(p, bnd, static_cast<const float*>(b),
                                  static_cast<const float*>(az_in), static_cast<float*>(phi),
                                  static_cast<float*>(az), mean_bcz, workspace, s);
  return project_field_t<double>(p, bnd, static_cast<const double*>(b),
                                 static_cast<const double*>(az_in), static_cast<double*>(phi),
                                 static_cast<double*>(az), mean_bcz, workspace, s);
}

extern "C" int nfgs_setup_jz(const nfgs_plan* p, const void* bx, const void* by, void* jz, int hny,
                             int hnx, int bfy, int bfx, int ijzbcor, int legacy, void* stream) {
  NFGS_REQUIRE(p && bx && by && jz, NFGS_ERR_ARG, "setup_jz: null argument");
  NFGS_REQUIRE(hny >= p->d.ny + bfy && hnx >= p->d.nx + bfx && bfy >= 0 && bfx >= 0,
               NFGS_ERR_ARG, "setup_jz: helper grid does not contain the calculation grid");
  hipStream_t s = as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    hipLaunchKernelGGL((jz_kernel<float>), dim3(1), dim3(1024), 0, s,
                       static_cast<const float*>(bx), static_cast<const float*>(by),
                       static_cast<float*>(jz), p->d.ny, p->d.nx, hny, hnx, bfy, bfx,
                       float(2.0 * p->d.dx), float(2.0 * p->d.dy), ijzbcor, legacy);
  else
    hipLaunchKernelGGL((jz_kernel<double>), dim3(1), dim3(1024), 0, s,
                       static_cast<const double*>(bx), static_cast<const double*>(by),
                       static_cast<double*>(jz), p->d.ny, p->d.nx, hny, hnx, bfy, bfx,
                       2.0 * p->d.dx, 2.0 * p->d.dy, ijzbcor, legacy);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

extern "C" int nfgs_setup_poisson2d_dirichlet(const nfgs_plan* p, const void* src, double scale,
                                              void* out, void* work, void* stream) {
  NFGS_REQUIRE(p && src && out && work && src != out, NFGS_ERR_ARG,
               "setup_poisson2d_dirichlet: bad argument");
  if (p->d.dtype == NFGS_F32)
    return poisson2d_dirichlet_t<float>(p, static_cast<const float*>(src), scale,
                                        static_cast<float*>(out), static_cast<float*>(work),
                                        as_stream(stream));
  return poisson2d_dirichlet_t<double>(p, static_cast<const double*>(src), scale,
                                       static_cast<double*>(out), static_cast<double*>(work),
                                       as_stream(stream));
}

extern "C" int nfgs_setup_plane_lap(const nfgs_plan* p, const void* in, double scale,
                                    int zero_edges, void* out, void* stream) {
  NFGS_REQUIRE(p && in && out && in != out, NFGS_ERR_ARG, "setup_plane_lap: bad argument");
  if (p->d.dtype == NFGS_F32)
    return plane_lap_t<float>(p, static_cast<const float*>(in), scale, zero_edges,
                              static_cast<float*>(out), as_stream(stream));
  return plane_lap_t<double>(p, static_cast<const double*>(in), scale, zero_edges,
                             static_cast<double*>(out), as_stream(stream));
}

extern "C" long nfgs_gaussian_workspace_bytes(const nfgs_plan* p) {
  if (!p) return 0;
  return (20 * p->plane + 4 * p->N) * (long)p->esize + 4096L * 8;
}

extern "C" int nfgs_setup_gaussian(const nfgs_plan* p, const void* bx, const void* by,
                                   const void* bz, const void* az_bottom, void* b_bg,
                                   void* bz_above, void* phi_bottom, void* dphi_target, void* d0,
                                   void* phi_above_hat, void* h_target, double* diag,
                                   void* workspace, void* stream) {
  NFGS_REQUIRE(p && bx && by && bz && az_bottom && b_bg && bz_above && phi_bottom && dphi_target &&
                   d0 && phi_above_hat && h_target && diag && workspace,
               NFGS_ERR_ARG, "setup_gaussian: null argument");
  hipStream_t s = as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    return gaussian_t<float>(
        p, static_cast<const float*>(bx), static_cast<const float*>(by),
        static_cast<const float*>(bz), static_cast<const float*>(az_bottom),
        static_cast<float*>(b_bg), static_cast<float*>(bz_above), static_cast<float*>(phi_bottom),
        static_cast<float*>(dphi_target), static_cast<float*>(d0),
        static_cast<float*>(phi_above_hat), static_cast<float*>(h_target), diag, workspace, s);
  return gaussian_t<double>(
      p, static_cast<const double*>(bx), static_cast<const double*>(by),
      static_cast<const double*>(bz), static_cast<const double*>(az_bottom),
      static_cast<double*>(b_bg), static_cast<double*>(bz_above), static_cast<double*>(phi_bottom),
      static_cast<double*>(dphi_target), static_cast<double*>(d0),
      static_cast<double*>(phi_above_hat), static_cast<double*>(h_target), diag, workspace, s);
}

extern "C" int nfgs_bottom_jz(const void* bx, const void* by, void* jz, int ny, int nx, double dx,
                              double dy, int dtype, int hny, int hnx, int bfy, int bfx,
                              int ijzbcor, int legacy, void* stream) {
  NFGS_REQUIRE(bx && by && jz && ny >= 1 && nx >= 1, NFGS_ERR_ARG, "bottom_jz: bad argument");
  NFGS_REQUIRE(dx > 0.0 && dy > 0.0, NFGS_ERR_ARG, "bottom_jz: spacing must be positive");
  NFGS_REQUIRE(dtype == NFGS_F32 || dtype == NFGS_F64, NFGS_ERR_ARG, "bottom_jz: bad dtype");
  NFGS_REQUIRE(hny >= ny + bfy && hnx >= nx + bfx && bfy >= 0 && bfx >= 0, NFGS_ERR_ARG,
               "bottom_jz: helper grid does not contain the calculation grid");
  hipStream_t s = as_stream(stream);
  if (dtype == NFGS_F32)
    hipLaunchKernelGGL((jz_kernel<float>), dim3(1), dim3(1024), 0, s,
                       static_cast<const float*>(bx), static_cast<const float*>(by),
                       static_cast<float*>(jz), ny, nx, hny, hnx, bfy, bfx, float(2.0 * dx),
                       float(2.0 * dy), ijzbcor, legacy);
  else
    hipLaunchKernelGGL((jz_kernel<double>), dim3(1), dim3(1024), 0, s,
                       static_cast<const double*>(bx), static_cast<const double*>(by),
                       static_cast<double*>(jz), ny, nx, hny, hnx, bfy, bfx, 2.0 * dx, 2.0 * dy,
                       ijzbcor, legacy);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

extern "C" int nfgs_lap_xy(const nfgs_plan* p, const void* in, void* out, int nplanes,
                           void* stream) {
  NFGS_REQUIRE(p && in && out && in != out && nplanes >= 1, NFGS_ERR_ARG,
               "lap_xy: bad argument");
  const int nx = p->d.nx, ny = p->d.ny;
  const long pl = p->plane;
  hipStream_t s = as_stream(stream);
  auto run = [&](auto tag) {
    using T = decltype(tag);
    const T* a = static_cast<const T*>(in);
    T* o = static_cast<T*>(out);
    const T idx2 = T(1.0 / (p->d.dx * p->d.dx)), idy2 = T(1.0 / (p->d.dy * p->d.dy));
    return launch_map((long)nplanes * pl, [=] __device__(long i) {
      const long q = i / pl, r = i - q * pl;
      const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
      o[i] = lap_xy_at(a + q * pl, y, x, ny, nx, (long)nx, idx2, idy2);
    }, s);
  };
  if (p->d.dtype == NFGS_F32) return run(0.0f);
  return run(0.0);
}
