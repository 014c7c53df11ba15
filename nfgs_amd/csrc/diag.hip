// Post-selection diagnostics: relative field differences and the NFGS
// source-attribution diagnostics of the selected state.

#include "util.h"

namespace nfgs {
namespace {

template <typename T>
int sq_diff_t(const nfgs_plan* p, const T* a, const T* b, long n, double* out, hipStream_t s) {
  NFGS_TRY(reduce_to(n, [=] __device__(long i) {
    const double d = (double)a[i] - (double)b[i];
    return d * d;
  }, p->red, out, 0, s));
  return reduce_to(n, [=] __device__(long i) { const double v = b[i]; return v * v; },
                   p->red + kRedBlocks, out + 1, 0, s);
}

// compute_source_diagnostics of the selected B (nfgs_solver.py:7014-7196).
template <typename T>
int source_diag_t(const nfgs_plan* p, const nfgs_boundary* bnd, const T* B, const T* d0,
                  const T* dphi_target, const T* bz_above, T* ws, double* out, hipStream_t s) {
  Ctx<T> c{p, s, p->d.nx, p->d.ny, p->d.nz, p->plane};
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const long pl = c.pl, N = p->N;
  const T dx = T(p->d.dx), dy = T(p->d.dy), dz = T(p->d.dz);
  const T* bg = static_cast<const T*>(bnd->b_bg);
  const T* bzo = static_cast<const T*>(bnd->bz_obs);
  const T* phib = static_cast<const T*>(bnd->phi_bottom);
  const T* lam = static_cast<const T*>(p->lam);
  const T* kk = static_cast<const T*>(p->kk);
  double* red = p->red;
  T* W[24];
  for (int i = 0; i < 24; ++i) W[i] = ws + (long)i * pl;
  T* bc3 = W[0];    // Bc_z on planes nz-3, nz-2, nz-1 (3 planes)
  T* h3 = W[3];     // their DCT2 (3 planes)
  T* t1 = W[6];     // transform scratch (3 planes)
  T* phi = W[9];    // Phi_c on planes 0, 1, 2, nz-3, nz-2, nz-1 (6 planes)
  T* ph3 = W[15];   // DCT2 of the top three Phi_c planes (3 planes)
  T* r1 = W[18];
  T* r2 = W[19];
  T* r3 = W[20];
  const T* Bz = B + 2 * N;
  const T* bgz = bg + 2 * N;
  const long top3 = (long)(nz - 3) * pl;

  NFGS_TRY(launch_map(3 * pl, [=] __device__(long i) { bc3[i] = Bz[top3 + i] - bgz[top3 + i]; },
                      s));
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, bc3, h3, t1, 3));
  // Poisson2d (endpoint Neumann, zero constant mode) of Bc_z on the planes
  // that Phi_c is needed on; plane 0 is Phi_above_bottom.
  const int planes[6] = {0, 1, 2, nz - 3, nz - 2, nz - 1};
  for (int q = 0; q < 6; ++q) {
    T* dst = phi + (long)q * pl;
    const int z = planes[q];
    if (z == 0) {
      NFGS_TRY(launch_map(pl, [=] __device__(long i) { dst[i] = phib[i]; }, s));
      continue;
    }
    const long zo = (long)z * pl;
    NFGS_TRY(launch_map(pl, [=] __device__(long i) { r1[i] = Bz[zo + i] - bgz[zo + i]; }, s));
    NFGS_TRY(xy_pair(c, NFGS_FORWARD, r1, r2, t1, 1));
    NFGS_TRY(launch_map(pl, [=] __device__(long i) {
      r2[i] = i == 0 ? T(0) : r2[i] / lam[i];
    }, s));
    NFGS_TRY(xy_pair(c, NFGS_INVERSE, r2, dst, t1, 1));
  }
  NFGS_TRY(xy_pair(c, NFGS_FORWARD, phi + 3 * pl, ph3, t1, 3));
  // Robin residuals and upward part at the top (nfgs_solver.py:7031-7060).
  const T two_dz = T(2) * dz;
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    const T der = (T(3) * h3[2 * pl + i] - T(4) * h3[pl + i] + h3[i]) / two_dz;
    const T pder = (T(3) * ph3[2 * pl + i] - T(4) * ph3[pl + i] + ph3[i]) / two_dz;
    r1[i] = i == 0 ? T(0) : der + kk[i] * h3[2 * pl + i];
    r2[i] = i == 0 ? T(0) : pder + kk[i] * ph3[2 * pl + i];
    const T ks = i == 0 ? T(1) : kk[i];
    r3[i] = i == 0 ? T(0) : T(0.5) * (h3[2 * pl + i] + der / ks);
  }, s));
  T* robin = W[21];
  T* probin = W[22];
  T* up = W[23];
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, r1, robin, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, r2, probin, t1, 1));
  NFGS_TRY(xy_pair(c, NFGS_INVERSE, r3, up, t1, 1));

  auto sq = [=](const T* a) {
    return [=] __device__(long i) { const double v = a[i]; return v * v; };
  };
  int q = 0;
  double* o = out;
  (void)q;
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const double v = (double)(bgz[i] + bz_above[i] - bzo[i]);
    return v * v;
  }, red, o + 0, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const double v = (double)(bgz[i] + bz_above[i] - bzo[i]);
    return v < 0 ? -v : v;
  }, red + kRedBlocks, o + 1, 1, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const double v = (double)((Bz[i] - bgz[i]) - bz_above[i]);
    return v * v;
  }, red + 2 * kRedBlocks, o + 2, 0, s));
  NFGS_TRY(reduce_to(pl, sq(robin), red + 3 * kRedBlocks, o + 3, 0, s));
  NFGS_TRY(reduce_to(pl, sq(probin), red + 4 * kRedBlocks, o + 4, 0, s));
  NFGS_TRY(reduce_to(pl, sq(up), red + 5 * kRedBlocks, o + 5, 0, s));
  NFGS_TRY(reduce_to(pl, sq(bc3 + 2 * pl), red + 6 * kRedBlocks, o + 6, 0, s));
  // bottom divergence of the selected field (nfgs_solver.py:7066-7081)
  const T* Bx = B;
  const T* By = B + N;
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const int y = (int)(i / nx), x = (int)(i - (long)y * nx);
    const T div = d1_plane(Bx, y, x, ny, nx, dx, true) + d1_plane(By, y, x, ny, nx, dy, false);
    const double v = (double)(div - d0[i]);
    return v * v;
  }, red + 7 * kRedBlocks, o + 7, 0, s));
  NFGS_TRY(reduce_to(pl, sq(d0), red + 8 * kRedBlocks, o + 8, 0, s));
  const T* p0 = phi;
  const T* p1 = phi + pl;
  const T* p2 = phi + 2 * pl;
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const T d = (T(-3) * p0[i] + T(4) * p1[i] - p2[i]) / two_dz;
    const double v = (double)(d - dphi_target[i]);
    return v * v;
  }, red + 9 * kRedBlocks, o + 9, 0, s));
  return NFGS_OK;
}

}  // namespace
}  // namespace nfgs

using namespace nfgs;

extern "C" int nfgs_sq_diff(const nfgs_plan* p, const void* a, const void* b, long n, double* out,
                            void* stream) {
  NFGS_REQUIRE(p && a && b && out && n >= 0, NFGS_ERR_ARG, "sq_diff: bad argument");
  if (p->d.dtype == NFGS_F32)
    return sq_diff_t<float>(p, static_cast<const float*>(a), static_cast<const float*>(b), n, out,
                            as_stream(stream));
  return sq_diff_t<double>(p, static_cast<const double*>(a), static_cast<const double*>(b), n, out,
                           as_stream(stream));
}

extern "C" long nfgs_source_diagnostics_workspace_bytes(const nfgs_plan* p) {
  return p ? 24L * p->plane * p->esize : 0;
}

extern "C" int nfgs_source_diagnostics(const nfgs_plan* p, const nfgs_boundary* bnd, const void* b,
                                       const void* d0, const void* dphi_target,
                                       const void* bz_above, void* workspace, double* out,
                                       void* stream) {
  NFGS_REQUIRE(p && bnd && b && d0 && dphi_target && bz_above && workspace && out, NFGS_ERR_ARG,
               "source_diagnostics: null argument");
  NFGS_REQUIRE(bnd->b_bg && bnd->bz_obs && bnd->phi_bottom, NFGS_ERR_STATE,
               "source_diagnostics: boundary is not set");
  NFGS_REQUIRE(p->d.nz >= 3, NFGS_ERR_ARG, "source_diagnostics: nz must be >= 3");
  if (p->d.dtype == NFGS_F32)
    return source_diag_t<float>(p, bnd, static_cast<const float*>(b),
                                static_cast<const float*>(d0),
                                static_cast<const float*>(dphi_target),
                                static_cast<const float*>(bz_above),
                                static_cast<float*>(workspace), out, as_stream(stream));
  return source_diag_t<double>(p, bnd, static_cast<const double*>(b),
                               static_cast<const double*>(d0),
                               static_cast<const double*>(dphi_target),
                               static_cast<const double*>(bz_above),
                               static_cast<double*>(workspace), out, as_stream(stream));
}
