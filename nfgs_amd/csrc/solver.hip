// One plain NFGS iteration at mu = 1, enqueued as a fixed kernel sequence.

#include "plan.h"

namespace {

inline void* at(const nfgs_plan* p, void* base, long elems) {
  return static_cast<char*>(base) + elems * p->esize;
}
inline const void* at(const nfgs_plan* p, const void* base, long elems) {
  return static_cast<const char*>(base) + elems * p->esize;
}

#define NFGS_TRY(expr)      \
  do {                      \
    int _rc = (expr);       \
    if (_rc) return _rc;    \
  } while (0)

}  // namespace

// Sources and both Poisson solves into (phi_out, az_out); the state pass is
// left to the caller.
static int advance_solves(const nfgs_plan* p, const nfgs_boundary* bnd, const nfgs_workspace* ws,
                   void* phi_out, void* az_out, int zero_source, int poloidal_mode, void* stream) {
  using namespace nfgs;
  NFGS_REQUIRE(p && bnd && ws && phi_out && az_out, NFGS_ERR_ARG, "advance: null argument");
  NFGS_REQUIRE(poloidal_mode == NFGS_POLOIDAL_DIRECT || poloidal_mode == NFGS_POLOIDAL_TWO_STAGE,
               NFGS_ERR_ARG, "advance: poloidal_mode must be NFGS_POLOIDAL_DIRECT or _TWO_STAGE");
  NFGS_REQUIRE(ws->jp && ws->s_bcz && ws->s_jz && ws->hat0 && ws->hat1 && ws->cscr &&
                   ws->partials && ws->metrics,
               NFGS_ERR_ARG, "advance: workspace buffer missing");
  hipStream_t s = as_stream(stream);
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
  const int nxi = nx - 2, nyi = ny - 2;
  const long pl = p->plane, ip = p->iplane;

  // Sources of both Poisson problems from the parallel current.
  if (zero_source) {
    NFGS_TRY(fill_zero(p, ws->s_bcz, p->N, s));
    NFGS_TRY(fill_zero(p, ws->s_jz, (long)nz * ip, s));
    NFGS_TRY(fill_zero_f64(ws->partials + (long)NFGS_Q_SRC2 * p->pblocks, p->pblocks, s));
  } else {
    NFGS_TRY(nfgs_sources(p, ws->jp, ws->s_bcz, ws->s_jz, ws->partials, stream));
  }

  // Phi_c: DCT-I x, DCT-I y, modal Robin sweep, inverse y... on planes 1..nz-1.
  NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_FORWARD, ws->s_bcz, 1, nx, pl, ws->hat0, 1, nx, pl,
                            ny, nz, s));
  NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_FORWARD, ws->hat0, nx, 1, pl, ws->hat1, nx, 1, pl,
                            nx, nz, s));
  if (poloidal_mode == NFGS_POLOIDAL_DIRECT) {
    NFGS_TRY(nfgs_robin_solve(p, bnd, ws->hat1, ws->hat0, ws->cscr, ws->partials, stream));
  } else {
    // Two-stage (nfgs_solver.py:5674-5725): B_c,z from the Robin solve with
    // the undivided source and lambda-scaled bottom data (two planes of
    // s_bcz, free once the forward transform has read it), back to space and
    // forward again, then Phi_c_hat = B_c,z_hat / lambda on planes 1..nz-1.
    NFGS_REQUIRE(bnd->phi_bottom_hat && (p->d.rho == 0.0 || bnd->q_hat), NFGS_ERR_STATE,
                 "advance: boundary tables are not set");
    NFGS_TRY(two_stage_tables(p, bnd, ws->s_bcz, s));
    NFGS_TRY(robin_modal(p, ws->s_bcz, at(p, ws->s_bcz, pl), ws->hat1, ws->hat0, ws->cscr,
                         ws->partials, false, s));
    NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_INVERSE, ws->hat0, 1, nx, pl, ws->hat1, 1, nx, pl,
                              ny, nz, s));
    NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_INVERSE, ws->hat1, nx, 1, pl, ws->hat0, nx, 1,
                              pl, nx, nz, s));
    NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_FORWARD, ws->hat0, 1, nx, pl, ws->hat1, 1, nx, pl,
                              ny, nz, s));
    NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_FORWARD, ws->hat1, nx, 1, pl, ws->hat0, nx, 1,
                              pl, nx, nz, s));
    NFGS_TRY(divide_by_lambda(p, at(p, ws->hat0, pl), nz - 1, s));
  }
  NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_INVERSE, at(p, ws->hat0, pl), 1, nx, pl,
                            at(p, ws->hat1, pl), 1, nx, pl, ny, nz - 1, s));
  NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_INVERSE, at(p, ws->hat1, pl), nx, 1, pl,
                            at(p, phi_out, pl), nx, 1, pl, nx, nz - 1, s));

  // Az: DST-I on the interior of planes 1..nz-2, Dirichlet sweep, inverse.
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_FORWARD, at(p, ws->s_jz, ip), 1, nxi, ip,
                            at(p, ws->hat0, ip), 1, nxi, ip, nyi, nz - 2, s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_FORWARD, at(p, ws->hat0, ip), nxi, 1, ip,
                            at(p, ws->hat1, ip), nxi, 1, ip, nxi, nz - 2, s));
  NFGS_TRY(nfgs_dirichlet_solve(p, bnd, ws->hat1, ws->hat0, ws->cscr, stream));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_INVERSE, at(p, ws->hat0, ip), 1, nxi, ip,
                            at(p, ws->hat1, ip), 1, nxi, ip, nyi, nz - 2, s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_INVERSE, at(p, ws->hat1, ip), nxi, 1, ip,
                            at(p, az_out, pl + nx + 1), nx, 1, pl, nxi, nz - 2, s));

  return NFGS_OK;
}

// New state: B and its face pass in one z-march (next J_par and the
// monitors; B itself only when b_out is given), then the reduction.
static int finish_state(const nfgs_plan* p, const nfgs_boundary* bnd, const nfgs_workspace* ws,
                 void* phi_out, void* az_out, void* b_out, void* stream) {
  NFGS_TRY(nfgs_state_pass(p, bnd, phi_out, az_out, b_out, ws->jp, nullptr, ws->partials, stream));
  return nfgs_reduce(p, ws->partials, ws->metrics, stream);
}

extern "C" int nfgs_advance_ex(const nfgs_plan* p, const nfgs_boundary* bnd,
                               const nfgs_workspace* ws, void* phi_out, void* az_out, void* b_out,
                               int zero_source, int poloidal_mode, void* stream) {
  NFGS_TRY(advance_solves(p, bnd, ws, phi_out, az_out, zero_source, poloidal_mode, stream));
  return finish_state(p, bnd, ws, phi_out, az_out, b_out, stream);
}

extern "C" int nfgs_advance_k0(const nfgs_plan* p, const nfgs_boundary* bnd,
                               const nfgs_workspace* ws, void* phi_out, void* az_out, void* b_out,
                               int zero_source, int poloidal_mode, const nfgs_k0_step* k0,
                               void* stream) {
  using namespace nfgs;
  NFGS_REQUIRE(k0 && bnd && k0->h_new == bnd->h_profile, NFGS_ERR_ARG,
               "advance_k0: the boundary's h_profile must be the step's new profile");
  NFGS_TRY(advance_solves(p, bnd, ws, phi_out, az_out, zero_source, poloidal_mode, stream));
  // the closure update reads the step's J_par (still in ws->jp) and the new
  // scalars; the gs_cfcf start keeps the old profile
  if (zero_source) {
    if (k0->h_new != k0->h_old)
      NFGS_HIP_CHECK(hipMemcpyAsync(k0->h_new, k0->h_old, (size_t)2 * p->d.nz * p->esize,
                                    hipMemcpyDeviceToDevice, as_stream(stream)));
  } else {
    NFGS_TRY(k0_update_scratch(p, k0, ws->jp, phi_out, az_out, ws->s_bcz, ws->hat0, ws->hat1,
                               as_stream(stream)));
  }
  // the sources and the solves are done with s_bcz and hat0: the spline
  // closure's J_z and wall current are built there
  NFGS_TRY(closure_state_pass(p, bnd, k0->kind, phi_out, az_out, b_out, ws->jp, nullptr,
                              ws->partials, ws->s_bcz, ws->hat0, as_stream(stream)));
  return nfgs_reduce(p, ws->partials, ws->metrics, stream);
}

namespace {

// Keep the bottom row of J_par where the observed B_z has the chosen sign
// (nfgs_solver.py:6694-6701): J_par[:, 0] = J[:, 0] * mask.
template <typename T>
__global__ __launch_bounds__(256) void polarity_mask_kernel(T* __restrict__ jp,
                                                            const T* __restrict__ bz_obs,
                                                            long plane, long N, int positive) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= plane) return;
  const T b = bz_obs[i];
  const T keep = (positive ? b > T(0) : b < T(0)) ? T(1) : T(0);
  jp[i] = jp[i] * keep;
  jp[N + i] = jp[N + i] * keep;
  jp[2 * N + i] = jp[2 * N + i] * keep;
}

// phi = mu phi + (1 - mu) phi_old on planes 1..nz-1; az likewise on the
// interior points of planes 1..nz-2 (nfgs_solver.py:6778-6790).  The fixed
// planes are left as they are.
template <typename T>
__global__ __launch_bounds__(256) void relax_kernel(T* __restrict__ phi,
                                                    const T* __restrict__ phi_old,
                                                    T* __restrict__ az,
                                                    const T* __restrict__ az_old, int nx, int ny,
                                                    int nz, T mu, T rest) {
#pragma clang fp contract(off)
  const long pl = (long)nx * ny;
  const long i = pl + (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= pl * nz) return;
  phi[i] = mu * phi[i] + rest * phi_old[i];
  const int z = (int)(i / pl);
  const long r = i - (long)z * pl;
  const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
  if (z < nz - 1 && x > 0 && x < nx - 1 && y > 0 && y < ny - 1)
    az[i] = mu * az[i] + rest * az_old[i];
}

}  // namespace

extern "C" int nfgs_advance_relaxed(const nfgs_plan* p, const nfgs_boundary* bnd,
                                    const nfgs_workspace* ws, const void* phi_old,
                                    const void* az_old, double mu, int polarity, void* phi_out,
                                    void* az_out, void* b_out, int poloidal_mode,
                                    const nfgs_k0_step* k0, void* stream) {
  using namespace nfgs;
  NFGS_REQUIRE(p && bnd && ws && phi_old && az_old && phi_out && az_out, NFGS_ERR_ARG,
               "advance_relaxed: null argument");
  NFGS_REQUIRE(mu > 0.0 && mu <= 1.0, NFGS_ERR_ARG, "advance_relaxed: mu must be in (0, 1]");
  NFGS_REQUIRE(polarity == NFGS_POLARITY_BOTH || polarity == NFGS_POLARITY_POSITIVE ||
                   polarity == NFGS_POLARITY_NEGATIVE,
               NFGS_ERR_ARG, "advance_relaxed: unknown polarity");
  NFGS_REQUIRE(!k0 || k0->h_new == bnd->h_profile, NFGS_ERR_ARG,
               "advance_relaxed: the boundary's h_profile must be the step's new profile");
  NFGS_REQUIRE(bnd->bz_obs && ws->jp, NFGS_ERR_STATE, "advance_relaxed: boundary is not set");
  hipStream_t s = as_stream(stream);
  const bool f32 = p->d.dtype == NFGS_F32;
  if (polarity != NFGS_POLARITY_BOTH) {
    const dim3 g((unsigned)((p->plane + 255) / 256));
    const int pos = polarity == NFGS_POLARITY_POSITIVE;
    if (f32)
      hipLaunchKernelGGL(polarity_mask_kernel<float>, g, dim3(256), 0, s,
                         static_cast<float*>(ws->jp), static_cast<const float*>(bnd->bz_obs),
                         p->plane, p->N, pos);
    else
      hipLaunchKernelGGL(polarity_mask_kernel<double>, g, dim3(256), 0, s,
                         static_cast<double*>(ws->jp), static_cast<const double*>(bnd->bz_obs),
                         p->plane, p->N, pos);
    NFGS_LAUNCH_CHECK();
  }
  NFGS_TRY(advance_solves(p, bnd, ws, phi_out, az_out, 0, poloidal_mode, stream));
  if (k0)
    NFGS_TRY(k0_update_scratch(p, k0, ws->jp, phi_out, az_out, ws->s_bcz, ws->hat0, ws->hat1, s));
  if (mu < 1.0) {
    const long n = p->N - p->plane;
    const dim3 g((unsigned)((n + 255) / 256));
    if (f32)
      hipLaunchKernelGGL(relax_kernel<float>, g, dim3(256), 0, s, static_cast<float*>(phi_out),
                         static_cast<const float*>(phi_old), static_cast<float*>(az_out),
                         static_cast<const float*>(az_old), p->d.nx, p->d.ny, p->d.nz,
                         (float)mu, (float)(1.0 - mu));
    else
      hipLaunchKernelGGL(relax_kernel<double>, g, dim3(256), 0, s, static_cast<double*>(phi_out),
                         static_cast<const double*>(phi_old), static_cast<double*>(az_out),
                         static_cast<const double*>(az_old), p->d.nx, p->d.ny, p->d.nz, mu,
                         1.0 - mu);
    NFGS_LAUNCH_CHECK();
  }
  NFGS_TRY(closure_state_pass(p, bnd, k0 ? k0->kind : 0, phi_out, az_out, b_out, ws->jp,
                              nullptr, ws->partials, ws->s_bcz, ws->hat0, s));
  return nfgs_reduce(p, ws->partials, ws->metrics, stream);
}

extern "C" int nfgs_advance(const nfgs_plan* p, const nfgs_boundary* bnd,
                            const nfgs_workspace* ws, void* phi_out, void* az_out, void* b_out,
                            int zero_source, void* stream) {
  return nfgs_advance_ex(p, bnd, ws, phi_out, az_out, b_out, zero_source, NFGS_POLOIDAL_DIRECT,
                         stream);
}
