// Batched z-direction Thomas sweeps in modal space.
//
// One lane owns one horizontal mode and walks its column of nz planes; the
// lanes of a wavefront hold consecutive modes, so every plane access is one
// contiguous 256-byte (f32) run.  The forward elimination stores c' and d'
// into caller scratch (c') and the output (d'); the back substitution then
// overwrites the output with the solution.  Arithmetic follows the
// reference recurrences row by row (nfgs_solver.py:5579-5624, :567-608).

#include "plan.h"

namespace nfgs {
namespace {

constexpr int KB = 32;

// u_r = d'_r - c'_r u_{r+1} for r = top .. 0 (row r lives on plane r+1);
// blocks of KB rows load d' and c' before any store of the block.
template <typename T>
__device__ __forceinline__ void back_substitute(T* out, const T* cs, long plane, long m, int top,
                                                T u) {
  for (int rb = top; rb >= 0; rb -= KB) {
    T dv[KB], cv[KB];
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb - i;
      if (r >= 0) {
        const long o = (long)(r + 1) * plane + m;
        dv[i] = out[o];
        cv[i] = cs[o];
      }
    }
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb - i;
      if (r >= 0) {
        u = dv[i] - cv[i] * u;
        out[(long)(r + 1) * plane + m] = u;
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void robin_sweep_kernel(
    const T* __restrict__ src, T* __restrict__ out, T* __restrict__ cs, const T* __restrict__ lam,
    const T* __restrict__ kk, const T* __restrict__ bh, const T* __restrict__ qh, int nz,
    long plane, T inv_dz2, T two_inv_dz2, T dz, T dz2, T two_dz, T rho, T one_m_rho, T kb,
    T fu_rho, T rho_over_dz, double* zm_partial, double norm) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= plane) return;
  if (m == 0) {
    // The (0,0) mode carries no correction; it is the zero-mode diagnostic.
    double acc = 0.0;
    for (int z = 0; z < nz; ++z) {
      const double v = (double)src[(long)z * plane] / norm;
      acc += v * v;
      out[(long)z * plane] = T(0);
    }
    if (zm_partial) *zm_partial = acc;
    return;
  }
  const T l = lam[m];
  const T lo = inv_dz2, up = inv_dz2;
  const T diag = l - two_inv_dz2;
  const T b = bh[m];
  T cp, dp;
  {
    const T s1 = src[plane + m] / l;
    T fd, fu, r0;
    if (rho > T(0)) {
      fd = one_m_rho * l - two_inv_dz2;
      fu = fu_rho;
      T fr = one_m_rho * s1;
      fr = fr - rho_over_dz * qh[m];
      r0 = fr - kb * b;
    } else {
      fd = diag;
      fu = up;
      r0 = s1 - lo * b;
    }
    cp = fu / fd;
    dp = r0 / fd;
    cs[plane + m] = cp;
    out[plane + m] = dp;
  }
  const int nu = nz - 1;
  // Rows are processed in blocks whose loads are all issued before the
  // block's stores, so KB row loads per lane are in flight at once.
  for (int rb = 1; rb < nu - 1; rb += KB) {
    T sv[KB];
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb + i;
      sv[i] = r < nu - 1 ? src[(long)(r + 1) * plane + m] : T(0);
    }
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb + i;
      if (r < nu - 1) {
        const T s = sv[i] / l;
        const T den = diag - lo * cp;
        cp = up / den;
        dp = (s - lo * dp) / den;
        cs[(long)(r + 1) * plane + m] = cp;
        out[(long)(r + 1) * plane + m] = dp;
      }
    }
  }
  // Top Robin row with u_{N-2} eliminated through the PDE at plane N-1.
  {
    const T s = src[(long)(nz - 2) * plane + m] / l;
    const T rl = (-l) * dz2 - T(2);
    const T rd = T(2) + two_dz * kk[m];
    const T rr = T(0) - s * dz2;
    const T den = rd - rl * cp;
    dp = (rr - rl * dp) / den;
  }
  T u = dp;
  out[(long)(nz - 1) * plane + m] = u;
  back_substitute(out, cs, plane, m, nu - 2, u);
  out[m] = b;
}

template <typename T>
__global__ __launch_bounds__(256) void dirichlet_sweep_kernel(
    const T* __restrict__ src, T* __restrict__ out, T* __restrict__ cs, const T* __restrict__ lamd,
    const T* __restrict__ bh, int nz, long plane, T inv_dz2, T two_inv_dz2) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= plane) return;
  const T l = lamd[m];
  const T a = inv_dz2, c = inv_dz2;
  const T bd = l - two_inv_dz2;
  const T u0 = bh[m];
  const int n = nz - 2;
  auto rhs = [&](int z) { return (-src[(long)z * plane + m]) / l; };
  if (n == 1) {
    out[plane + m] = (rhs(1) - a * u0) / bd;
  } else {
    T cp = c / bd;
    T dp = (rhs(1) - a * u0) / bd;
    cs[plane + m] = cp;
    out[plane + m] = dp;
    for (int ib = 1; ib < n; ib += KB) {
      T sv[KB];
#pragma unroll
      for (int q = 0; q < KB; ++q) {
        const int i = ib + q;
        sv[q] = i < n ? src[(long)(i + 1) * plane + m] : T(0);
      }
#pragma unroll
      for (int q = 0; q < KB; ++q) {
        const int i = ib + q;
        if (i < n) {
          const T d = (-sv[q]) / l;
          const T ce = (i == n - 1) ? T(0) : c;
          const T den = bd - a * cp;
          cp = ce / den;
          dp = (d - a * dp) / den;
          cs[(long)(i + 1) * plane + m] = cp;
          out[(long)(i + 1) * plane + m] = dp;
        }
      }
    }
    back_substitute(out, cs, plane, m, n - 2, dp);
  }
  out[m] = u0;
  out[(long)(nz - 1) * plane + m] = T(0);
}

template <typename T>
int robin_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* src, void* out, void* scr,
            double* partials, hipStream_t s) {
  const nfgs_grid_desc& d = p->d;
  const double dz = d.dz, dz2 = dz * dz, rho = d.rho;
  const double norm = 4.0 * (d.nx - 1) * (d.ny - 1);
  double* zm = partials ? partials + (long)NFGS_Q_ZERO_MODE * p->pblocks : nullptr;
  const unsigned nb = (unsigned)((p->plane + 255) / 256);
  hipLaunchKernelGGL((robin_sweep_kernel<T>), dim3(nb), dim3(256), 0, s,
                     static_cast<const T*>(src), static_cast<T*>(out), static_cast<T*>(scr),
                     static_cast<const T*>(p->lam), static_cast<const T*>(p->kk),
                     static_cast<const T*>(bnd->phi_bottom_hat), static_cast<const T*>(bnd->q_hat),
                     d.nz, p->plane, T(1.0 / dz2), T(2.0 / dz2), T(dz), T(dz2), T(2.0 * dz),
                     T(rho), T(1.0 - rho), T((1.0 + 0.5 * rho) / dz2), T((1.0 - 0.5 * rho) / dz2),
                     T(rho / dz), zm, norm);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

template <typename T>
int dirichlet_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* src, void* out,
                void* scr, hipStream_t s) {
  const double dz2 = p->d.dz * p->d.dz;
  const unsigned nb = (unsigned)((p->iplane + 255) / 256);
  hipLaunchKernelGGL((dirichlet_sweep_kernel<T>), dim3(nb), dim3(256), 0, s,
                     static_cast<const T*>(src), static_cast<T*>(out), static_cast<T*>(scr),
                     static_cast<const T*>(p->lamd), static_cast<const T*>(bnd->az_bottom_hat),
                     p->d.nz, p->iplane, T(1.0 / dz2), T(2.0 / dz2));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace
}  // namespace nfgs

extern "C" int nfgs_robin_solve(const nfgs_plan* p, const nfgs_boundary* bnd, const void* src_hat,
                                void* out_hat, void* scratch, double* partials, void* stream) {
  NFGS_REQUIRE(p && bnd && src_hat && out_hat && scratch, NFGS_ERR_ARG,
               "robin_solve: null argument");
  NFGS_REQUIRE(bnd->phi_bottom_hat && (p->d.rho == 0.0 || bnd->q_hat), NFGS_ERR_STATE,
               "robin_solve: boundary tables are not set");
  NFGS_REQUIRE(src_hat != out_hat && out_hat != scratch, NFGS_ERR_ARG,
               "robin_solve: buffers must be distinct");
  if (p->d.dtype == NFGS_F32)
    return nfgs::robin_t<float>(p, bnd, src_hat, out_hat, scratch, partials,
                                nfgs::as_stream(stream));
  return nfgs::robin_t<double>(p, bnd, src_hat, out_hat, scratch, partials,
                               nfgs::as_stream(stream));
}

extern "C" int nfgs_dirichlet_solve(const nfgs_plan* p, const nfgs_boundary* bnd,
                                    const void* src_hat, void* out_hat, void* scratch,
                                    void* stream) {
  NFGS_REQUIRE(p && bnd && src_hat && out_hat && scratch, NFGS_ERR_ARG,
               "dirichlet_solve: null argument");
  NFGS_REQUIRE(bnd->az_bottom_hat, NFGS_ERR_STATE, "dirichlet_solve: boundary tables are not set");
  NFGS_REQUIRE(src_hat != out_hat && out_hat != scratch, NFGS_ERR_ARG,
               "dirichlet_solve: buffers must be distinct");
  if (p->d.dtype == NFGS_F32)
    return nfgs::dirichlet_t<float>(p, bnd, src_hat, out_hat, scratch, nfgs::as_stream(stream));
  return nfgs::dirichlet_t<double>(p, bnd, src_hat, out_hat, scratch, nfgs::as_stream(stream));
}
