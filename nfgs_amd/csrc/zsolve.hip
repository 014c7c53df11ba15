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

// Rows are processed in blocks of KB; the loads of block b+1 are issued
// before block b is computed, so two blocks of loads per lane are in flight.
// Loads past the row range are clamped to the last row and never used.
//
// Forward rows r = lo .. hi (row r lives on plane r+1): step(r, v) consumes
// the loaded value of row r.
template <typename T, typename Step>
__device__ __forceinline__ void forward_rows(const T* __restrict__ src, long plane, long m, int lo,
                                             int hi, Step step) {
  if (hi < lo) return;
  T va[KB], vb[KB];
  auto load = [&](T (&v)[KB], int rb) {
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb + i < hi ? rb + i : hi;
      v[i] = src[(long)(r + 1) * plane + m];
    }
  };
  auto run = [&](const T (&v)[KB], int rb) {
#pragma unroll
    for (int i = 0; i < KB; ++i)
      if (rb + i <= hi) step(rb + i, v[i]);
  };
  load(va, lo);
  for (int rb = lo; rb <= hi; rb += 2 * KB) {
    load(vb, rb + KB);
    run(va, rb);
    load(va, rb + 2 * KB);
    run(vb, rb + KB);
  }
}

// u_r = d'_r - c'_r u_{r+1} for r = top .. 0 (row r lives on plane r+1).
template <typename T>
__device__ __forceinline__ void back_substitute(T* out, const T* cs, long plane, long m, int top,
                                                T u) {
  if (top < 0) return;
  T da[KB], ca[KB], db[KB], cb[KB];
  auto load = [&](T (&d)[KB], T (&c)[KB], int rb) {
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb - i > 0 ? rb - i : 0;
      const long o = (long)(r + 1) * plane + m;
      d[i] = out[o];
      c[i] = cs[o];
    }
  };
  auto run = [&](const T (&d)[KB], const T (&c)[KB], int rb) {
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int r = rb - i;
      if (r >= 0) {
        u = d[i] - c[i] * u;
        out[(long)(r + 1) * plane + m] = u;
      }
    }
  };
  load(da, ca, top);
  for (int rb = top; rb >= 0; rb -= 2 * KB) {
    load(db, cb, rb - KB);
    run(da, ca, rb);
    load(da, ca, rb - 2 * KB);
    run(db, cb, rb - KB);
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
    // rows r = -1 .. nz-2 cover planes 0 .. nz-1 (blocked like every lane)
    double acc = 0.0;
    forward_rows(src, plane, m, -1, nz - 2, [&](int r, T sv) {
      const double v = (double)sv / norm;
      acc += v * v;
      out[(long)(r + 1) * plane] = T(0);
    });
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
  forward_rows(src, plane, m, 1, nu - 2, [&](int r, T sv) {
    const T s = sv / l;
    const T den = diag - lo * cp;
    cp = up / den;
    dp = (s - lo * dp) / den;
    cs[(long)(r + 1) * plane + m] = cp;
    out[(long)(r + 1) * plane + m] = dp;
  });
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
    forward_rows(src, plane, m, 1, n - 1, [&](int i, T sv) {
      const T d = (-sv) / l;
      const T ce = (i == n - 1) ? T(0) : c;
      const T den = bd - a * cp;
      cp = ce / den;
      dp = (d - a * dp) / den;
      cs[(long)(i + 1) * plane + m] = cp;
      out[(long)(i + 1) * plane + m] = dp;
    });
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
