// Small device helpers shared by the setup and diagnostic entry points:
// point maps, deterministic two-stage reductions and plane-stack transform
// compositions built on the paired transform kernel.
#pragma once

#include "plan.h"

namespace nfgs {
namespace {

constexpr int kRedBlocks = 256;

template <typename F>
__global__ __launch_bounds__(256) void map_kernel(long n, F f) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) f(i);
}

template <typename F>
int launch_map(long n, F f, hipStream_t s) {
  if (n <= 0) return NFGS_OK;
  long nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL((map_kernel<F>), dim3((unsigned)nb), dim3(256), 0, s, n, f);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

// Two-stage deterministic reduction: op 0 = sum, op 1 = max.
template <typename F>
__global__ __launch_bounds__(256) void red1_kernel(long n, F f, double* part, int op) {
  __shared__ double sh[256];
  double v = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double x = f(i);
    v = op ? (x > v ? x : v) : v + x;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double a = sh[threadIdx.x], b = sh[threadIdx.x + w];
      sh[threadIdx.x] = op ? (b > a ? b : a) : a + b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void red2_kernel(const double* part, int nb, double* out, int op) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v = 0.0;
  for (int b = 0; b < nb; ++b) v = op ? (part[b] > v ? part[b] : v) : v + part[b];
  *out = v;
}

template <typename F>
int reduce_to(long n, F f, double* part, double* out, int op, hipStream_t s) {
  hipLaunchKernelGGL((red1_kernel<F>), dim3(kRedBlocks), dim3(256), 0, s, n, f, part, op);
  NFGS_LAUNCH_CHECK();
  hipLaunchKernelGGL(red2_kernel, dim3(1), dim3(64), 0, s, part, kRedBlocks, out, op);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

#define NFGS_TRY(expr)   \
  do {                   \
    int _rc = (expr);    \
    if (_rc) return _rc; \
  } while (0)

// ---------------------------------------------------------------------------
// Transform helpers on plane stacks (row stride nx, plane stride ny*nx).
// ---------------------------------------------------------------------------
template <typename T>
struct Ctx {
  const nfgs_plan* p;
  hipStream_t s;
  int nx, ny, nz;
  long pl;
};

// dct2 / idct2 = transform along x then along y (nfgs_solver.py:4354, :4459)
template <typename T>
int xy_pair(const Ctx<T>& c, int dir, const T* in, T* out, T* tmp, int nplanes) {
  NFGS_TRY(launch_transform(c.p, &c.p->dct_x, dir, in, 1, c.nx, c.pl, tmp, 1, c.nx, c.pl, c.ny,
                            nplanes, c.s));
  return launch_transform(c.p, &c.p->dct_y, dir, tmp, c.nx, 1, c.pl, out, c.nx, 1, c.pl, c.nx,
                          nplanes, c.s);
}

// out[:, 1:-1] = DST_x(DCT_y(in)[:, 1:-1])  (nfgs_solver.py:4420-4427)
template <typename T>
int dstx_dcty(const Ctx<T>& c, int dir, const T* in, T* out, T* tmp, int nplanes) {
  NFGS_TRY(launch_transform(c.p, &c.p->dct_y, dir, in, c.nx, 1, c.pl, tmp, c.nx, 1, c.pl, c.nx,
                            nplanes, c.s));
  NFGS_TRY(fill_zero(c.p, out, c.pl * nplanes, c.s));
  return launch_transform(c.p, &c.p->dst_x, dir, tmp + 1, 1, c.nx, c.pl, out + 1, 1, c.nx, c.pl,
                          c.ny, nplanes, c.s);
}

// out[1:-1, :] = DST_y(DCT_x(in)[1:-1, :])  (nfgs_solver.py:4429-4436)
template <typename T>
int dctx_dsty(const Ctx<T>& c, int dir, const T* in, T* out, T* tmp, int nplanes) {
  NFGS_TRY(launch_transform(c.p, &c.p->dct_x, dir, in, 1, c.nx, c.pl, tmp, 1, c.nx, c.pl, c.ny,
                            nplanes, c.s));
  NFGS_TRY(fill_zero(c.p, out, c.pl * nplanes, c.s));
  return launch_transform(c.p, &c.p->dst_y, dir, tmp + c.nx, c.nx, 1, c.pl, out + c.nx, c.nx, 1,
                          c.pl, c.nx, nplanes, c.s);
}

// Negated spectral gradient -grad_h of a field whose DCT2 is hat
// (nfgs_solver.py:4400-4418 with the sign of B_h = -grad_h folded in).
template <typename T>
int neg_spectral_grad(const Ctx<T>& c, const T* hat, T* gx_neg, T* gy_neg, T* t1, T* t2,
                      int nplanes) {
  const nfgs_plan* p = c.p;
  const T* ax = static_cast<const T*>(p->ax);
  const T* ay = static_cast<const T*>(p->ay);
  const int nx = c.nx, ny = c.ny;
  const long pl = c.pl;
  const long n = pl * nplanes;
  NFGS_TRY(launch_map(n, [=] __device__(long i) { t1[i] = ax[i % nx] * hat[i]; }, c.s));
  NFGS_TRY(launch_transform(p, &p->dct_y, NFGS_INVERSE, t1, nx, 1, pl, t2, nx, 1, pl, nx, nplanes,
                            c.s));
  NFGS_TRY(fill_zero(p, gx_neg, n, c.s));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_INVERSE, t2 + 1, 1, nx, pl, gx_neg + 1, 1, nx, pl,
                            ny, nplanes, c.s));
  NFGS_TRY(launch_map(n, [=] __device__(long i) { t1[i] = ay[(i % pl) / nx] * hat[i]; }, c.s));
  NFGS_TRY(launch_transform(p, &p->dct_x, NFGS_INVERSE, t1, 1, nx, pl, t2, 1, nx, pl, ny, nplanes,
                            c.s));
  NFGS_TRY(fill_zero(p, gy_neg, n, c.s));
  return launch_transform(p, &p->dst_y, NFGS_INVERSE, t2 + nx, nx, 1, pl, gy_neg + nx, nx, 1, pl,
                          nx, nplanes, c.s);
}

template <typename T>
__device__ __forceinline__ T d1_plane(const T* a, int y, int x, int ny, int nx, T h, bool along_x) {
  const long off = (long)y * nx + x;
  return along_x ? fd1(a + off, 1, x, nx, h) : fd1(a + off, (long)nx, y, ny, h);
}

}  // namespace
}  // namespace nfgs
