// Method-independent scoring of output fields: the common second-order curl
// and divergence (nfgs/diagnostics/common_fd2.py:23-61), sigma_J on a box
// (common_fd2.py:64-83, util/metrics.py:66-92), the reference-field
// agreement metrics (util/metrics.py:128-176) and the normalised divergence
// of the controlled evaluation (recompute_td_referee_diagnostics.py:344-400).  FP64 on (3, nz, ny, nx)
// component-first arrays of any size, independent of a plan, as the
// reference evaluates them ("common_fd2_float64_full_volume_then_crop").

#include <algorithm>
#include <mutex>
#include <string>

#include "nfgs_common.h"

namespace nfgs {
namespace {

// d/d(axis) at index i of n samples with stride s: centred inside, the
// second-order one-sided forms at the ends; n = 2 takes the forward
// difference at both nodes and n = 1 gives 0 (common_fd2.py:23-37).
__device__ __forceinline__ double fd2(const double* f, long s, int i, int n, double h) {
  if (n < 3) return n == 1 ? 0.0 : (f[(1 - i) * s] - f[-i * s]) / h;
  if (i == 0) return (-3.0 * f[0] + 4.0 * f[s] - f[2 * s]) / (2.0 * h);
  if (i == n - 1) return (3.0 * f[0] - 4.0 * f[-s] + f[-2 * s]) / (2.0 * h);
  return (f[s] - f[-s]) / (2.0 * h);
}

__global__ __launch_bounds__(256) void curl_div_kernel(const double* __restrict__ b, int nz, int ny,
                                                       int nx, double dx, double dy, double dz,
                                                       double* __restrict__ j,
                                                       double* __restrict__ div) {
  const long N = (long)nz * ny * nx, pl = (long)ny * nx;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int z = (int)(i / pl);
  const long r = i - z * pl;
  const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
  const double *bx = b + i, *by = b + N + i, *bz = b + 2 * N + i;
  const double dbz_dy = fd2(bz, nx, y, ny, dy), dby_dz = fd2(by, pl, z, nz, dz);
  const double dbx_dz = fd2(bx, pl, z, nz, dz), dbz_dx = fd2(bz, 1, x, nx, dx);
  const double dby_dx = fd2(by, 1, x, nx, dx), dbx_dy = fd2(bx, nx, y, ny, dy);
  if (j) {
    j[i] = dbz_dy - dby_dz;
    j[N + i] = dbx_dz - dbz_dx;
    j[2 * N + i] = dby_dx - dbx_dy;
  }
  if (div) div[i] = fd2(bx, 1, x, nx, dx) + fd2(by, nx, y, ny, dy) + fd2(bz, pl, z, nz, dz);
}

// Per-block sums over the box [z0,z1) x [y0,y1) x [x0,x1) (NFGS_SC_* order).
constexpr int kScores = NFGS_SC_COUNT;
__global__ __launch_bounds__(256) void scores_kernel(const double* __restrict__ bref,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ j,
                                                     const double* __restrict__ dv, int nz, int ny,
                                                     int nx, int z0, int y0, int x0, int bz_n,
                                                     int by_n, int bx_n, double b2min,
                                                     double* __restrict__ part) {
  __shared__ double red[kScores * 16];
  const long N = (long)nz * ny * nx, pl = (long)ny * nx;
  const long M = (long)bz_n * by_n * bx_n, bpl = (long)by_n * bx_n;
  double acc[kScores];
#pragma unroll
  for (int q = 0; q < kScores; ++q) acc[q] = 0.0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < M;
       k += (long)gridDim.x * blockDim.x) {
    const int z = (int)(k / bpl);
    const long r = k - z * bpl;
    const int y = (int)(r / bx_n), x = (int)(r - (long)y * bx_n);
    const long i = (long)(z0 + z) * pl + (long)(y0 + y) * nx + (x0 + x);
    const double s0 = b[i], s1 = b[N + i], s2 = b[2 * N + i];
    const double bn = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
    acc[NFGS_SC_BN] += bn;
    if (dv) acc[NFGS_SC_DIVABS] += fabs(dv[i]);
    if (bref) {
      const double r0 = bref[i], r1 = bref[N + i], r2 = bref[2 * N + i];
      const double rn = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
      const double dot = r0 * s0 + r1 * s1 + r2 * s2;
      const double e0 = s0 - r0, e1 = s1 - r1, e2 = s2 - r2;
      const double en = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
      acc[NFGS_SC_DOT] += dot;
      acc[NFGS_SC_REF2] += r0 * r0 + r1 * r1 + r2 * r2;
      acc[NFGS_SC_SOL2] += s0 * s0 + s1 * s1 + s2 * s2;
      acc[NFGS_SC_COS] += dot / (rn * bn + 1e-12);
      acc[NFGS_SC_DIFF] += en;
      acc[NFGS_SC_REFN] += rn;
      acc[NFGS_SC_DIFF_REL] += en / (rn + 1e-12);
    }
    if (j) {
      const double j0 = j[i], j1 = j[N + i], j2 = j[2 * N + i];
      const double c0 = j1 * s2 - j2 * s1, c1 = j2 * s0 - j0 * s2, c2 = j0 * s1 - j1 * s0;
      acc[NFGS_SC_JXB] += sqrt(c0 * c0 + c1 * c1 + c2 * c2) / fmax(bn, b2min);
      acc[NFGS_SC_JN] += sqrt(j0 * j0 + j1 * j1 + j2 * j2);
    }
  }
  block_sum<kScores>(acc, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < kScores; ++q) part[(long)q * gridDim.x + blockIdx.x] = acc[q];
}

__global__ __launch_bounds__(256) void scores_final_kernel(const double* __restrict__ part,
                                                           int nb, double* __restrict__ out) {
  __shared__ double red[kScores * 16];
  double acc[kScores];
#pragma unroll
  for (int q = 0; q < kScores; ++q) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[(long)q * nb + k];
    acc[q] = s;
  }
  block_sum<kScores>(acc, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < kScores; ++q) out[q] = acc[q];
}

constexpr int kScoreBlocks = 1024;

}  // namespace
}  // namespace nfgs

using namespace nfgs;

extern "C" int nfgs_fd2_curl_div(const double* b, int nz, int ny, int nx, double dx, double dy,
                                 double dz, double* j, double* div, void* stream) {
  NFGS_REQUIRE(b && (j || div) && nz >= 1 && ny >= 1 && nx >= 1, NFGS_ERR_ARG,
               "fd2_curl_div: bad argument");
  NFGS_REQUIRE(dx > 0 && dy > 0 && dz > 0, NFGS_ERR_ARG, "spacing must be positive");
  const long N = (long)nz * ny * nx;
  hipLaunchKernelGGL(curl_div_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                     as_stream(stream), b, nz, ny, nx, dx, dy, dz, j, div);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

extern "C" long nfgs_scores_workspace_bytes(void) {
  return (long)kScores * kScoreBlocks * (long)sizeof(double);
}

extern "C" int nfgs_field_scores(const double* b_ref, const double* b, const double* j,
                                 const double* div, int nz, int ny, int nx, int z0, int z1, int y0,
                                 int y1, int x0, int x1, double b2min, double* workspace,
                                 double* out, void* stream) {
  NFGS_REQUIRE(b && out && workspace, NFGS_ERR_ARG, "field_scores: bad argument");
  NFGS_REQUIRE(0 <= z0 && z0 < z1 && z1 <= nz && 0 <= y0 && y0 < y1 && y1 <= ny && 0 <= x0 &&
                   x0 < x1 && x1 <= nx,
               NFGS_ERR_ARG, "field_scores: the box is empty or outside the field");
  const long M = (long)(z1 - z0) * (y1 - y0) * (x1 - x0);
  const int nb = (int)std::min<long>(kScoreBlocks, (M + 255) / 256);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(scores_kernel, dim3(nb), dim3(256), 0, s, b_ref, b, j, div, nz, ny, nx, z0,
                     y0, x0, z1 - z0, y1 - y0, x1 - x0, b2min, workspace);
  NFGS_LAUNCH_CHECK();
  hipLaunchKernelGGL(scores_final_kernel, dim3(1), dim3(256), 0, s, workspace, nb, out);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

// ---- relative error of an outward bottom field (the CCI comparison) -------
//
// One workgroup walks the window twice: the first pass sums |c-b|^2 and
// |b|^2 and finds max|b|; the second counts, per threshold
// t_k = frac_k * (max|b| + eps), the pixels with |b| > t_k and sums
// |c-b| / |b| over them.  Fixed-order tree sums make the result independent
// of scheduling.

namespace nfgs {
namespace {

constexpr int kMaxCuts = 16;
constexpr int kOutwardThreads = 1024;

struct CutList {
  double frac[kMaxCuts];
  int n;
};

// sum (or max) of one value per thread over the workgroup, valid in every
// thread afterwards; `lds` holds kOutwardThreads doubles
__device__ double tree_reduce(double v, double* lds, bool take_max) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int w = kOutwardThreads / 2; w > 0; w >>= 1) {
    if (t < w) lds[t] = take_max ? fmax(lds[t], lds[t + w]) : lds[t] + lds[t + w];
    __syncthreads();
  }
  const double r = lds[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kOutwardThreads) void outward_errors_kernel(
    const double* __restrict__ b, const double* __restrict__ c, int ny, int nx, int y0, int x0,
    int wy, int wx, CutList cuts, double eps, double* __restrict__ out) {
  __shared__ double lds[kOutwardThreads];
  const long pl = (long)ny * nx;
  const int M = wy * wx;
  double d2 = 0.0, b2 = 0.0, bmax = 0.0;
  for (int k = threadIdx.x; k < M; k += kOutwardThreads) {
    const int y = k / wx, x = k - y * wx;
    const long i = (long)(y0 + y) * nx + (x0 + x);
    const double t0 = b[i], t1 = b[pl + i], t2 = b[2 * pl + i];
    const double e0 = c[i] - t0, e1 = c[pl + i] - t1, e2 = c[2 * pl + i] - t2;
    const double m2 = t0 * t0 + t1 * t1 + t2 * t2;
    d2 += e0 * e0 + e1 * e1 + e2 * e2;
    b2 += m2;
    bmax = fmax(bmax, sqrt(m2));
  }
  d2 = tree_reduce(d2, lds, false);
  b2 = tree_reduce(b2, lds, false);
  bmax = tree_reduce(bmax, lds, true);
  double cnt[kMaxCuts], rel[kMaxCuts], thr[kMaxCuts];
  for (int q = 0; q < kMaxCuts; ++q) {
    cnt[q] = 0.0;
    rel[q] = 0.0;
    thr[q] = q < cuts.n ? cuts.frac[q] * (bmax + eps) : 0.0;
  }
  for (int k = threadIdx.x; k < M; k += kOutwardThreads) {
    const int y = k / wx, x = k - y * wx;
    const long i = (long)(y0 + y) * nx + (x0 + x);
    const double t0 = b[i], t1 = b[pl + i], t2 = b[2 * pl + i];
    const double e0 = c[i] - t0, e1 = c[pl + i] - t1, e2 = c[2 * pl + i] - t2;
    const double mag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    const double q_rel = sqrt(e0 * e0 + e1 * e1 + e2 * e2) / mag;
    for (int q = 0; q < cuts.n; ++q) {
      const bool keep = mag > thr[q];
      cnt[q] += keep ? 1.0 : 0.0;
      rel[q] += keep ? q_rel : 0.0;
    }
  }
  for (int q = 0; q < cuts.n; ++q) {
    cnt[q] = tree_reduce(cnt[q], lds, false);
    rel[q] = tree_reduce(rel[q], lds, false);
  }
  if (threadIdx.x == 0) {
    out[0] = d2;
    out[1] = b2;
    out[2] = bmax;
    out[3] = (double)M;
    for (int q = 0; q < cuts.n; ++q) {
      out[4 + 2 * q] = cnt[q];
      out[5 + 2 * q] = rel[q];
    }
  }
}

}  // namespace
}  // namespace nfgs

extern "C" int nfgs_outward_errors(const double* b, const double* c, int ny, int nx, int y0,
                                   int y1, int x0, int x1, const double* frac, int ncut,
                                   double eps, double* out, void* stream) {
  NFGS_REQUIRE(b && c && out && (ncut == 0 || frac), NFGS_ERR_ARG,
               "outward_errors: bad argument");
  NFGS_REQUIRE(ncut >= 0 && ncut <= kMaxCuts, NFGS_ERR_ARG,
               "outward_errors: at most 16 thresholds");
  NFGS_REQUIRE(0 <= y0 && y0 < y1 && y1 <= ny && 0 <= x0 && x0 < x1 && x1 <= nx, NFGS_ERR_ARG,
               "outward_errors: the window is empty or outside the plane");
  NFGS_REQUIRE((long)(y1 - y0) * (x1 - x0) < (1L << 30), NFGS_ERR_ARG,
               "outward_errors: window too large");
  CutList cuts{};
  cuts.n = ncut;
  for (int q = 0; q < ncut; ++q) cuts.frac[q] = frac[q];
  hipLaunchKernelGGL(outward_errors_kernel, dim3(1), dim3(kOutwardThreads), 0, as_stream(stream),
                     b, c, ny, nx, y0, x0, y1 - y0, x1 - x0, cuts, eps, out);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

// ---- field scores without a workspace argument ----------------------------
//
// nfgs_field_scores_static runs the kernels of nfgs_field_scores with their
// per-block partials in `scratch`, a device array of this library: it
// allocates nothing, for callers that cannot allocate (code under graph
// capture, tight serving loops).  The partials of a launch are written for
// blocks [0, nb) before the final kernel reads exactly those, so nothing needs
// clearing between calls; calls that overlap on one GPU share the array and
// race.

static __device__ __align__(256) double scratch[(long)nfgs::kScores * nfgs::kScoreBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device.  Host code uses this address only; the array's own name is a host
// stand-in and must never be passed to a kernel or a copy.
int scratch_address(double** out) {
  int dev = -1;
  NFGS_HIP_CHECK(hipGetDevice(&dev));
  NFGS_REQUIRE(dev >= 0 && dev < kMaxDevices, NFGS_ERR_ARG,
               "field_scores_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      nfgs::set_error(std::string("field_scores_static: hipGetSymbolAddress: ") +
                      hipGetErrorString(e));
      return NFGS_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return NFGS_OK;
}
}  // namespace

extern "C" long nfgs_field_scores_static_capacity(void) {
  return (long)kScores * kScoreBlocks;
}

extern "C" int nfgs_field_scores_static(const double* b_ref, const double* b, const double* j,
                                        const double* div, int nz, int ny, int nx, int z0, int z1,
                                        int y0, int y1, int x0, int x1, double b2min, double* out,
                                        void* stream) {
  NFGS_REQUIRE(b && out, NFGS_ERR_ARG, "field_scores_static: bad argument");
  NFGS_REQUIRE(0 <= z0 && z0 < z1 && z1 <= nz && 0 <= y0 && y0 < y1 && y1 <= ny && 0 <= x0 &&
                   x0 < x1 && x1 <= nx,
               NFGS_ERR_ARG, "field_scores_static: the box is empty or outside the field");
  const long M = (long)(z1 - z0) * (y1 - y0) * (x1 - x0);
  const long nb = std::min<long>(kScoreBlocks, (M + 255) / 256);
  NFGS_REQUIRE(nb * kScores <= nfgs_field_scores_static_capacity(), NFGS_ERR_ARG,
               "field_scores_static: the partials exceed the static scratch");
  double* part = nullptr;
  const int rc = scratch_address(&part);
  if (rc != NFGS_OK) return rc;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(scores_kernel, dim3((unsigned)nb), dim3(256), 0, s, b_ref, b, j, div, nz, ny,
                     nx, z0, y0, x0, z1 - z0, y1 - y0, x1 - x0, b2min, part);
  NFGS_LAUNCH_CHECK();
  hipLaunchKernelGGL(scores_final_kernel, dim3(1), dim3(256), 0, s, part, (int)nb, out);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}
