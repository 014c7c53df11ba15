// Shared definitions for the NFGS gfx950 kernels and their C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/nfgs_hip.h"

namespace nfgs {

// Thread-local last-error text returned by nfgs_last_error().
void set_error(const std::string& msg);

#define NFGS_HIP_CHECK(expr)                                                              \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) {                                                               \
      ::nfgs::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));              \
      return NFGS_ERR_HIP;                                                                \
    }                                                                                     \
  } while (0)

#define NFGS_REQUIRE(cond, code, msg)                                                     \
  do {                                                                                    \
    if (!(cond)) {                                                                        \
      ::nfgs::set_error(msg);                                                             \
      return code;                                                                        \
    }                                                                                     \
  } while (0)

// Launch-then-check helper: kernel launch errors surface as NFGS_ERR_HIP.
#define NFGS_LAUNCH_CHECK()                                                               \
  do {                                                                                    \
    hipError_t _e = hipGetLastError();                                                    \
    if (_e != hipSuccess) {                                                               \
      ::nfgs::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));         \
      return NFGS_ERR_HIP;                                                                \
    }                                                                                     \
  } while (0)

// Second-order centred first derivative with one-sided ends (needs n >= 3),
// evaluated with the operand order of the reference stencil.
template <typename T>
__device__ __forceinline__ T fd1(const T* f, long s, int i, int n, T h) {
  if (i > 0 && i < n - 1) return (f[s] - f[-s]) / (T(2) * h);
  if (i == 0) return (T(-3) * f[0] + T(4) * f[s] - f[2 * s]) / (T(2) * h);
  return (T(3) * f[0] - T(4) * f[-s] + f[-2 * s]) / (T(2) * h);
}

// Same derivative on a value accessor: g(o) returns the sample at offset o.
template <typename T, typename G>
__device__ __forceinline__ T fd1g(G g, int i, int n, T h) {
  if (i > 0 && i < n - 1) return (g(1) - g(-1)) / (T(2) * h);
  if (i == 0) return (T(-3) * g(0) + T(4) * g(1) - g(2)) / (T(2) * h);
  return (T(3) * g(0) - T(4) * g(-1) + g(-2)) / (T(2) * h);
}

// The same derivatives scaled by a precomputed r = 1/(2h) instead of a
// division by 2h (the hot stencils; differs from the division by at most an
// ulp per derivative).
template <typename T>
__device__ __forceinline__ T fd1r(const T* f, long s, int i, int n, T r) {
  if (i > 0 && i < n - 1) return (f[s] - f[-s]) * r;
  if (i == 0) return (T(-3) * f[0] + T(4) * f[s] - f[2 * s]) * r;
  return (T(3) * f[0] - T(4) * f[-s] + f[-2 * s]) * r;
}
template <typename T, typename G>
__device__ __forceinline__ T fd1gr(G g, int i, int n, T r) {
  if (i > 0 && i < n - 1) return (g(1) - g(-1)) * r;
  if (i == 0) return (T(-3) * g(0) + T(4) * g(1) - g(2)) * r;
  return (T(3) * g(0) - T(4) * g(-1) + g(-2)) * r;
}

// Square root and quotient of the convergence monitors (sigma_J, |B|).  In
// FP32 they use the hardware instructions (about 1 ulp, like the reference's
// own FP32 monitors, without the correctly rounded sequences); in FP64 they
// are the IEEE operations.
__device__ __forceinline__ float msqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float mdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double mdiv(double a, double b) { return a / b; }

// Mirror-Neumann 5-point lateral Laplacian at (y, x) of a (ny, nx) plane with
// row stride ld, accumulated in the reference's term order.
template <typename T>
__device__ __forceinline__ T lap_xy_at(const T* p, int y, int x, int ny, int nx, long ld,
                                       T idx2, T idy2) {
  const T* c = p + (long)y * ld + x;
  T acc = T(0);
  if (x > 0 && x < nx - 1) acc += (c[1] - T(2) * c[0] + c[-1]) * idx2;
  if (y > 0 && y < ny - 1) acc += (c[ld] - T(2) * c[0] + c[-ld]) * idy2;
  if (y == 0) acc += (T(2) * c[ld] - T(2) * c[0]) * idy2;
  if (y == ny - 1) acc += (T(2) * c[-ld] - T(2) * c[0]) * idy2;
  if (x == 0) acc += (T(2) * c[1] - T(2) * c[0]) * idx2;
  if (x == nx - 1) acc += (T(2) * c[-1] - T(2) * c[0]) * idx2;
  return acc;
}

// Cross-lane reads without LDS traffic: DPP inside a row of 16 lanes and
// ds_swizzle (bit-mask mode) inside a group of 32.  Every lane of the wave
// must be active where they are used.
constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kDppRor4 = 0x124;  // row_ror:4
constexpr int kDppRor8 = 0x128;  // row_ror:8 (= lane ^ 8 inside the row)
constexpr int kSwzXor16 = 0x1F | (0x10 << 10);  // and 0x1f, or 0, xor 0x10: lane ^ 16
template <int C>
__device__ __forceinline__ unsigned dpp_u32(unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, C, 0xF, 0xF, false);
}
template <int C>
__device__ __forceinline__ double dpp_f64(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = dpp_u32<C>((unsigned)u), hi = dpp_u32<C>((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ unsigned swz16_u32(unsigned x) {
  return (unsigned)__builtin_amdgcn_ds_swizzle((int)x, kSwzXor16);
}
__device__ __forceinline__ double swz16_f64(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = swz16_u32((unsigned)u), hi = swz16_u32((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double readlane_f64(double x, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// largest value over the 16 lanes of a row, in every lane of the row
__device__ __forceinline__ unsigned row_max_u32(unsigned x) {
  x = max(x, dpp_u32<kDppXor1>(x));
  x = max(x, dpp_u32<kDppXor2>(x));
  x = max(x, dpp_u32<kDppRor4>(x));
  return max(x, dpp_u32<kDppRor8>(x));
}
// sum over the 8 lanes of a row whose index has the parity of this lane's
__device__ __forceinline__ double row_parity_sum_f64(double x) {
  x += dpp_f64<kDppXor2>(x);
  x += dpp_f64<kDppRor4>(x);
  return x + dpp_f64<kDppRor8>(x);
}
// sum over the 64 lanes of a wavefront (fixed order), in every lane
__device__ __forceinline__ double wave_sum_f64(double x) {
  x += dpp_f64<kDppXor1>(x);
  x = row_parity_sum_f64(x);
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}

// Block-wide sum of NQ doubles in a fixed order; result valid in thread 0.
// Every thread of the block takes part; blockDim.x must be a multiple of 64
// and at most 1024.
template <int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double* lds /* [NQ*16] */) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_sum_f64(v[q]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) lds[q * 16 + wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[q * 16 + w];
      v[q] = s;
    }
  }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Workgroup barrier that orders LDS only.  __syncthreads() is a release
// fence on all address spaces, which waits for every outstanding global load
// and store (vmcnt(0)) and so drains register prefetches at each barrier.
// Kernels that never read, within one launch, global memory written by
// another thread of that launch only need LDS ordering.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// element at byte offset o of a plane (uniform base + 32-bit lane offset)
template <typename T>
__device__ __forceinline__ T ld(const T* base, unsigned o) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + o);
}
template <typename T>
__device__ __forceinline__ void st(T* base, unsigned o, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + o) = v;
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs, each
// with its own L2.  xcd_tile maps launch-order workgroup w of n to a logical
// tile so that every XCD owns one contiguous range of tiles: neighbouring
// tiles (shared halo and cache lines) then meet in the same L2.  A bijection
// on [0, n) for any n.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int w, int n) {
  const int per = n / kXcds, extra = n % kXcds;
  const int x = w % kXcds, i = w / kXcds;
  return x * per + (x < extra ? x : extra) + i;
}

}  // namespace nfgs
