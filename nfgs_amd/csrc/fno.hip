// Learned warm start: the kernels of a small cosine-basis spectral neural
// operator (forward, backward and its optimiser step) that maps the gs_cfcf
// state (its B and parallel current J_par) and rho to a correction of J_par
// towards the converged state's.  One solver step from the corrected current
// (the exact Poisson solves) then gives the start state.  No counterpart in
// the reference, whose iterate() only starts from gs_cfcf
// (nfgs_solver.py:6613-6645).
//
// Network (fixed shape): lift 8 -> 32 channels; four spectral layers
//   v' = act(W v + b + E_x E_y E_z (R_mode (P_z P_y P_x v)))
// with P_a the truncated orthonormal DCT-II along axis a (16 modes, a dense
// n x 16 basis) and E_a = P_a^T, R a 32 x 32 channel mix per retained mode
// (16^3 modes), act = GELU except after the last layer; head 32 -> 64 (GELU)
// -> 3.  Weights are bfloat16 (FP32 master copies in the optimiser),
// activations and accumulation FP32.
//
// Layouts: a volume of C channels is (C, nz, ny, nx), x fastest.  The
// spectral chain keeps the mode index in place of the axis it replaces:
// after P_x (C, nz, ny, 16), after P_y (C, nz, 16, 16), after P_z
// (C, 16, 16, 16).  R is (mode, out, in) with mode = (kz*16 + ky)*16 + kx.

#include "util.h"

namespace nfgs {
namespace {

constexpr int kW = NFGS_FNO_WIDTH;
constexpr int kM = NFGS_FNO_MODES;
constexpr int kNM = kM * kM * kM;

__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even; a NaN stays a (quiet) NaN.
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// GELU(x) = x * Phi(x) with the normal distribution function Phi.
__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.7071067811865476f));
}
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.7071067811865476f)) +
         x * 0.3989422804014327f * expf(-0.5f * x * x);
}

// ---------------------------------------------------------------------------
// Pointwise layer: out[o][p] = act(sum_i W[o][i] in[i][p] + b[o] + s[o][p]),
// s = E_x(spec) when SPEC (spec is (COUT, nz*ny, 16)).  One thread per node,
// the weights staged once per block in LDS as FP32.
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int ACT, bool SPEC>
__global__ __launch_bounds__(256) void pw_forward(const float* __restrict__ in,
                                                  const uint16_t* __restrict__ w,
                                                  const float* __restrict__ bias,
                                                  const float* __restrict__ spec,
                                                  const float* __restrict__ bx, int nx, int rows,
                                                  float* __restrict__ out) {
  __shared__ float ws[COUT * CIN];
  __shared__ float bs[COUT];
  for (int k = threadIdx.x; k < COUT * CIN; k += 256) ws[k] = bf16f(w[k]);
  for (int k = threadIdx.x; k < COUT; k += 256) bs[k] = bias[k];
  __syncthreads();
  const long N = (long)rows * nx;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  float xin[CIN];
#pragma unroll
  for (int i = 0; i < CIN; ++i) xin[i] = in[(long)i * N + p];
  float sb[kM];
  long srow = 0;
  if constexpr (SPEC) {
    const long row = p / nx;
    const int x = (int)(p - row * nx);
#pragma unroll
    for (int k = 0; k < kM; ++k) sb[k] = bx[x * kM + k];
    srow = row * kM;
  }
  for (int o = 0; o < COUT; ++o) {
    float acc = bs[o];
#pragma unroll
    for (int i = 0; i < CIN; ++i) acc = fmaf(ws[o * CIN + i], xin[i], acc);
    if constexpr (SPEC) {
      const float4* s4 = reinterpret_cast<const float4*>(spec + (long)o * rows * kM + srow);
#pragma unroll
      for (int q = 0; q < kM / 4; ++q) {
        const float4 v = s4[q];
        acc = fmaf(sb[4 * q], v.x, acc);
        acc = fmaf(sb[4 * q + 1], v.y, acc);
        acc = fmaf(sb[4 * q + 2], v.z, acc);
        acc = fmaf(sb[4 * q + 3], v.w, acc);
      }
    }
    out[(long)o * N + p] = ACT == NFGS_FNO_GELU ? gelu(acc) : acc;
  }
}

// Backward of pw_forward.  The gradient reaching the layer's output is
// g = gout + E_x(gspec) (gspec, when GSPEC, is (COUT, nz*ny, 16)); then
// gz = g * act'(z) (z recomputed from in, W, b and spec) and gin = W^T gz.
// gz may be null (identity activation without gspec: gz is gout itself).
template <int CIN, int COUT, int ACT, bool SPEC, bool GSPEC>
__global__ __launch_bounds__(256) void pw_backward(
    const float* __restrict__ in, const uint16_t* __restrict__ w, const float* __restrict__ bias,
    const float* __restrict__ spec, const float* __restrict__ bx, int nx, int rows,
    const float* __restrict__ gout, const float* __restrict__ gspec, float* __restrict__ gz,
    float* __restrict__ gin) {
  __shared__ float ws[COUT * CIN];
  __shared__ float bs[COUT];
  for (int k = threadIdx.x; k < COUT * CIN; k += 256) ws[k] = bf16f(w[k]);
  for (int k = threadIdx.x; k < COUT; k += 256) bs[k] = bias[k];
  __syncthreads();
  const long N = (long)rows * nx;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  constexpr bool RECOMPUTE = ACT == NFGS_FNO_GELU;
  float xin[RECOMPUTE ? CIN : 1];
  if constexpr (RECOMPUTE) {
#pragma unroll
    for (int i = 0; i < CIN; ++i) xin[i] = in[(long)i * N + p];
  }
  float sb[kM];
  long srow = 0;
  if constexpr (SPEC || GSPEC) {
    const long row = p / nx;
    const int x = (int)(p - row * nx);
#pragma unroll
    for (int k = 0; k < kM; ++k) sb[k] = bx[x * kM + k];
    srow = row * kM;
  }
  float gacc[CIN];
#pragma unroll
  for (int i = 0; i < CIN; ++i) gacc[i] = 0.0f;
  const bool want_gin = gin != nullptr;
  for (int o = 0; o < COUT; ++o) {
    float g = gout[(long)o * N + p];
    if constexpr (GSPEC) {
      const float4* s4 = reinterpret_cast<const float4*>(gspec + (long)o * rows * kM + srow);
#pragma unroll
      for (int q = 0; q < kM / 4; ++q) {
        const float4 v = s4[q];
        g = fmaf(sb[4 * q], v.x, g);
        g = fmaf(sb[4 * q + 1], v.y, g);
        g = fmaf(sb[4 * q + 2], v.z, g);
        g = fmaf(sb[4 * q + 3], v.w, g);
      }
    }
    if constexpr (RECOMPUTE) {
      float z = bs[o];
#pragma unroll
      for (int i = 0; i < CIN; ++i) z = fmaf(ws[o * CIN + i], xin[i], z);
      if constexpr (SPEC) {
        const float4* s4 = reinterpret_cast<const float4*>(spec + (long)o * rows * kM + srow);
#pragma unroll
        for (int q = 0; q < kM / 4; ++q) {
          const float4 v = s4[q];
          z = fmaf(sb[4 * q], v.x, z);
          z = fmaf(sb[4 * q + 1], v.y, z);
          z = fmaf(sb[4 * q + 2], v.z, z);
          z = fmaf(sb[4 * q + 3], v.w, z);
        }
      }
      g *= gelu_grad(z);
    }
    if (gz) gz[(long)o * N + p] = g;
    if (want_gin) {
#pragma unroll
      for (int i = 0; i < CIN; ++i) gacc[i] = fmaf(ws[o * CIN + i], g, gacc[i]);
    }
  }
  if (want_gin) {
#pragma unroll
    for (int i = 0; i < CIN; ++i) gin[(long)i * N + p] = gacc[i];
  }
}

// ---------------------------------------------------------------------------
// Weight gradients of a pointwise layer: gw[o][i] = sum_p a[o][p] b[i][p],
// gb[o] = sum_p a[o][p].  Each block takes chunks of 64 nodes, staged in LDS
// (the next chunk's loads are in flight in registers while the current one
// is used); a thread owns an RB x 4 block of (o, i) pairs over a share of
// the chunk's nodes, read four nodes at a time, and the shares are summed in
// LDS at the end.  One FP32 partial per block and pair, stored pair-major;
// the partials are summed in FP64 in a fixed order, one wavefront per pair.
// ---------------------------------------------------------------------------
constexpr int kOsBlocks = 2048;
constexpr int kOsQ = 64;

template <int CA, int CB>
struct OuterShape {
  static constexpr int RB = CA < 4 ? CA : 4;
  static constexpr int NPB = (CA / RB) * (CB / 4);  // pair blocks
  static constexpr int G = 256 / NPB;                // node groups
  static constexpr int QG = kOsQ / G;                // nodes per group and chunk
  static constexpr int COUNT = CA * CB + CA;         // partials per block
  static_assert(CA % RB == 0 && CB % 4 == 0 && 256 % NPB == 0 && kOsQ % G == 0 && QG % 4 == 0,
                "shape");
};

template <int CA, int CB>
__global__ __launch_bounds__(256) void outer_sum_kernel(const float* __restrict__ a,
                                                        const float* __restrict__ b, long N,
                                                        float* __restrict__ part) {
  using S = OuterShape<CA, CB>;
  constexpr int RB = S::RB, G = S::G, QG = S::QG;
  constexpr int LD = kOsQ + 4;  // 16-byte rows
  constexpr int NA = (CA * kOsQ + 255) / 256, NBL = (CB * kOsQ + 255) / 256;
  __shared__ __align__(16) float la[CA][LD];
  __shared__ __align__(16) float lb[CB][LD];
  __shared__ float red[256][RB * 4 + 1];
  const int t = threadIdx.x;
  const int pb = t % S::NPB, grp = t / S::NPB;
  const int o0 = (pb / (CB / 4)) * RB, i0 = (pb % (CB / 4)) * 4;
  float acc[RB][4];
  float accb[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    accb[r] = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0f;
  }
  float ra[NA], rb[NBL];
  auto fetch = [&](long base) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int k = t + 256 * u;
      const long p = base + (k % kOsQ);
      ra[u] = (k < CA * kOsQ && p < N) ? a[(long)(k / kOsQ) * N + p] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < NBL; ++u) {
      const int k = t + 256 * u;
      const long p = base + (k % kOsQ);
      rb[u] = (k < CB * kOsQ && p < N) ? b[(long)(k / kOsQ) * N + p] : 0.0f;
    }
  };
  const long stride = (long)gridDim.x * kOsQ;
  long base = (long)blockIdx.x * kOsQ;
  if (base < N) fetch(base);
  for (; base < N; base += stride) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int k = t + 256 * u;
      if (k < CA * kOsQ) la[k / kOsQ][k % kOsQ] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < NBL; ++u) {
      const int k = t + 256 * u;
      if (k < CB * kOsQ) lb[k / kOsQ][k % kOsQ] = rb[u];
    }
    __syncthreads();
    if (base + stride < N) fetch(base + stride);
#pragma unroll
    for (int qq = 0; qq < QG; qq += 4) {
      const int q = grp * QG + qq;
      float4 av[RB], bv[4];
#pragma unroll
      for (int r = 0; r < RB; ++r) av[r] = *reinterpret_cast<const float4*>(&la[o0 + r][q]);
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = *reinterpret_cast<const float4*>(&lb[i0 + c][q]);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        accb[r] += (av[r].x + av[r].y) + (av[r].z + av[r].w);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          float v = acc[r][c];
          v = fmaf(av[r].x, bv[c].x, v);
          v = fmaf(av[r].y, bv[c].y, v);
          v = fmaf(av[r].z, bv[c].z, v);
          v = fmaf(av[r].w, bv[c].w, v);
          acc[r][c] = v;
        }
      }
    }
    __syncthreads();
  }
  // sum the node groups of each pair block
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) red[t][r * 4 + c] = acc[r][c];
  __syncthreads();
  const long nbk = gridDim.x, blk = blockIdx.x;
  for (int k = t; k < S::NPB * RB * 4; k += 256) {
    const int pbk = k / (RB * 4), e = k % (RB * 4);
    float sum = 0.0f;
    for (int g = 0; g < G; ++g) sum += red[g * S::NPB + pbk][e];
    const int o = (pbk / (CB / 4)) * RB + e / 4, i = (pbk % (CB / 4)) * 4 + e % 4;
    part[(long)(o * CB + i) * nbk + blk] = sum;
  }
  __syncthreads();
  // bias sums: the pair blocks of column 0 hold every o
#pragma unroll
  for (int r = 0; r < RB; ++r) red[t][r] = accb[r];
  __syncthreads();
  if (t < S::NPB && (t % (CB / 4)) == 0) {
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      float sum = 0.0f;
      for (int g = 0; g < G; ++g) sum += red[g * S::NPB + t][r];
      part[(long)(CA * CB + (t / (CB / 4)) * RB + r) * nbk + blk] = sum;
    }
  }
}

// One wavefront per output: lanes take every 64th block partial, then a
// fixed butterfly over the wavefront.
__global__ __launch_bounds__(256) void sum_partials_kernel(const float* __restrict__ part,
                                                           int nblocks, int count, int nw,
                                                           float* __restrict__ gw,
                                                           float* __restrict__ gb, int accumulate) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= count) return;
  const float* src = part + (long)j * nblocks;
  double sum = 0.0;
  for (int k = lane; k < nblocks; k += 64) sum += (double)src[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if (lane == 0) {
    float* dst = j < nw ? gw + j : gb + (j - nw);
    *dst = accumulate ? *dst + (float)sum : (float)sum;
  }
}

// ---------------------------------------------------------------------------
// Truncated DCT-II along one axis of an (outer, n, inner) array.
//   project: out[o][k][i] = sum_j B[j][k] in[o][j][i]   (k < 16)
//   expand:  out[o][j][i] = sum_k B[j][k] in[o][k][i]   (j < n)
// B is the n x 16 orthonormal basis.  inner > 1: one thread per (o, i), the
// basis read as uniform (scalar) loads.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void axis_project_inner(const float* __restrict__ in,
                                                          float* __restrict__ out,
                                                          const float* __restrict__ B, long outer,
                                                          int n, long inner) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= outer * inner) return;
  const long o = t / inner, i = t - o * inner;
  const float* src = in + o * n * inner + i;
  float acc[kM];
#pragma unroll
  for (int k = 0; k < kM; ++k) acc[k] = 0.0f;
  for (int j = 0; j < n; ++j) {
    const float x = src[(long)j * inner];
#pragma unroll
    for (int k = 0; k < kM; ++k) acc[k] = fmaf(B[j * kM + k], x, acc[k]);
  }
  float* dst = out + o * kM * inner + i;
#pragma unroll
  for (int k = 0; k < kM; ++k) dst[(long)k * inner] = acc[k];
}

__global__ __launch_bounds__(256) void axis_expand_inner(const float* __restrict__ in,
                                                         float* __restrict__ out,
                                                         const float* __restrict__ B, long outer,
                                                         int n, long inner) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= outer * inner) return;
  const long o = t / inner, i = t - o * inner;
  const float* src = in + o * kM * inner + i;
  float x[kM];
#pragma unroll
  for (int k = 0; k < kM; ++k) x[k] = src[(long)k * inner];
  float* dst = out + o * n * inner + i;
  for (int j = 0; j < n; ++j) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < kM; ++k) acc = fmaf(B[j * kM + k], x[k], acc);
    dst[(long)j * inner] = acc;
  }
}

// x axis: R rows per block copied linearly into LDS; each wavefront forms a
// 16-row x 16-mode tile on v_mfma_f32_16x16x4f32 (A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k, output row 4(l/16) + r of lane l in register r).
typedef float f32x4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4v mma(float a, float b, f32x4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int R>
__global__ __launch_bounds__(R * 4) void axis_project_rows(const float* __restrict__ in,
                                                           float* __restrict__ out,
                                                           const float* __restrict__ B, long rows,
                                                           int n) {
  extern __shared__ __align__(16) float lrow[];  // R rows of n
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long r0 = (long)blockIdx.x * R;
  const int nr = (int)((rows - r0) < R ? (rows - r0) : R);
  const long cnt = (long)nr * n;
  const float* src = in + r0 * n;
  // r0 * n * 4 bytes is a multiple of 16 (R is a multiple of 16)
  const long c4 = cnt >> 2;
  for (long k = t; k < c4; k += R * 4)
    reinterpret_cast<float4*>(lrow)[k] = reinterpret_cast<const float4*>(src)[k];
  for (long k = (c4 << 2) + t; k < cnt; k += R * 4) lrow[k] = src[k];
  __syncthreads();
  const int i = lane & 15, kk = lane >> 4;
  const int row = w * 16 + i;
  const bool rowok = row < nr;
  const float* lr = lrow + row * n;
  f32x4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int j0 = 0; j0 < n; j0 += 4) {
    const int j = j0 + kk;
    const float a = (rowok && j < n) ? lr[j] : 0.0f;
    const float b = j < n ? B[j * kM + i] : 0.0f;
    acc = mma(a, b, acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int orow = w * 16 + 4 * kk + r;
    if (orow < nr) out[(r0 + orow) * kM + i] = acc[r];
  }
}

__global__ __launch_bounds__(256) void axis_expand_rows(const float* __restrict__ in,
                                                        float* __restrict__ out,
                                                        const float* __restrict__ B, long rows,
                                                        int n) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= rows * n) return;
  const long r = t / n;
  const int j = (int)(t - r * n);
  const float* src = in + r * kM;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < kM; ++k) acc = fmaf(B[j * kM + k], src[k], acc);
  out[t] = acc;
}

// ---------------------------------------------------------------------------
// Mode mixing: out[o][m] = sum_i R[m][o][i] in[i][m]; its transpose
// out[i][m] = sum_o R[m][o][i] in[o][m]; and the weight gradient
// gr[m][o][i] += g[o][m] u[i][m].  Arrays of kW channels over nm modes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mix_kernel(const uint16_t* __restrict__ R,
                                                  const float* __restrict__ in,
                                                  float* __restrict__ out, int nm, int transpose) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nm * kW) return;
  const int m = (int)(t / kW), c = (int)(t % kW);
  const uint16_t* r = R + (long)m * kW * kW;
  float acc = 0.0f;
  if (!transpose) {
#pragma unroll 8
    for (int i = 0; i < kW; ++i) acc = fmaf(bf16f(r[c * kW + i]), in[(long)i * nm + m], acc);
  } else {
#pragma unroll 8
    for (int o = 0; o < kW; ++o) acc = fmaf(bf16f(r[o * kW + c]), in[(long)o * nm + m], acc);
  }
  out[(long)c * nm + m] = acc;
}

__global__ __launch_bounds__(256) void mix_grad_kernel(const float* __restrict__ g,
                                                       const float* __restrict__ u,
                                                       float* __restrict__ gr, int nm) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nm * kW * kW) return;
  const int m = (int)(t / (kW * kW));
  const int o = (int)((t / kW) % kW), i = (int)(t % kW);
  gr[t] = fmaf(g[(long)o * nm + m], u[(long)i * nm + m], gr[t]);
}

// ---------------------------------------------------------------------------
// Loss: the mean of (out - target)^2 over every node of the output channels;
// gout = d loss / d out.
// ---------------------------------------------------------------------------
constexpr int kLossBlocks = 256;

__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ out,
                                                   const float* __restrict__ tgt, long n,
                                                   float scale, float* __restrict__ gout,
                                                   double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
    const float d = out[p] - tgt[p];
    gout[p] = 2.0f * scale * d;
    acc += (double)d * d;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void loss_sum_kernel(const double* __restrict__ part, int nb, double scale,
                                double* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b];
  *loss = s * scale;
}

// ---------------------------------------------------------------------------
// Adam on FP32 master weights; writes the bfloat16 copy the kernels use and
// clears the gradient for the next accumulation.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ master,
                                                   uint16_t* __restrict__ wb,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   float* __restrict__ g, long n, float lr,
                                                   float b1, float b2, float eps, float c1,
                                                   float c2, float gscale) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * gscale;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float w = master[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  master[i] = w;
  wb[i] = to_bf16(w);
  g[i] = 0.0f;
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ src,
                                                      uint16_t* __restrict__ dst, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = to_bf16(src[i]);
}

// ---------------------------------------------------------------------------
// Inputs, targets and the start state.
// ---------------------------------------------------------------------------
template <typename T>
int features_t(const T* b, const T* jp, int nx, int ny, int nz, float cb, float cj, float rho,
               float* feat, hipStream_t s) {
  const long pl = (long)nx * ny, N = pl * nz;
  const float zs = nz > 1 ? 1.0f / (float)(nz - 1) : 0.0f;
  return launch_map(N, [=] __device__(long p) {
    const int z = (int)(p / pl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      feat[c * N + p] = (float)b[c * N + p] * cb;
      feat[(3 + c) * N + p] = (float)jp[c * N + p] * cj;
    }
    feat[6 * N + p] = rho;
    feat[7 * N + p] = (float)z * zs;
  }, s);
}

template <typename T>
int targets_t(const T* jp_gs, const T* jp, long n, double inv_scale, float* tgt, hipStream_t s) {
  return launch_map(n, [=] __device__(long p) {
    tgt[p] = (float)(((double)jp[p] - (double)jp_gs[p]) * inv_scale);
  }, s);
}

template <typename T>
int apply_t(const T* jp_gs, const float* net, long n, double scale, T* jp, hipStream_t s) {
  return launch_map(n, [=] __device__(long p) {
    jp[p] = (T)((double)jp_gs[p] + scale * (double)net[p]);
  }, s);
}

bool grid_ok(int nx, int ny, int nz) {
  return nx >= kM && ny >= kM && nz >= kM && (long)nx * ny * nz * 64 < (1L << 40);
}

}  // namespace
}  // namespace nfgs

using namespace nfgs;

#define FNO_STREAM(s) (static_cast<hipStream_t>(s))
#define FNO_GRID_CHECK(nx, ny, nz)                                                     \
  NFGS_REQUIRE(grid_ok(nx, ny, nz), NFGS_ERR_ARG,                                       \
               "learned start: every grid axis needs at least 16 nodes")

extern "C" {

int nfgs_fno_pointwise(const float* in, int cin, const uint16_t* weight, const float* bias,
                       const float* spec, const float* basis_x, int cout, int activation,
                       int nx, int ny, int nz, float* out, void* stream) {
  NFGS_REQUIRE(in && weight && bias && out, NFGS_ERR_ARG, "fno pointwise: null pointer");
  FNO_GRID_CHECK(nx, ny, nz);
  const int rows = ny * nz;
  const long N = (long)rows * nx;
  const dim3 grid((unsigned)((N + 255) / 256)), blk(256);
  hipStream_t s = FNO_STREAM(stream);
  const bool sp = spec != nullptr;
  NFGS_REQUIRE(!sp || basis_x, NFGS_ERR_ARG, "fno pointwise: spectral input without basis");
#define PWF(CI, CO, A, S)                                                                       \
  hipLaunchKernelGGL((pw_forward<CI, CO, A, S>), grid, blk, 0, s, in, weight, bias, spec, basis_x, \
                     nx, rows, out)
  if (cin == 8 && cout == kW && activation == NFGS_FNO_IDENTITY && !sp) PWF(8, kW, 0, false);
  else if (cin == kW && cout == kW && activation == NFGS_FNO_GELU && sp) PWF(kW, kW, 1, true);
  else if (cin == kW && cout == kW && activation == NFGS_FNO_IDENTITY && sp) PWF(kW, kW, 0, true);
  else if (cin == kW && cout == 64 && activation == NFGS_FNO_GELU && !sp) PWF(kW, 64, 1, false);
  else if (cin == 64 && cout == 3 && activation == NFGS_FNO_IDENTITY && !sp) PWF(64, 3, 0, false);
  else NFGS_REQUIRE(false, NFGS_ERR_ARG, "fno pointwise: unsupported layer shape");
#undef PWF
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_pointwise_backward(const float* in, int cin, const uint16_t* weight,
                                const float* bias, const float* spec, const float* basis_x,
                                int cout, int activation, const float* grad_out,
                                const float* grad_out_spec, int nx, int ny, int nz,
                                float* grad_z, float* grad_in, void* stream) {
  NFGS_REQUIRE(weight && bias && grad_out, NFGS_ERR_ARG, "fno pointwise backward: null pointer");
  FNO_GRID_CHECK(nx, ny, nz);
  const bool gelu_act = activation == NFGS_FNO_GELU;
  NFGS_REQUIRE(!gelu_act || in, NFGS_ERR_ARG, "fno pointwise backward: GELU needs the input");
  NFGS_REQUIRE(grad_z || (!gelu_act && !grad_out_spec), NFGS_ERR_ARG,
               "fno pointwise backward: grad_z may be omitted only when it equals grad_out");
  NFGS_REQUIRE(!(spec || grad_out_spec) || basis_x, NFGS_ERR_ARG,
               "fno pointwise backward: spectral input without basis");
  const int rows = ny * nz;
  const long N = (long)rows * nx;
  const dim3 grid((unsigned)((N + 255) / 256)), blk(256);
  hipStream_t s = FNO_STREAM(stream);
  const bool sp = spec != nullptr, gs = grad_out_spec != nullptr;
#define PWB(CI, CO, A, S, GS)                                                                  \
  hipLaunchKernelGGL((pw_backward<CI, CO, A, S, GS>), grid, blk, 0, s, in, weight, bias, spec,   \
                     basis_x, nx, rows, grad_out, grad_out_spec, grad_z, grad_in)
  if (cin == 64 && cout == 3 && !gelu_act && !sp && !gs) PWB(64, 3, 0, false, false);
  else if (cin == kW && cout == 64 && gelu_act && !sp && !gs) PWB(kW, 64, 1, false, false);
  else if (cin == kW && cout == kW && !gelu_act && !gs) PWB(kW, kW, 0, false, false);
  else if (cin == kW && cout == kW && gelu_act && sp && gs) PWB(kW, kW, 1, true, true);
  else if (cin == 8 && cout == kW && !gelu_act && !sp && gs && !grad_in) PWB(8, kW, 0, false, true);
  else NFGS_REQUIRE(false, NFGS_ERR_ARG, "fno pointwise backward: unsupported layer shape");
#undef PWB
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

long nfgs_fno_outer_sum_workspace_bytes(int ca, int cb) {
  return (long)kOsBlocks * (ca * cb + ca) * (long)sizeof(float);
}

int nfgs_fno_outer_sum(const float* a, int ca, const float* b, int cb, long npoints,
                       float* grad_w, float* grad_b, int accumulate, void* workspace,
                       void* stream) {
  NFGS_REQUIRE(a && b && grad_w && grad_b && workspace && npoints > 0, NFGS_ERR_ARG,
               "fno outer sum: null pointer or empty input");
  hipStream_t s = FNO_STREAM(stream);
  float* part = static_cast<float*>(workspace);
  const long nb = (npoints + kOsQ - 1) / kOsQ;
  const unsigned blocks = (unsigned)(nb < kOsBlocks ? nb : kOsBlocks);
#define OS(A, B)                                                                               \
  hipLaunchKernelGGL((outer_sum_kernel<A, B>), dim3(blocks), dim3(256), 0, s, a, b, npoints, part)
  if (ca == kW && cb == 8) OS(kW, 8);
  else if (ca == kW && cb == kW) OS(kW, kW);
  else if (ca == 64 && cb == kW) OS(64, kW);
  else if (ca == 3 && cb == 64) OS(3, 64);
  else NFGS_REQUIRE(false, NFGS_ERR_ARG, "fno outer sum: unsupported shape");
#undef OS
  NFGS_LAUNCH_CHECK();
  const int count = ca * cb + ca;
  hipLaunchKernelGGL(sum_partials_kernel, dim3((count + 3) / 4), dim3(256), 0, s, part,
                     (int)blocks, count, ca * cb, grad_w, grad_b, accumulate);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_axis(const float* in, float* out, const float* basis, long outer, int n, long inner,
                  int direction, void* stream) {
  NFGS_REQUIRE(in && out && basis, NFGS_ERR_ARG, "fno axis: null pointer");
  NFGS_REQUIRE(outer > 0 && inner > 0 && n >= kM && n <= 4096, NFGS_ERR_ARG,
               "fno axis: the axis needs 16 to 4096 nodes");
  NFGS_REQUIRE(direction == NFGS_FNO_PROJECT || direction == NFGS_FNO_EXPAND, NFGS_ERR_ARG,
               "fno axis: direction must be NFGS_FNO_PROJECT or NFGS_FNO_EXPAND");
  hipStream_t s = FNO_STREAM(stream);
  if (inner == 1 && direction == NFGS_FNO_PROJECT) {
    // rows per block so that R n floats fit in 64 KiB of LDS
    NFGS_REQUIRE(n <= 1024, NFGS_ERR_ARG, "fno axis: rows longer than 1024 nodes");
    const int R = n <= 256 ? 64 : (n <= 512 ? 32 : 16);
    const size_t lds = (size_t)R * n * sizeof(float);
    const dim3 g((unsigned)((outer + R - 1) / R));
    if (R == 64)
      hipLaunchKernelGGL(axis_project_rows<64>, g, dim3(256), lds, s, in, out, basis, outer, n);
    else if (R == 32)
      hipLaunchKernelGGL(axis_project_rows<32>, g, dim3(128), lds, s, in, out, basis, outer, n);
    else
      hipLaunchKernelGGL(axis_project_rows<16>, g, dim3(64), lds, s, in, out, basis, outer, n);
  } else if (inner == 1) {
    hipLaunchKernelGGL(axis_expand_rows, dim3((unsigned)((outer * n + 255) / 256)), dim3(256), 0,
                       s, in, out, basis, outer, n);
  } else {
    const unsigned g = (unsigned)((outer * inner + 255) / 256);
    if (direction == NFGS_FNO_PROJECT)
      hipLaunchKernelGGL(axis_project_inner, dim3(g), dim3(256), 0, s, in, out, basis, outer, n,
                         inner);
    else
      hipLaunchKernelGGL(axis_expand_inner, dim3(g), dim3(256), 0, s, in, out, basis, outer, n,
                         inner);
  }
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_mix(const uint16_t* weight, const float* in, float* out, int nmodes, int transpose,
                 void* stream) {
  NFGS_REQUIRE(weight && in && out && nmodes > 0, NFGS_ERR_ARG, "fno mix: null pointer");
  const long n = (long)nmodes * kW;
  hipLaunchKernelGGL(mix_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     FNO_STREAM(stream), weight, in, out, nmodes, transpose);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_mix_grad(const float* grad_out, const float* in, float* grad_weight, int nmodes,
                      void* stream) {
  NFGS_REQUIRE(grad_out && in && grad_weight && nmodes > 0, NFGS_ERR_ARG,
               "fno mix grad: null pointer");
  const long n = (long)nmodes * kW * kW;
  hipLaunchKernelGGL(mix_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     FNO_STREAM(stream), grad_out, in, grad_weight, nmodes);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

long nfgs_fno_loss_workspace_bytes(void) { return (long)kLossBlocks * (long)sizeof(double); }

int nfgs_fno_loss(const float* out, const float* target, long n, float* grad_out, double* loss,
                  void* workspace, void* stream) {
  NFGS_REQUIRE(out && target && grad_out && loss && workspace && n > 0, NFGS_ERR_ARG,
               "fno loss: null pointer or empty output");
  const double scale = 1.0 / (double)n;
  hipStream_t s = FNO_STREAM(stream);
  double* part = static_cast<double*>(workspace);
  hipLaunchKernelGGL(loss_kernel, dim3(kLossBlocks), dim3(256), 0, s, out, target, n,
                     (float)scale, grad_out, part);
  NFGS_LAUNCH_CHECK();
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, s, part, kLossBlocks, scale, loss);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_adam(float* master, uint16_t* weight, float* m, float* v, float* grad, long n,
                  double lr, double beta1, double beta2, double eps, int step, double grad_scale,
                  void* stream) {
  NFGS_REQUIRE(master && weight && m && v && grad && n > 0 && step >= 1, NFGS_ERR_ARG,
               "fno adam: null pointer, empty tensor or step < 1");
  const double c1 = 1.0 - pow(beta1, step), c2 = 1.0 - pow(beta2, step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     FNO_STREAM(stream), master, weight, m, v, grad, n, (float)lr, (float)beta1,
                     (float)beta2, (float)eps, (float)c1, (float)c2, (float)grad_scale);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_to_bf16(const float* src, uint16_t* dst, long n, void* stream) {
  NFGS_REQUIRE(src && dst && n > 0, NFGS_ERR_ARG, "fno to_bf16: null pointer");
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     FNO_STREAM(stream), src, dst, n);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

int nfgs_fno_features(int dtype, const void* b, const void* jp, int nx, int ny, int nz, double c_b,
                      double c_j, double rho, float* feat, void* stream) {
  NFGS_REQUIRE(b && jp && feat, NFGS_ERR_ARG, "fno features: null pointer");
  FNO_GRID_CHECK(nx, ny, nz);
  hipStream_t s = FNO_STREAM(stream);
  if (dtype == NFGS_F32)
    return features_t<float>((const float*)b, (const float*)jp, nx, ny, nz, (float)c_b,
                             (float)c_j, (float)rho, feat, s);
  NFGS_REQUIRE(dtype == NFGS_F64, NFGS_ERR_ARG, "fno features: dtype must be F32 or F64");
  return features_t<double>((const double*)b, (const double*)jp, nx, ny, nz, (float)c_b,
                            (float)c_j, (float)rho, feat, s);
}

int nfgs_fno_targets(int dtype, const void* jp_gs, const void* jp, long n, double inv_scale,
                     float* target, void* stream) {
  NFGS_REQUIRE(jp_gs && jp && target && n > 0, NFGS_ERR_ARG, "fno targets: null pointer");
  hipStream_t s = FNO_STREAM(stream);
  if (dtype == NFGS_F32)
    return targets_t<float>((const float*)jp_gs, (const float*)jp, n, inv_scale, target, s);
  NFGS_REQUIRE(dtype == NFGS_F64, NFGS_ERR_ARG, "fno targets: dtype must be F32 or F64");
  return targets_t<double>((const double*)jp_gs, (const double*)jp, n, inv_scale, target, s);
}

int nfgs_fno_apply(int dtype, const void* jp_gs, const float* net, long n, double scale, void* jp,
                   void* stream) {
  NFGS_REQUIRE(jp_gs && net && jp && n > 0, NFGS_ERR_ARG, "fno apply: null pointer");
  hipStream_t s = FNO_STREAM(stream);
  if (dtype == NFGS_F32)
    return apply_t<float>((const float*)jp_gs, net, n, scale, (float*)jp, s);
  NFGS_REQUIRE(dtype == NFGS_F64, NFGS_ERR_ARG, "fno apply: dtype must be F32 or F64");
  return apply_t<double>((const double*)jp_gs, net, n, scale, (double*)jp, s);
}

}  // extern "C"
