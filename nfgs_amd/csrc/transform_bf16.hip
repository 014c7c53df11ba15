// Lateral type-I transforms as split-bf16 x3 MFMA products with FP32
// accumulation: the precision form of the lateral GEMMs.
//
// Every FP32 input x and every matrix entry t is split into two bf16 terms,
// x = xh + xl and t = th + tl (16 significant bits), and x t is taken as
// xh th + xh tl + xl th, each product exact in FP32 and summed in FP32 by
// v_mfma_f32_16x16x16_bf16 (the dropped xl tl term and the representation
// leave a relative error near 2^-16 per product).  The plan holds th, tl in
// the MFMA operand layout (PairedMatrix::bfr).
//
// A workgroup of 512 lanes (8 wavefronts) takes 128 vectors; each wavefront
// owns 16 of them.  The contraction runs over chunks of 16 paired rows (32
// elements): the chunk's data is staged in LDS as FP32 and split per lane,
// the matrix fragments of up to kBNB output blocks are staged in LDS once
// for the 8 wavefronts.  ELEM_CONTIG (elements adjacent, the x axis)
// computes D = X T^T (lanes hold consecutive outputs); otherwise D = T X
// (lanes hold consecutive vectors).  The data and matrix fragments of a lane
// have the same shape in both: 4 paired rows 4*(lane/16)..+3 of vector or
// output lane%16.

#include "plan.h"

namespace nfgs {
namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBVec = 128;  // vectors per workgroup
constexpr int kBNB = 9;     // 16-row output blocks per group
constexpr int kBXS = 20;    // LDS stride (floats) of one vector's 16 paired elements

// bf16 of an FP32 value, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ void split4(const float* p, s16x4& hi, s16x4& lo) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const unsigned h = bf16_rne(p[t]);
    hi[t] = (short)h;
    lo[t] = (short)bf16_rne(p[t] - __uint_as_float(h << 16));
  }
}

__device__ __forceinline__ f32x4 mma(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ s16x4 as_s16x4(uint2 v) {
  s16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

template <bool ELEM_CONTIG>
__global__ __launch_bounds__(512) void paired_transform_bf16x3(
    const float* __restrict__ in, long ies, long ivs, long ips, float* __restrict__ out,
    long oes, long ovs, long ops, const uint2* __restrict__ frag, int n, int h, int nchunks,
    int nblocks, int nvec, long nvec_total) {
  __shared__ float xe[kBVec * kBXS];
  __shared__ float xo[kBVec * kBXS];
  __shared__ uint2 mf[4][kBNB][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long g0 = (long)blockIdx.x * kBVec;

  // staging slots: 128 vectors x 32 elements, 8 per lane, read along the
  // contiguous axis
  long ibase[8];
  int sv[8], sj[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int e = tid + 512 * i;
    sv[i] = ELEM_CONTIG ? (e >> 5) : (e & (kBVec - 1));
    sj[i] = ELEM_CONTIG ? (e & 31) : (e >> 7);
    const long g = g0 + sv[i];
    if (g < nvec_total) {
      const long p = g / nvec, q = g - p * nvec;
      ibase[i] = p * ips + q * ivs;
    } else {
      ibase[i] = -1;
    }
  }

  const int ngroups = (nblocks + kBNB - 1) / kBNB;
  for (int grp = 0; grp < ngroups; ++grp) {
    const int b0 = grp * kBNB;
    const int nbg = min(kBNB, nblocks - b0);
    f32x4 accE[kBNB], accO[kBNB];
#pragma unroll
    for (int b = 0; b < kBNB; ++b) {
      accE[b] = f32x4{0, 0, 0, 0};
      accO[b] = f32x4{0, 0, 0, 0};
    }
    for (int c = 0; c < nchunks; ++c) {
      lds_barrier();  // the previous chunk is consumed
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int j = c * 32 + sj[i];
        const float v = (ibase[i] >= 0 && j < n) ? in[ibase[i] + (long)j * ies] : 0.0f;
        const int idx = sv[i] * kBXS + (sj[i] >> 1);
        if (sj[i] & 1)
          xo[idx] = v;
        else
          xe[idx] = v;
      }
      for (int e = tid; e < 4 * kBNB * 64; e += 512) {
        const int a = e / (kBNB * 64), r = e - a * (kBNB * 64), b = r >> 6, l = r & 63;
        mf[a][b][l] = b < nbg ? frag[(((long)a * nchunks + c) * nblocks + b0 + b) * 64 + l]
                              : make_uint2(0u, 0u);
      }
      lds_barrier();
      const int fo = (wave * 16 + (lane & 15)) * kBXS + 4 * (lane >> 4);
      s16x4 eh, el, oh, ol;
      split4(xe + fo, eh, el);
      split4(xo + fo, oh, ol);
#pragma unroll
      for (int b = 0; b < kBNB; ++b) {
        if (b < nbg) {
          const s16x4 Eh = as_s16x4(mf[0][b][lane]), El = as_s16x4(mf[1][b][lane]);
          const s16x4 Oh = as_s16x4(mf[2][b][lane]), Ol = as_s16x4(mf[3][b][lane]);
          if (ELEM_CONTIG) {
            accE[b] = mma(el, Eh, accE[b]);
            accE[b] = mma(eh, El, accE[b]);
            accE[b] = mma(eh, Eh, accE[b]);
            accO[b] = mma(ol, Oh, accO[b]);
            accO[b] = mma(oh, Ol, accO[b]);
            accO[b] = mma(oh, Oh, accO[b]);
          } else {
            accE[b] = mma(Eh, el, accE[b]);
            accE[b] = mma(El, eh, accE[b]);
            accE[b] = mma(Eh, eh, accE[b]);
            accO[b] = mma(Oh, ol, accO[b]);
            accO[b] = mma(Ol, oh, accO[b]);
            accO[b] = mma(Oh, oh, accO[b]);
          }
        }
      }
    }
    // out[k] = E + O, out[n-1-k] = E - O
#pragma unroll
    for (int b = 0; b < kBNB; ++b) {
      if (b >= nbg) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int k, v;
        if (ELEM_CONTIG) {
          v = wave * 16 + 4 * (lane >> 4) + r;
          k = (b0 + b) * 16 + (lane & 15);
        } else {
          k = (b0 + b) * 16 + 4 * (lane >> 4) + r;
          v = wave * 16 + (lane & 15);
        }
        const long g = g0 + v;
        if (k < h && g < nvec_total) {
          const long p = g / nvec, q = g - p * nvec;
          float* o = out + p * ops + q * ovs;
          const float e = accE[b][r], od = accO[b][r];
          o[(long)k * oes] = e + od;
          const int k2 = n - 1 - k;
          if (k2 != k) o[(long)k2 * oes] = e - od;
        }
      }
    }
  }
}

}  // namespace

int launch_bf16x3(const PairedMatrix* m, int dir, const float* in, long ies, long ivs, long ips,
                  float* out, long oes, long ovs, long ops, int nvec, long total, hipStream_t s) {
  NFGS_REQUIRE(m->bfr[dir] != nullptr, NFGS_ERR_STATE,
               "transform: the plan holds no split-bf16 matrices (FP32 plans only)");
  const long nblk = (total + kBVec - 1) / kBVec;
  NFGS_REQUIRE(nblk < (1L << 31), NFGS_ERR_ARG, "transform: too many vectors");
  const uint2* f = static_cast<const uint2*>(m->bfr[dir]);
  if (ies == 1 && oes == 1) {
    hipLaunchKernelGGL((paired_transform_bf16x3<true>), dim3((unsigned)nblk), dim3(512), 0, s, in,
                       ies, ivs, ips, out, oes, ovs, ops, f, m->n, m->h, m->bchunks, m->bblocks,
                       nvec, total);
  } else if (ivs == 1 && ovs == 1) {
    hipLaunchKernelGGL((paired_transform_bf16x3<false>), dim3((unsigned)nblk), dim3(512), 0, s, in,
                       ies, ivs, ips, out, oes, ovs, ops, f, m->n, m->h, m->bchunks, m->bblocks,
                       nvec, total);
  } else {
    set_error("transform: need unit element strides or unit vector strides on both sides");
    return NFGS_ERR_ARG;
  }
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace nfgs
