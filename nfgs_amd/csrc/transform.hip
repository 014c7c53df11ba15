// Lateral type-I transforms (DCT-I / DST-I and inverses) as paired MFMA GEMMs.
//
// A workgroup of 256 lanes (4 wavefronts) takes 64 vectors of length n.  The
// transform matrix is split into its even-column half E and odd-column half O
// (see PairedMatrix), so each vector costs two (h x n/2) products instead of
// one (n x n) product.  The contraction dimension is streamed through LDS in
// chunks of 16 even + 16 odd elements; each wavefront owns 16 vectors and all
// output rows of the current group, accumulating with v_mfma 16x16x4 (f32 or
// f64) and finishing with out[k] = E+O, out[n-1-k] = E-O.
//
// Two layouts share the kernel: ELEM_CONTIG (elements of a vector are
// adjacent, the x axis) computes C = X * T^T so that lanes hold consecutive
// output elements; otherwise (vectors adjacent, the y axis) it computes
// C = T * X so that lanes hold consecutive vectors.  Either way stores are
// 64-byte runs.

#include "plan.h"

namespace nfgs {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Both take A[i][k] from lane i + 16k and B[k][j] from lane j + 16k; the
// output row held in accumulator register r of lane l differs:
// f32 gives D[4*(l/16) + r][l%16], f64 gives D[l/16 + 4r][l%16].
template <>
struct Mfma<float> {
  using acc = f32x4;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l4, int r) { return 4 * l4 + r; }
};
template <>
struct Mfma<double> {
  using acc = f64x4;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l4, int r) { return l4 + 4 * r; }
};

constexpr int kVec = 64;    // vectors per workgroup
constexpr int kChunk = 16;  // paired rows (j') per LDS chunk
constexpr int kNB = 9;      // 16-row output blocks per group (144 outputs)

template <typename T, bool ELEM_CONTIG>
__global__ __launch_bounds__(256) void paired_transform_kernel(
    const T* __restrict__ in, long ies, long ivs, long ips, T* __restrict__ out, long oes,
    long ovs, long ops, const T* __restrict__ et, const T* __restrict__ ot, int n, int h,
    int rows, int cols_pad, int nvec, long nvec_total) {
  using A = typename Mfma<T>::acc;
  // Data chunk: ELEM_CONTIG -> [vector][j'] (row padded to 17);
  //             otherwise   -> [j'][vector].
  constexpr int XS = ELEM_CONTIG ? kVec * (kChunk + 1) : kChunk * kVec;
  __shared__ T xe[XS];
  __shared__ T xo[XS];
  __shared__ T me[kChunk * kNB * 16];
  __shared__ T mo[kChunk * kNB * 16];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int l16 = lane & 15;
  const int l4 = lane >> 4;
  const long g0 = (long)blockIdx.x * kVec;

  // Base offsets of the 8 (vector, j) slots this thread loads per chunk.
  long ibase[8];
  int islot_v[8], islot_j[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int e = tid + 256 * i;
    const int v = ELEM_CONTIG ? (e >> 5) : (e & 63);
    const int jj = ELEM_CONTIG ? (e & 31) : (e >> 6);
    islot_v[i] = v;
    islot_j[i] = jj;
    const long g = g0 + v;
    if (g < nvec_total) {
      const long p = g / nvec, q = g - p * nvec;
      ibase[i] = p * ips + q * ivs;
    } else {
      ibase[i] = -1;
    }
  }

  const int hblocks = (h + 15) >> 4;
  const int ngroups = (hblocks + kNB - 1) / kNB;
  const int nchunks = (rows + kChunk - 1) / kChunk;

  for (int grp = 0; grp < ngroups; ++grp) {
    const int b0 = grp * kNB;
    const int nbg = min(kNB, hblocks - b0);
    A accE[kNB], accO[kNB];
#pragma unroll
    for (int b = 0; b < kNB; ++b) {
      accE[b] = A{0, 0, 0, 0};
      accO[b] = A{0, 0, 0, 0};
    }

    for (int c = 0; c < nchunks; ++c) {
      __syncthreads();  // previous chunk fully consumed
      // ---- stage the data chunk: original j in [32c, 32c+32) ----
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int j = c * 2 * kChunk + islot_j[i];
        T val = T(0);
        if (ibase[i] >= 0 && j < n) val = in[ibase[i] + (long)j * ies];
        const int jp = islot_j[i] >> 1;
        const int v = islot_v[i];
        const int idx = ELEM_CONTIG ? v * (kChunk + 1) + jp : jp * kVec + v;
        if (islot_j[i] & 1)
          xo[idx] = val;
        else
          xe[idx] = val;
      }
      // ---- stage the matrix chunk: rows [16c, 16c+16), cols of this group ----
      const int mcols = kNB * 16;
      for (int e = tid; e < kChunk * mcols; e += 256) {
        const int r = e / mcols, col = e - r * mcols;
        const int gr = c * kChunk + r, gc = b0 * 16 + col;
        T ve = T(0), vo = T(0);
        if (gr < rows && gc < cols_pad) {
          ve = et[(long)gr * cols_pad + gc];
          vo = ot[(long)gr * cols_pad + gc];
        }
        me[e] = ve;
        mo[e] = vo;
      }
      __syncthreads();

      const int valid = min(kChunk, rows - c * kChunk);
      const int ksteps = (valid + 3) >> 2;
      for (int ks = 0; ks < ksteps; ++ks) {
        const int jl = ks * 4 + l4;
        if (ELEM_CONTIG) {
          const T aE = xe[(wave * 16 + l16) * (kChunk + 1) + jl];
          const T aO = xo[(wave * 16 + l16) * (kChunk + 1) + jl];
#pragma unroll
          for (int b = 0; b < kNB; ++b) {
            if (b < nbg) {
              const T bE = me[jl * mcols + b * 16 + l16];
              const T bO = mo[jl * mcols + b * 16 + l16];
              accE[b] = Mfma<T>::mma(aE, bE, accE[b]);
              accO[b] = Mfma<T>::mma(aO, bO, accO[b]);
            }
          }
        } else {
          const T bE = xe[jl * kVec + wave * 16 + l16];
          const T bO = xo[jl * kVec + wave * 16 + l16];
#pragma unroll
          for (int b = 0; b < kNB; ++b) {
            if (b < nbg) {
              const T aE = me[jl * mcols + b * 16 + l16];
              const T aO = mo[jl * mcols + b * 16 + l16];
              accE[b] = Mfma<T>::mma(aE, bE, accE[b]);
              accO[b] = Mfma<T>::mma(aO, bO, accO[b]);
            }
          }
        }
      }
    }

    // ---- epilogue: out[k] = E + O, out[n-1-k] = E - O ----
#pragma unroll
    for (int b = 0; b < kNB; ++b) {
      if (b >= nbg) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int k, v;
        if (ELEM_CONTIG) {
          v = wave * 16 + Mfma<T>::row(l4, r);
          k = (b0 + b) * 16 + l16;
        } else {
          k = (b0 + b) * 16 + Mfma<T>::row(l4, r);
          v = wave * 16 + l16;
        }
        const long g = g0 + v;
        if (k < h && g < nvec_total) {
          const long p = g / nvec, q = g - p * nvec;
          T* o = out + p * ops + q * ovs;
          const T e = accE[b][r], od = accO[b][r];
          o[(long)k * oes] = e + od;
          const int k2 = n - 1 - k;
          if (k2 != k) o[(long)k2 * oes] = e - od;
        }
      }
    }
  }
}

template <typename T>
int launch_t(const PairedMatrix* m, int dir, const void* in, long ies, long ivs, long ips,
             void* out, long oes, long ovs, long ops, int nvec, int nplanes, hipStream_t s) {
  const long total = (long)nvec * nplanes;
  if (total == 0) return NFGS_OK;
  const long nblk = (total + kVec - 1) / kVec;
  NFGS_REQUIRE(nblk < (1L << 31), NFGS_ERR_ARG, "transform: too many vectors");
  const T* ti = static_cast<const T*>(in);
  T* to = static_cast<T*>(out);
  const T* et = static_cast<const T*>(m->et[dir]);
  const T* ot = static_cast<const T*>(m->ot[dir]);
  if (ies == 1 && oes == 1) {
    hipLaunchKernelGGL((paired_transform_kernel<T, true>), dim3((unsigned)nblk), dim3(256), 0, s,
                       ti, ies, ivs, ips, to, oes, ovs, ops, et, ot, m->n, m->h, m->rows,
                       m->cols_pad, nvec, total);
  } else if (ivs == 1 && ovs == 1) {
    hipLaunchKernelGGL((paired_transform_kernel<T, false>), dim3((unsigned)nblk), dim3(256), 0, s,
                       ti, ies, ivs, ips, to, oes, ovs, ops, et, ot, m->n, m->h, m->rows,
                       m->cols_pad, nvec, total);
  } else {
    set_error("transform: need unit element strides or unit vector strides on both sides");
    return NFGS_ERR_ARG;
  }
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace

int launch_transform(const nfgs_plan* p, const PairedMatrix* m, int direction, const void* in,
                     long ies, long ivs, long ips, void* out, long oes, long ovs, long ops,
                     int nvec, int nplanes, hipStream_t s) {
  NFGS_REQUIRE(direction == NFGS_FORWARD || direction == NFGS_INVERSE, NFGS_ERR_ARG,
               "transform: direction must be NFGS_FORWARD or NFGS_INVERSE");
  NFGS_REQUIRE(in != nullptr && out != nullptr && in != out, NFGS_ERR_ARG,
               "transform: in and out must be distinct non-null buffers");
  NFGS_REQUIRE(nvec >= 0 && nplanes >= 0, NFGS_ERR_ARG, "transform: negative counts");
  if (p->d.dtype == NFGS_F32)
    return launch_t<float>(m, direction, in, ies, ivs, ips, out, oes, ovs, ops, nvec, nplanes, s);
  return launch_t<double>(m, direction, in, ies, ivs, ips, out, oes, ovs, ops, nvec, nplanes, s);
}

}  // namespace nfgs

extern "C" int nfgs_transform(const nfgs_plan* plan, int kind, int direction, const void* in,
                              long in_elem_stride, long in_vec_stride, long in_plane_stride,
                              void* out, long out_elem_stride, long out_vec_stride,
                              long out_plane_stride, int n, int nvec, int nplanes, void* stream) {
  NFGS_REQUIRE(plan != nullptr, NFGS_ERR_ARG, "transform: null plan");
  const PairedMatrix* m = nfgs::find_matrix(plan, kind, n);
  NFGS_REQUIRE(m != nullptr, NFGS_ERR_ARG,
               "transform: no matrix of this kind and length in the plan");
  return nfgs::launch_transform(plan, m, direction, in, in_elem_stride, in_vec_stride,
                                in_plane_stride, out, out_elem_stride, out_vec_stride,
                                out_plane_stride, nvec, nplanes, nfgs::as_stream(stream));
}
