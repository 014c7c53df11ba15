// Lateral type-I transforms (DCT-I / DST-I and inverses) as paired MFMA GEMMs.
//
// The transform matrix is split into its even-column half E and odd-column
// half O (see PairedMatrix), so each vector costs two (h x n/2) products
// instead of one (n x n) product, finished with out[k] = E+O,
// out[n-1-k] = E-O.  Three kernels share that form:
//
// * FP32 plans with n <= 258: fixed-point digit products summed exactly on
//   the int8 MFMA (launch_i8), as accurate as FP64 products for FP32 output;
// * vectors whose contraction fits the registers (Q <= 33 paired rows per
//   lane group): the matrix-stationary persistent kernel (launch_v2);
// * longer vectors: 128 vectors per workgroup, K chunks of the data and of
//   nine 16-row output blocks of the matrix staged in LDS per step, the next
//   chunk's loads in flight during the current chunk's MFMAs
//   (paired_transform_wide).
//
// The last two compute in FP64 (v_mfma_f64_16x16x4) for both plan dtypes:
// FP32 data is widened when staged and each output is rounded once when
// stored.  Two layouts share each kernel: ELEM_CONTIG (elements of a vector
// are adjacent, the x axis) computes C = X * T^T so that lanes hold
// consecutive output elements; otherwise (vectors adjacent, the y axis) it
// computes C = T * X so that lanes hold consecutive vectors.

#include <algorithm>
#include <type_traits>

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

// Exact-arithmetic (FP64 product) paired transform for long vectors.  A
// workgroup of 512 lanes (8 wavefronts) takes 128 vectors and one group of
// NB 16-row output blocks; each wavefront owns 16 vectors.  The contraction
// runs over chunks of 16 paired rows: the chunk's data (widened to FP64) and
// the group's 16 x 16*NB matrix entries of E and O are staged in LDS, in two
// buffers with one barrier per chunk: after its MFMAs on chunk c a wave
// writes chunk c+1 (held in registers) to the other buffer and issues the
// loads of chunk c+2 (lds_barrier leaves them in flight).  For f64 16x16x4 lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator r holds
// D[l/16 + 4r][l%16].  Every output is rounded once to the storage type.
constexpr int kWVec = 128;  // vectors per workgroup
constexpr int kWNB = 9;     // most 16-row output blocks per group
constexpr int kWXS = 17;    // LDS stride (doubles) of a vector's 16 paired elements

template <typename TS, bool ELEM_CONTIG, int NB>
__global__ __launch_bounds__(512) void paired_transform_wide(
    const TS* __restrict__ in, long ies, long ivs, long ips, TS* __restrict__ out, long oes,
    long ovs, long ops, const double* __restrict__ et, const double* __restrict__ ot, int n,
    int h, int rows, int cols_pad, int nvec, long nvec_total, int ngroups) {
  // LDS row stride (doubles) of a matrix chunk: 16 NB = 16 mod 32, so the
  // four rows a wavefront reads at once (k = 4 ks + lane/16) fall on two
  // disjoint halves of the 64 banks, each bank hit twice (the minimum for
  // 64 eight-byte reads)
  constexpr int MC = NB * 16;
  static_assert(NB % 2 == 1, "odd block counts keep the matrix rows conflict-free");
  constexpr int MS = (16 * NB * 16 + 511) / 512;  // matrix entries per lane per chunk
  // two buffers: chunk c is read from buffer c&1 while chunk c+1 is written
  // to the other
  __shared__ double xe[2][kWVec * kWXS];
  __shared__ double xo[2][kWVec * kWXS];
  __shared__ double me[2][16 * MC];
  __shared__ double mo[2][16 * MC];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, l4 = lane >> 4;
  // workgroup -> (vector tile, output group): the groups of one tile are
  // adjacent logical workgroups, so they read the tile's data through one L2
  const int lt = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = lt / ngroups, grp = lt - tile * ngroups;
  const long g0 = (long)tile * kWVec;
  const int b0 = grp * NB;
  const int nchunks = (rows + 15) >> 4;
  // loader slot i of a lane: vector sv[i] of the tile, element sj[i] of a
  // chunk's 32 (paired row sj/2, parity sj&1)
  long ibase[8];
  int sv[8], sj[8];
  unsigned vmask = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int e = tid + 512 * i;
    sv[i] = ELEM_CONTIG ? (e >> 5) : (e & (kWVec - 1));
    sj[i] = ELEM_CONTIG ? (e & 31) : (e >> 7);
    const long g = g0 + sv[i];
    ibase[i] = 0;
    if (g < nvec_total) {
      const long p = g / nvec, q = g - p * nvec;
      ibase[i] = p * ips + q * ivs;
      vmask |= 1u << i;
    }
  }
  // matrix slot i of a lane: entry (mrow, column) of the group's chunk at
  // element offset moff + 16*c*cols_pad (the matrices are far below 4 GiB)
  int moff[MS], mrow[MS];
  unsigned mmask = 0;
#pragma unroll
  for (int i = 0; i < MS; ++i) {
    const int e = tid + 512 * i;
    const int r = e / (NB * 16), col = e - r * (NB * 16);
    mrow[i] = r;
    moff[i] = r * cols_pad + b0 * 16 + col;
    if (e < 16 * NB * 16 && b0 * 16 + col < cols_pad) mmask |= 1u << i;
  }
  TS rd[8];
  double rme[MS], rmo[MS];
  auto load_chunk = [&](int c) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int j = c * 32 + sj[i];
      rd[i] = ((vmask >> i) & 1) && j < n ? in[ibase[i] + (long)j * ies] : TS(0);
    }
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      const bool ok = ((mmask >> i) & 1) && mrow[i] < rows - c * 16;
      const unsigned o = (unsigned)(moff[i] + c * 16 * cols_pad) * 8u;
      rme[i] = ok ? ld(et, o) : 0.0;
      rmo[i] = ok ? ld(ot, o) : 0.0;
    }
  };
  auto store_chunk = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int idx = sv[i] * kWXS + (sj[i] >> 1);
      if (sj[i] & 1)
        xo[buf][idx] = (double)rd[i];
      else
        xe[buf][idx] = (double)rd[i];
    }
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      const int e = tid + 512 * i;
      if (e < 16 * NB * 16) {
        const int r = e / (NB * 16), col = e - r * (NB * 16);
        me[buf][r * MC + col] = rme[i];
        mo[buf][r * MC + col] = rmo[i];
      }
    }
  };
  f64x4 accE[NB], accO[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    accE[b] = f64x4{0, 0, 0, 0};
    accO[b] = f64x4{0, 0, 0, 0};
  }
  load_chunk(0);
  store_chunk(0);
  if (nchunks > 1) load_chunk(1);
  lds_barrier();
  const int v0 = (wave * 16 + l16) * kWXS;
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int k = ks * 4 + l4;
      const double de = xe[buf][v0 + k], dod = xo[buf][v0 + k];
      // every block of the group runs (blocks past h multiply zero matrix
      // entries or rows the epilogue drops), so the loop has no branches
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const double m0 = me[buf][k * MC + b * 16 + l16];
        const double m1 = mo[buf][k * MC + b * 16 + l16];
        if (ELEM_CONTIG) {
          accE[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(de, m0, accE[b], 0, 0, 0);
          accO[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(dod, m1, accO[b], 0, 0, 0);
        } else {
          accE[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(m0, de, accE[b], 0, 0, 0);
          accO[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(m1, dod, accO[b], 0, 0, 0);
        }
      }
    }
    // chunk c+1 into the other buffer (its last readers finished before the
    // previous barrier), then the loads of chunk c+2
    if (c + 1 < nchunks) {
      store_chunk(buf ^ 1);
      if (c + 2 < nchunks) load_chunk(c + 2);
    }
    lds_barrier();
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int k, v;
      if (ELEM_CONTIG) {
        v = wave * 16 + l4 + 4 * r;
        k = (b0 + b) * 16 + l16;
      } else {
        k = (b0 + b) * 16 + l4 + 4 * r;
        v = wave * 16 + l16;
      }
      const long g = g0 + v;
      if (k < h && g < nvec_total) {
        const long p = g / nvec, q = g - p * nvec;
        TS* o = out + p * ops + q * ovs;
        const double e = accE[b][r], od = accO[b][r];
        o[(long)k * oes] = TS(e + od);
        const int k2 = n - 1 - k;
        if (k2 != k) o[(long)k2 * oes] = TS(e - od);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// v2: matrix-stationary persistent kernel.  A workgroup has one wavefront per
// 16-row output block (NKB blocks per wave); each wave keeps its E and O
// fragments for every contraction step in registers for the whole kernel.
// Tiles of V vectors stream through LDS: the next tile's global loads are
// issued before the current tile's MFMAs, so HBM latency hides behind the
// matrix work.  LDS holds the even and odd halves of the tile de-interleaved.
// ---------------------------------------------------------------------------
template <typename T, bool EC, int QMAX, int NKB, int PF, int V, int QC, typename TS = T>
__global__ __launch_bounds__(512) void paired_transform_v2(
    const TS* __restrict__ in, long ies, long ivs, long ips, TS* __restrict__ out, long oes,
    long ovs, long ops, const T* __restrict__ et, const T* __restrict__ ot, int n, int h,
    int ne, int cols_pad, int nvec, long nvec_total, int S, int hb, int rem) {
  using A = typename Mfma<T>::acc;
  typedef T vec4 __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // Contraction order: lane group l4 takes j' = l4*Q + ks for ks < Q, so a
  // lane's successive steps are adjacent in LDS and one 4-wide LDS read feeds
  // four MFMAs.  In LDS the tile row of vector v holds segment l4 at
  // [l4*Qp, l4*Qp + Q) (Qp = Q rounded up to 4; the gap stays zero).  QC > 0
  // fixes Q at compile time so the MFMA sequence has no branches.
  const int Q = QC > 0 ? QC : (ne + 3) >> 2;
  const int Qp = (Q + 3) & ~3;
  T* xe = reinterpret_cast<T*>(smem_raw);
  T* xo = xe + V * S;
  T* mrem = xo + V * S;  // [2][rem][ne]: E then O columns of the remainder rows
  const int trash = 2 * V * S + 2 * rem * ne;  // word after mrem, never read
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, l4 = lane >> 4;
  auto pos = [&](int jp) { const int sg = jp / Q; return sg * Qp + (jp - sg * Q); };

  for (int e = tid; e < 2 * V * S; e += nthr) xe[e] = T(0);
  for (int e = tid; e < 2 * rem * ne; e += nthr) {
    const int half = e / (rem * ne), rr = (e / ne) % rem, jp = e % ne;
    const T* src = half ? ot : et;
    mrem[e] = src[(long)jp * cols_pad + hb * 16 + rr];
  }

  T mE[NKB][QMAX], mO[NKB][QMAX];
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb) {
    const int blk = wave * NKB + kb;
    const int col = blk * 16 + l16;
#pragma unroll
    for (int ks = 0; ks < QMAX; ++ks) {
      const int jp = l4 * Q + ks;
      const bool ok = blk < hb && ks < Q;
      mE[kb][ks] = ok ? et[(long)jp * cols_pad + col] : T(0);
      mO[kb][ks] = ok ? ot[(long)jp * cols_pad + col] : T(0);
    }
  }

  const long ntiles = (nvec_total + V - 1) / V;
  // Loader mapping: each thread owns one vector of the tile and a strided
  // subset of its elements (EC: nthr/V consecutive threads per vector;
  // otherwise V consecutive threads cover V consecutive vectors).
  const int myv = EC ? tid / (nthr / V) : tid % V;
  const int j0 = EC ? tid % (nthr / V) : tid / V;
  const int jstep = nthr / V;
  TS pf[PF];  // prefetched elements in the storage type (widened at commit)
  bool pvalid = false;
  // Loads are unconditional: a vector past the end reads vector 0 and
  // elements past n re-read element n-1.  The raw values go to registers and
  // are masked only when they are committed to LDS, so nothing waits on the
  // loads before the MFMA work they overlap.
  // Offset of vector tp*nvec + tq + v in a (plane, vector) layout.  The
  // tile base is divided once per tile; when nvec >= V a tile offset wraps
  // into the next plane at most once.
  auto vec_off = [&](long tp, long tq, int v, long ps, long vs) -> long {
    long q = tq + v, p = tp;
    if (nvec >= V) {
      const bool w = q >= nvec;
      q -= w ? nvec : 0;
      p += w ? 1 : 0;
    } else {
      const long d = q / nvec;
      p += d;
      q -= d * nvec;
    }
    return p * ps + q * vs;
  };
  // (plane, vector) coordinates of tile t's first vector, advanced by the
  // grid stride without a divide
  const long dstep = (long)gridDim.x * V;
  const long sp = dstep / nvec, sq = dstep - sp * nvec;
  auto advance = [&](long& tp, long& tq) {
    tp += sp;
    tq += sq;
    if (tq >= nvec) {
      tq -= nvec;
      ++tp;
    }
  };
  auto load_tile = [&](long t, long tp, long tq) {
    pvalid = t * V + myv < nvec_total;
    const long base = pvalid ? vec_off(tp, tq, myv, ips, ivs) : 0;
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int j = j0 + jstep * i;
      const int jc = j < n ? j : n - 1;
      pf[i] = in[base + (long)jc * ies];
    }
  };
  long tile = xcd_tile(blockIdx.x, gridDim.x);
  long ctp = tile * V / nvec, ctq = tile * V - ctp * nvec;
  if (tile < ntiles) load_tile(tile, ctp, ctq);
  lds_barrier();

  for (; tile < ntiles; tile += gridDim.x) {
    // Every register is stored (elements past n go to a trash word after
    // the tile), so the waits on the loads sit in straight-line code.
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int j = j0 + jstep * i;
      const int idx = j < n ? (j & 1) * V * S + myv * S + pos(j >> 1) : trash;
      xe[idx] = pvalid ? T(pf[i]) : T(0);
    }
    lds_barrier();
    // the last tile re-reads itself instead of branching around the loads
    const long next = tile + gridDim.x;
    const long otp = ctp, otq = ctq;
    advance(ctp, ctq);
    if (next < ntiles)
      load_tile(next, ctp, ctq);
    else
      load_tile(tile, otp, otq);
    // keep the loads ahead of the MFMA section they overlap
    __builtin_amdgcn_sched_barrier(0);

    A accE[V / 16][NKB], accO[V / 16][NKB];
#pragma unroll
    for (int g = 0; g < V / 16; ++g)
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        accE[g][kb] = A{0, 0, 0, 0};
        accO[g][kb] = A{0, 0, 0, 0};
      }
    if (QC > 0) {
      // static schedule: fragments of step group kq+1 are read while the
      // MFMAs of group kq run
      constexpr int KQ = (QC + 3) / 4;
      vec4 ce[V / 16], co[V / 16];
      const int rb = l16 * S + l4 * Qp;
#pragma unroll
      for (int g = 0; g < V / 16; ++g) {
        ce[g] = *reinterpret_cast<const vec4*>(xe + g * 16 * S + rb);
        co[g] = *reinterpret_cast<const vec4*>(xo + g * 16 * S + rb);
      }
#pragma unroll
      for (int kq = 0; kq < KQ; ++kq) {
        vec4 ne_[V / 16], no_[V / 16];
        if (kq + 1 < KQ) {
#pragma unroll
          for (int g = 0; g < V / 16; ++g) {
            ne_[g] = *reinterpret_cast<const vec4*>(xe + g * 16 * S + rb + (kq + 1) * 4);
            no_[g] = *reinterpret_cast<const vec4*>(xo + g * 16 * S + rb + (kq + 1) * 4);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int ks = kq * 4 + u;
          if (ks < QC) {
#pragma unroll
            for (int g = 0; g < V / 16; ++g) {
#pragma unroll
              for (int kb = 0; kb < NKB; ++kb) {
                if (EC) {
                  accE[g][kb] = Mfma<T>::mma(ce[g][u], mE[kb][ks], accE[g][kb]);
                  accO[g][kb] = Mfma<T>::mma(co[g][u], mO[kb][ks], accO[g][kb]);
                } else {
                  accE[g][kb] = Mfma<T>::mma(mE[kb][ks], ce[g][u], accE[g][kb]);
                  accO[g][kb] = Mfma<T>::mma(mO[kb][ks], co[g][u], accO[g][kb]);
                }
              }
            }
          }
        }
        if (kq + 1 < KQ) {
#pragma unroll
          for (int g = 0; g < V / 16; ++g) {
            ce[g] = ne_[g];
            co[g] = no_[g];
          }
        }
      }
    } else {
#pragma unroll
    for (int kq = 0; kq < (QMAX + 3) / 4; ++kq) {
      if (kq * 4 < Q) {
#pragma unroll
        for (int g = 0; g < V / 16; ++g) {
          const int ix = (g * 16 + l16) * S + l4 * Qp + kq * 4;
          const vec4 ve = *reinterpret_cast<const vec4*>(xe + ix);
          const vec4 vo = *reinterpret_cast<const vec4*>(xo + ix);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int ks = kq * 4 + u;
            if (ks < QMAX && ks < Q) {
#pragma unroll
              for (int kb = 0; kb < NKB; ++kb) {
                if (EC) {
                  accE[g][kb] = Mfma<T>::mma(ve[u], mE[kb][ks], accE[g][kb]);
                  accO[g][kb] = Mfma<T>::mma(vo[u], mO[kb][ks], accO[g][kb]);
                } else {
                  accE[g][kb] = Mfma<T>::mma(mE[kb][ks], ve[u], accE[g][kb]);
                  accO[g][kb] = Mfma<T>::mma(mO[kb][ks], vo[u], accO[g][kb]);
                }
              }
            }
          }
        }
      }
    }
    }
#pragma unroll
    for (int g = 0; g < V / 16; ++g)
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        const int blk = wave * NKB + kb;
        if (blk >= hb) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int k, v;
          if (EC) {
            v = g * 16 + Mfma<T>::row(l4, r);
            k = blk * 16 + l16;
          } else {
            k = blk * 16 + Mfma<T>::row(l4, r);
            v = g * 16 + l16;
          }
          const long gi = tile * V + v;
          if (k < h && gi < nvec_total) {
            TS* o = out + vec_off(otp, otq, v, ops, ovs);
            const T e = accE[g][kb][r], od = accO[g][kb][r];
            o[(long)k * oes] = TS(e + od);
            const int k2 = n - 1 - k;
            if (k2 != k) o[(long)k2 * oes] = TS(e - od);
          }
        }
      }
    // remainder rows: 16 lanes per (vector, row) pair, shuffle reduction
    for (int pr = tid >> 4; pr < V * rem; pr += nthr >> 4) {
      const int v = pr % V, rr = pr / V;
      const int part = tid & 15;
      const T* cE = mrem + rr * ne;
      const T* cO = mrem + (rem + rr) * ne;
      T e = T(0), od = T(0);
      for (int jp = part; jp < ne; jp += 16) {
        const int ix = v * S + pos(jp);
        e += cE[jp] * xe[ix];
        od += cO[jp] * xo[ix];
      }
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 16);
        od += __shfl_xor(od, off, 16);
      }
      const int k = hb * 16 + rr;
      const long gi = tile * V + v;
      if (part == 0 && k < h && gi < nvec_total) {
        TS* o = out + vec_off(otp, otq, v, ops, ovs);
        o[(long)k * oes] = TS(e + od);
        const int k2 = n - 1 - k;
        if (k2 != k) o[(long)k2 * oes] = TS(e - od);
      }
    }
    lds_barrier();
  }
}

// ---------------------------------------------------------------------------
// FP32 plans: exact-integer products on the int8 MFMA.
//
// Each tile of 16 vectors is brought to fixed point with a power-of-two scale
// per vector (|X| < 2^30) and split into four signed base-256 digits,
// X = sum_p a_p 256^(3-p) with a_p in [-128, 127]; the matrices are held the
// same way (PairedMatrix::dig).  v_mfma_i32_16x16x64_i8 sums digit products
// exactly in int32, and the ten digit pairs of weight >= 256^-3 relative to
// the leading pair are kept, in four int32 accumulators by weight.  The
// accumulators combine exactly in FP64 and each output is rounded once to
// FP32, so the result is as accurate as an FP64-arithmetic product (the
// dropped pairs weigh about 2^-30 of the vector's largest element) while the
// products run at int8 rate.  A tail paired row (n = 128*nch + 1 or + 2) is
// added in FP64, and so is a remainder output row h-1 when h = 16*hb + 1: it
// is a dot product of the raw elements with row h-1 of T.
//
// Loader lanes are (vector v, offset ro < 32) pairs holding elements
// j = ro + 32r: all of one parity, so four of them (r = 4c..4c+3) pack into
// one digit word per digit plane with two rounds of v_perm_b32.  The K order
// this defines (paired row j' = ro/2 + 16r) is the order of the matrix
// digits; integer sums do not depend on it.  ELEM_CONTIG puts a vector's 32
// offsets in one half-wavefront (coalesced along the vector) and otherwise
// 16 vectors times 4 offsets per wavefront (coalesced across vectors).  Wave w
// owns output rows [16w, 16w+16) with its matrix digits in registers; the
// data digits of a tile are read from LDS by every wave.
//
// The tiles of a workgroup run as a pipeline with one barrier per step: step
// i computes the products of tile i, the digits of tile i+1, the per-vector
// maxima of tile i+2 and issues the loads of tile i+4 (four register sets,
// rotated by unrolling the step loop four times).  Addresses of a full tile
// are a uniform base (its first vector) plus a lane offset that does not
// change between tiles except across a plane boundary.  FULL (n >= 128*nch)
// drops the masks of elements past n.
// ---------------------------------------------------------------------------
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef f32x4 f32x4u __attribute__((aligned(4)));

// (plane, vector) of the first vector of a workgroup's successive tiles
struct TileCursor {
  int p, q;
  __device__ __forceinline__ void next(int dp, int dq, int nvec) {
    p += dp;
    q += dq;
    if (q >= nvec) {
      q -= nvec;
      ++p;
    }
  }
  // next() when go holds, by selects (no branch)
  __device__ __forceinline__ void next_if(bool go, int dp, int dq, int nvec) {
    const int nq = q + dq, w = nq >= nvec ? 1 : 0;
    p = go ? p + dp + w : p;
    q = go ? nq - w * nvec : q;
  }
};

template <bool EC, int NCH, bool TAIL, bool REM, bool FULL>
__global__ __launch_bounds__(512) void paired_transform_i8(
    const float* __restrict__ in, long ies, long ivs, long ips, float* __restrict__ out,
    long oes, long ovs, long ops, const i32x4* __restrict__ dig, const double* __restrict__ mtail,
    const double* __restrict__ mrem, int shift, int n, int h, int hb, int nvec, int nvec_total) {
  constexpr int RSB = 64 * NCH + 16;  // bytes per vector row of one digit plane
  constexpr int PLANE = 16 * RSB;     // one (half, digit) plane of a tile
  constexpr int DBUF = 8 * PLANE;     // one tile: 2 halves x 4 digits
  constexpr int NR = 4 * NCH;         // main elements per loader lane
  __shared__ __align__(16) unsigned char dl[2 * DBUF];
  __shared__ unsigned smax[3][16];
  __shared__ double ssc[2][16];  // 2^-(fixed-point exponent), NaN for non-finite vectors
  __shared__ double stail[2][2][16];  // tail elements of the E and O halves
  __shared__ double smt[2][16 * 8];
  __shared__ double rpart[REM && !EC ? 2 : 1][8][2][16];
  __shared__ double sjunk[64];  // target of LDS writes by lanes that own no slot
  __shared__ double smr[REM ? NR + 1 : 1][32];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l16 = lane & 15, g4 = lane >> 4;
  const int lv = EC ? 2 * wave + (lane >> 5) : l16;
  const int lro = EC ? (lane & 31) : 4 * wave + g4;
  const int lhalf = lro & 1, lm = lro >> 1;
  const bool mw = wave < hb;
  const int krem = 16 * hb;
  const double qnan = __longlong_as_double(0x7ff8000000000000LL);

  i32x4 md[2][4][NCH];
#pragma unroll
  for (int hf = 0; hf < 2; ++hf)
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int c = 0; c < NCH; ++c)
        md[hf][p][c] = mw ? dig[(((hf * 4 + p) * NCH + c) * hb + wave) * 64 + lane]
                          : i32x4{0, 0, 0, 0};
  if (TAIL)
    for (int e = tid; e < 32 * hb; e += blockDim.x) smt[e / (16 * hb)][e % (16 * hb)] = mtail[e];
  // remainder-row weights by element (zero past n): smr[i][lro] weighs element
  // lro + 32 i, smr[NR][lro] the tail element 128*NCH + lro
  if (REM)
    for (int e = tid; e < (NR + 1) * 32; e += blockDim.x) {
      const int i = e >> 5, ro = e & 31;
      const int el = i < NR ? ro + 32 * i : 128 * NCH + ro;
      smr[i][ro] = (i < NR || ro < 2) && el < n ? mrem[el] : 0.0;
    }
  if (tid < 48) smax[tid >> 4][tid & 15] = 0u;
  // The set-up loads are complete before the pipeline starts (vmcnt(0)), so
  // the waits inside it count only tile loads.
  __builtin_amdgcn_s_waitcnt(0x0F70);

  const int ntiles = (nvec_total + 15) >> 4;
  const int G = gridDim.x;
  const int tile0 = xcd_tile(blockIdx.x, G);
  const int cnt = tile0 < ntiles ? (ntiles - 1 - tile0) / G + 1 : 0;
  // tile cursors for the loads, the digits (remainder row) and the products
  const int dp = (int)((16L * G) / nvec), dq = (int)(16L * G - (long)dp * nvec);
  TileCursor c0;
  c0.p = (tile0 * 16) / nvec;
  c0.q = tile0 * 16 - c0.p * nvec;
  TileCursor cl = c0, cd = c0, cp = c0;
  // a full tile whose vectors cross at most one plane boundary takes the
  // uniform-base addressing
  auto full_tile = [&](int j) { return (FULL || nvec >= 16) && (tile0 + j * G) * 16 + 15 < nvec_total; };
  const long iwrap = ips - (long)nvec * ivs, owrap = ops - (long)nvec * ovs;

  auto voff = [&](int gv, long ps, long vs) {
    const int p = gv / nvec, q = gv - p * nvec;
    return (long)p * ps + (long)q * vs;
  };
  // Byte offsets of full tiles in 32-bit arithmetic: the host keeps both
  // arrays below 4 GiB, so offsets of real elements are exact modulo 2^32.
  // A tile's offset is a uniform part (its first vector) plus a
  // lane-constant part plus the plane wrap of lanes past the plane's end.
  const unsigned ies4 = (unsigned)ies * 4u, ivs4 = (unsigned)ivs * 4u, ips4 = (unsigned)ips * 4u;
  const unsigned oes4 = (unsigned)oes * 4u, ovs4 = (unsigned)ovs * 4u, ops4 = (unsigned)ops * 4u;
  const unsigned iwrap4 = (unsigned)iwrap * 4u, owrap4 = (unsigned)owrap * 4u;
  auto tile_base = [&](const TileCursor& c, unsigned ps4, unsigned vs4) {
    return (unsigned)c.p * ps4 + (unsigned)c.q * vs4;
  };
  const unsigned lin = (unsigned)lv * ivs4 + (unsigned)lro * ies4;  // load lane part
  // Output byte offset of vector v of the tile at cursor c (global tile t).
  // When a plane holds at least 16 vectors a tile spans at most two planes,
  // so the cursor form holds for every tile; lanes past the last vector get
  // an offset they never store to.
  const bool wide = FULL || nvec >= 16;  // FULL kernels run on planes of >= 16 vectors
  auto out_off = [&](const TileCursor& c, int t, int v) -> unsigned {
    if (wide) return tile_base(c, ops4, ovs4) + (unsigned)v * ovs4 + (c.q + v >= nvec ? owrap4 : 0u);
    return (unsigned)(voff(t * 16 + v, ops, ovs) * 4);
  };
  const int etail = min(128 * NCH + lro, n - 1);  // tail element (clamped)
  // Loads of the j-th tile of this workgroup.  Lanes past the last vector
  // read vector 0 and elements past n re-read element n-1 (masked when
  // digits are formed).  Every step issues the same loads, so the waits on
  // them are counted exactly; steps past the last tile read vector 0.
  auto load = [&](int j, float (&r)[NR], float& rt) {
    // byte offset of element ro of the lane's vector (arrays below 4 GiB)
    unsigned o;
    if (FULL) {
      // planes of >= 16 vectors: the cursor form holds for every tile
      const bool ok = j < cnt && (tile0 + j * G) * 16 + lv < nvec_total;
      o = ok ? tile_base(cl, ips4, ivs4) + lin + (cl.q + lv >= nvec ? iwrap4 : 0u)
             : (unsigned)lro * ies4;
    } else {
      int gv = (tile0 + j * G) * 16 + lv;
      gv = (j < cnt && gv < nvec_total) ? gv : 0;
      o = (unsigned)((voff(gv, ips, ivs) + (long)lro * ies) * 4);
    }
    const unsigned es = (unsigned)(ies * 4);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int e = lro + 32 * i;
      r[i] = ld(in, o + (unsigned)((FULL || e < n ? 32 * i : n - 1 - lro)) * es);
    }
    if (TAIL || REM) rt = ld(in, o + (unsigned)(etail - lro) * es);
    cl.next(dp, dq, nvec);
  };
  auto has_tail_elem = [&]() { return (TAIL || REM) && lro < 2 && 128 * NCH + lro < n; };
  auto valid = [&](int i) { return FULL || lro + 32 * i < n; };

  // ---- largest magnitude per vector (|x| bits compare as unsigned) ----
  auto vmax = [&](int j, const float (&r)[NR], float rt) {
    unsigned mx = 0u;
#pragma unroll
    for (int i = 0; i < NR; ++i)
      if (valid(i)) mx = max(mx, __float_as_uint(r[i]) & 0x7fffffffu);
    if (has_tail_elem()) mx = max(mx, __float_as_uint(rt) & 0x7fffffffu);
    // ELEM_CONTIG: a vector spans two rows, each reduced by DPP; otherwise
    // rows hold offsets g4 = 0..3 of 16 vectors, paired by the swizzle.  One
    // lane per row (pair) merges into the tile's maxima.
    if (EC) {
      mx = row_max_u32(mx);
      if ((lro & 15) == 0) atomicMax(&smax[j % 3][lv], mx);
    } else {
      mx = max(mx, swz16_u32(mx));
      if ((g4 & 1) == 0) atomicMax(&smax[j % 3][lv], mx);
    }
  };

  // ---- digits of tile j, in registers: the words of the four digit planes,
  // the scale exponent and the remainder-row sums (no memory writes) ----
  struct Digits {
    unsigned w[4 * NCH];
    int s;
    bool bad;
  };
  auto digit_words = [&](int j, const float (&r)[NR], float rt, Digits& D) {
    const unsigned mb = smax[(j + 3) % 3][lv];
    D.bad = mb >= 0x7f800000u;
    D.s = (mb == 0u || D.bad) ? 0 : 156 - (int)(mb >> 23);  // max|x| * 2^s < 2^30
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      unsigned d[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ri = 4 * c + i;
        const float x = valid(ri) ? r[ri] : 0.0f;
        const int X = (int)ldexpf(x, D.s);
        d[i] = ((unsigned)X + 0x808080u) ^ 0x808080u;  // bytes = signed digits
      }
      const unsigned p01l = __builtin_amdgcn_perm(d[1], d[0], 0x05010400u);
      const unsigned p01h = __builtin_amdgcn_perm(d[1], d[0], 0x07030602u);
      const unsigned p23l = __builtin_amdgcn_perm(d[3], d[2], 0x05010400u);
      const unsigned p23h = __builtin_amdgcn_perm(d[3], d[2], 0x07030602u);
      // digit plane p holds byte 3-p of every element
      D.w[4 * c + 0] = __builtin_amdgcn_perm(p23h, p01h, 0x07060302u);
      D.w[4 * c + 1] = __builtin_amdgcn_perm(p23h, p01h, 0x05040100u);
      D.w[4 * c + 2] = __builtin_amdgcn_perm(p23l, p01l, 0x07060302u);
      D.w[4 * c + 3] = __builtin_amdgcn_perm(p23l, p01l, 0x05040100u);
    }
  };

  // ---- digits of tile j into LDS with its scale and tail (branch-free:
  // lanes that own no slot write to sjunk) ----
  auto digit_store = [&](int j, const Digits& D, float rt) {
    const int buf = j & 1;
    const bool sw = EC ? lro == 0 : (wave == 0 && g4 == 0);
    *(sw ? &ssc[buf][lv] : &sjunk[lane]) = D.bad ? qnan : ldexp(1.0, 24 - shift - D.s);
    unsigned* dst = reinterpret_cast<unsigned*>(dl + buf * DBUF + lhalf * 4 * PLANE + lv * RSB) + lm;
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int p = 0; p < 4; ++p) dst[(p * PLANE + c * 64) / 4] = D.w[4 * c + p];
    if (TAIL) *(lro < 2 ? &stail[buf][lro & 1][lv] : &sjunk[lane]) = has_tail_elem() ? (double)rt : 0.0;
  };
  // ---- remainder row of tile j: an FP64 dot product of the raw elements
  // per parity (elements past n re-read element n-1 of the vector and weigh
  // zero), stored (ELEM_CONTIG) or left as partials ----
  auto digit_rem = [&](int j, const Digits& D, const float (&r)[NR], float rt, bool real) {
    if (!REM) return;
    const int buf = j & 1;
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < NR; ++i) acc = fma(smr[i][lro], (double)r[i], acc);
    acc = fma(smr[NR][lro], (double)rt, acc);
    if (EC) {
      // same-parity lanes of each row, the two rows of the half-wavefront,
      // then E (even) with O (odd)
      acc = row_parity_sum_f64(acc);
      acc += swz16_f64(acc);
      const double re = acc, ro = dpp_f64<kDppXor1>(acc);
      const int gv = (tile0 + j * G) * 16 + lv;
      if (real && lro == 0 && gv < nvec_total) {
        const unsigned ob = out_off(cd, tile0 + j * G, lv);
        const float nanf = __int_as_float(0x7fc00000);
        st(out, ob + (unsigned)krem * 4u, D.bad ? nanf : (float)(re + ro));
        if (n - 1 - krem != krem)
          st(out, ob + (unsigned)(n - 1 - krem) * 4u, D.bad ? nanf : (float)(re - ro));
      }
    } else {
      acc += __shfl_xor(acc, 32);  // offsets g4 and g4 + 2 share a parity
      *(g4 < 2 ? &rpart[buf][wave][g4 & 1][lv] : &sjunk[lane]) = acc;
    }
  };

  // FP64 value of the four accumulators, times the vector's scale
  auto combine = [&](const i32x4 (&a)[4], int r) -> double {
    const int t01 = a[0][r] * 256 + a[1][r];
    const int t23 = a[2][r] * 256 + a[3][r];
    return fma((double)t01, 65536.0, (double)t23);
  };

  // ---- products of tile j (waves without rows multiply zero digits) ----
  auto mfma_tile = [&](int j, i32x4 (&acc)[2][4]) {
    const int buf = j & 1;
#pragma unroll
    for (int hf = 0; hf < 2; ++hf)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[hf][q] = i32x4{0, 0, 0, 0};
    const unsigned char* xb = dl + buf * DBUF + l16 * RSB + 16 * g4;
    constexpr int PI[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};  // matrix digit
    constexpr int PJ[10] = {3, 2, 1, 0, 2, 1, 0, 1, 0, 0};  // data digit
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        i32x4 xd[4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
          xd[p] = *reinterpret_cast<const i32x4*>(xb + (hf * 4 + p) * PLANE + c * 64);
#pragma unroll
        for (int u = 0; u < 10; ++u) {
          const int L = PI[u] + PJ[u];
          acc[hf][L] = EC ? __builtin_amdgcn_mfma_i32_16x16x64_i8(md[hf][PI[u]][c], xd[PJ[u]],
                                                                 acc[hf][L], 0, 0, 0)
                          : __builtin_amdgcn_mfma_i32_16x16x64_i8(xd[PJ[u]], md[hf][PI[u]][c],
                                                                 acc[hf][L], 0, 0, 0);
        }
      }
  };

  // ---- out[k] = E + O and out[n-1-k] = E - O of tile j ----
  // Accumulator register r of lane l holds row 4*(l/16) + r, column l%16:
  // ELEM_CONTIG rows are output rows (4 adjacent k per lane), otherwise rows
  // are vectors (4 adjacent vectors per lane).  The values are formed by
  // every lane; stores are predicated (real: j is one of this block's tiles).
  auto epilogue = [&](int j, const i32x4 (&acc)[2][4], bool real) {
    const int buf = j & 1;
    const int t = tile0 + j * G;
    f32x4 vp, vm;
    if (EC) {
      const int v = l16, gv = t * 16 + v;
      const double sc = ssc[buf][v];
      const int k0 = 16 * wave + 4 * g4;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + r;
        double e = combine(acc[0], r) * sc, od = combine(acc[1], r) * sc;
        if (TAIL) {
          e = fma(smt[0][k], stail[buf][0][v], e);
          od = fma(smt[1][k], stail[buf][1][v], od);
        }
        vp[r] = (float)(e + od);
        vm[r] = (float)(e - od);
      }
      if (real && mw && gv < nvec_total) {
        // byte offset of the vector's element 0 (elements contiguous)
        const unsigned ob = out_off(cp, t, v);
        if (k0 + 3 < h && n - 4 - k0 > k0 + 3) {
          st(reinterpret_cast<f32x4u*>(out), ob + (unsigned)k0 * 4u, vp);
          st(reinterpret_cast<f32x4u*>(out), ob + (unsigned)(n - 4 - k0) * 4u,
             f32x4{vm[3], vm[2], vm[1], vm[0]});
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int k = k0 + r;
            if (k < h) {
              st(out, ob + (unsigned)k * 4u, vp[r]);
              if (n - 1 - k != k) st(out, ob + (unsigned)(n - 1 - k) * 4u, vm[r]);
            }
          }
        }
      }
    } else {
      const int k = 16 * wave + l16;
      const int v0 = 4 * g4, gv0 = t * 16 + v0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double sc = ssc[buf][v0 + r];
        double e = combine(acc[0], r) * sc, od = combine(acc[1], r) * sc;
        if (TAIL) {
          e = fma(smt[0][k], stail[buf][0][v0 + r], e);
          od = fma(smt[1][k], stail[buf][1][v0 + r], od);
        }
        vp[r] = (float)(e + od);
        vm[r] = (float)(e - od);
      }
      const int k2 = n - 1 - k;
      if (real && mw && k < h) {
        // the four vectors sit in one plane unless a boundary falls inside them
        const bool one = full_tile(j) && (cp.q + v0 >= nvec || cp.q + v0 + 3 < nvec);
        if (one) {
          // vectors contiguous: the lane's four vectors are one 16-byte run
          const unsigned ob = tile_base(cp, ops4, ovs4) + (unsigned)v0 * 4u +
                              (cp.q + v0 >= nvec ? owrap4 : 0u);
          st(reinterpret_cast<f32x4u*>(out), ob + (unsigned)k * oes4, vp);
          if (k2 != k) st(reinterpret_cast<f32x4u*>(out), ob + (unsigned)k2 * oes4, vm);
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int gv = gv0 + r;
            if (gv < nvec_total) {
              const unsigned ob = out_off(cp, t, v0 + r);
              st(out, ob + (unsigned)k * oes4, vp[r]);
              if (k2 != k) st(out, ob + (unsigned)k2 * oes4, vm[r]);
            }
          }
        }
      }
      if (REM && wave == 0 && lane < 16) {
        // remainder row of the tile from the eight waves' partials
        const int gv = t * 16 + lane;
        if (real && gv < nvec_total) {
          double e = 0.0, od = 0.0;
          for (int w = 0; w < 8; ++w) {
            e += rpart[buf][w][0][lane];
            od += rpart[buf][w][1][lane];
          }
          if (ssc[buf][lane] != ssc[buf][lane]) e = qnan;  // non-finite vector
          const unsigned ob = out_off(cp, t, lane);
          st(out, ob + (unsigned)krem * oes4, (float)(e + od));
          if (n - 1 - krem != krem) st(out, ob + (unsigned)(n - 1 - krem) * oes4, (float)(e - od));
        }
      }
    }
    cp.next(dp, dq, nvec);
  };

  // One pipeline step: products of tile i, digits of i+1, maxima of i+2,
  // loads of i+4.  Tile j's raw elements live in register set j % 4.  In an
  // ELEM_CONTIG steady step (every part on a real tile) the digit arithmetic
  // of tile i+1,
  // the matrix products of tile i, the digit writes and the epilogue
  // arithmetic form one branch-free region, so the digit work issues between
  // the products.  The parts of a step depend on each other only through the
  // barrier, so half of the waves take the products last: a SIMD's two
  // waves then issue matrix products and digit work at the same time.
  const bool late = (wave >> 2) & 1;
  auto work_steady = [&](int i, float (&rd)[NR], float td) {
    Digits D;
    digit_words(i + 1, rd, td, D);
    i32x4 acc[2][4];
    mfma_tile(i, acc);
    digit_store(i + 1, D, td);
    epilogue(i, acc, true);
    digit_rem(i + 1, D, rd, td, true);
    cd.next(dp, dq, nvec);
  };
  auto work_guarded = [&](int i, float (&rd)[NR], float td) {
    if (i >= 0 && i < cnt) {
      i32x4 acc[2][4];
      mfma_tile(i, acc);
      epilogue(i, acc, true);
    }
    Digits D;
    digit_words(i + 1, rd, td, D);
    digit_store(i + 1, D, td);
    digit_rem(i + 1, D, rd, td, i + 1 >= 0 && i + 1 < cnt);
    cd.next_if(i + 1 >= 0, dp, dq, nvec);
  };
  auto step = [&](bool steady, int i, float (&rd)[NR], float& td, float (&rx)[NR], float& tx,
                  float (&rl)[NR], float& tl) {
    if (!late) {
      if (steady) work_steady(i, rd, td);
      else work_guarded(i, rd, td);
    }
    vmax(i + 2, rx, tx);
    if (tid < 16) smax[(i + 3) % 3][tid] = 0u;  // tile i+3 gathers here next step
    load(i + 4, rl, tl);
    if (late) {
      if (steady) work_steady(i, rd, td);
      else work_guarded(i, rd, td);
    }
    lds_barrier();
  };

  // set 3 is read (as stale data) by the first step before its first load
  float r0[NR] = {}, r1[NR] = {}, r2[NR] = {}, r3[NR] = {};
  float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
  load(0, r0, t0);
  load(1, r1, t1);
  lds_barrier();
  for (int i = -2; i < cnt; i += 4) {
    // (vectors across lanes hold more registers per lane in the epilogue:
    // there the digit work follows the products)
    if (EC && i >= 0 && i + 4 < cnt) {
      step(true, i, r3, t3, r0, t0, r2, t2);
      step(true, i + 1, r0, t0, r1, t1, r3, t3);
      step(true, i + 2, r1, t1, r2, t2, r0, t0);
      step(true, i + 3, r2, t2, r3, t3, r1, t1);
    } else {
      step(false, i, r3, t3, r0, t0, r2, t2);
      step(false, i + 1, r0, t0, r1, t1, r3, t3);
      step(false, i + 2, r1, t1, r2, t2, r0, t0);
      step(false, i + 3, r2, t2, r3, t3, r1, t1);
    }
  }
}

int cu_count() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

// Launch v2 when the shape fits its register plan; returns -1 when it does not.
template <typename T, bool EC, typename TS>
int launch_v2(const PairedMatrix* m, int dir, const TS* in, long ies, long ivs, long ips, TS* out,
              long oes, long ovs, long ops, int nvec, long total, hipStream_t s) {
  // 16 vectors per tile: with FP64 operands wider tiles do not fit the
  // register file next to the matrix fragments.
  constexpr int V = 16;
  constexpr int QMAX = 33, NKB = 1, PF = 9;
  const int h = m->h, ne = m->rows, Q = (ne + 3) / 4;
  int hb = (h + 15) / 16, rem = 0;
  if (h >= 16 && h % 16 <= 2) {
    hb = h / 16;
    rem = h % 16;
  }
  if (Q > QMAX || hb > 8) return -1;
  const int nthr = 64 * hb;
  if (m->n > PF * (nthr / V)) return -1;
  // row stride >= 4*Qp, congruent to 4 mod 32 words: the 16-byte reads of
  // eight consecutive rows land on distinct bank groups
  const int Qp = (Q + 3) & ~3;
  int S = 4 * Qp;
  while (S % 32 != 4) ++S;
  const size_t lds = (2ul * V * S + 2ul * rem * ne + 1) * sizeof(T);
  if (lds > 64 * 1024) return -1;
  const long ntiles = (total + V - 1) / V;
  const long grid = std::min<long>(ntiles, 2L * cu_count());
#define NFGS_V2(QC_)                                                                      \
  hipLaunchKernelGGL((paired_transform_v2<T, EC, QMAX, NKB, PF, V, QC_, TS>), dim3((unsigned)grid), \
                     dim3(nthr), lds, s, in, ies, ivs, ips, out, oes, ovs, ops,                 \
                     static_cast<const T*>(m->et[dir]), static_cast<const T*>(m->ot[dir]),     \
                     m->n, h, ne, m->cols_pad, nvec, total, S, hb, rem)
  switch (Q) {
    case 33: NFGS_V2(33); break;
    case 32: NFGS_V2(32); break;
    case 25: NFGS_V2(25); break;
    case 24: NFGS_V2(24); break;
    case 17: NFGS_V2(17); break;
    case 16: NFGS_V2(16); break;
    case 13: NFGS_V2(13); break;
    case 12: NFGS_V2(12); break;
    case 9: NFGS_V2(9); break;
    case 8: NFGS_V2(8); break;
    default: NFGS_V2(0); break;
  }
#undef NFGS_V2
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

// Exact-integer FP32 transform (paired_transform_i8); persistent workgroups,
// as many per CU as fit.
template <bool EC, int NCH, bool TAIL, bool REM, bool FULL>
int launch_i8_t(const PairedMatrix* m, int dir, const float* in, long ies, long ivs, long ips,
                float* out, long oes, long ovs, long ops, int nvec, int total, hipStream_t s) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, paired_transform_i8<EC, NCH, TAIL, REM, FULL>, 512, 0) != hipSuccess ||
        nb < 1)
      nb = 1;
    per_cu = nb;
  }
  const int ntiles = (total + 15) / 16;
  const int pcu = m->blocks_per_cu > 0 ? std::min(m->blocks_per_cu, per_cu) : per_cu;
  const int grid = std::min(ntiles, pcu * cu_count());
  hipLaunchKernelGGL((paired_transform_i8<EC, NCH, TAIL, REM, FULL>), dim3(grid), dim3(512), 0, s,
                     in, ies, ivs, ips, out, oes, ovs, ops,
                     static_cast<const i32x4*>(m->dig[dir]),
                     static_cast<const double*>(m->mtail[dir]),
                     static_cast<const double*>(m->mrem[dir]), m->shift[dir], m->n, m->h, m->hb,
                     nvec, total);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

// Vectors of at least 128*nch elements, on planes of at least 16 vectors,
// use the unmasked kernels with the exact tail/remainder variant; others
// share a masked kernel that always adds the (then zero) tail row.
template <bool EC>
int launch_i8(const PairedMatrix* m, int dir, const float* in, long ies, long ivs, long ips,
              float* out, long oes, long ovs, long ops, int nvec, int total, hipStream_t s) {
#define NFGS_I8(NCH_, TAIL_, REM_, FULL_)                                                      \
  return launch_i8_t<EC, NCH_, TAIL_, REM_, FULL_>(m, dir, in, ies, ivs, ips, out, oes, ovs, ops, \
                                                  nvec, total, s)
  const bool full = m->n >= 128 * m->nch && nvec >= 16;
  if (m->nch == 1) {
    if (!full) {
      if (m->rem) NFGS_I8(1, true, true, false);
      NFGS_I8(1, true, false, false);
    }
    if (m->tail) {
      if (m->rem) NFGS_I8(1, true, true, true);
      NFGS_I8(1, true, false, true);
    }
    if (m->rem) NFGS_I8(1, false, true, true);
    NFGS_I8(1, false, false, true);
  }
  if (!full) {
    if (m->rem) NFGS_I8(2, true, true, false);
    NFGS_I8(2, true, false, false);
  }
  if (m->tail) {
    if (m->rem) NFGS_I8(2, true, true, true);
    NFGS_I8(2, true, false, true);
  }
  if (m->rem) NFGS_I8(2, false, true, true);
  NFGS_I8(2, false, false, true);
#undef NFGS_I8
}

template <typename TS, int NB>
int launch_wide_nb(const PairedMatrix* m, int dir, const TS* in, long ies, long ivs, long ips,
                   TS* out, long oes, long ovs, long ops, int nvec, long total, int ngroups,
                   hipStream_t s) {
  const long nblk = (total + kWVec - 1) / kWVec * ngroups;
  NFGS_REQUIRE(nblk < (1L << 31), NFGS_ERR_ARG, "transform: too many vectors");
  const double* et = static_cast<const double*>(m->et[dir]);
  const double* ot = static_cast<const double*>(m->ot[dir]);
  if (ies == 1 && oes == 1) {
    hipLaunchKernelGGL((paired_transform_wide<TS, true, NB>), dim3((unsigned)nblk), dim3(512), 0,
                       s, in, ies, ivs, ips, out, oes, ovs, ops, et, ot, m->n, m->h, m->rows,
                       m->cols_pad, nvec, total, ngroups);
  } else if (ivs == 1 && ovs == 1) {
    hipLaunchKernelGGL((paired_transform_wide<TS, false, NB>), dim3((unsigned)nblk), dim3(512), 0,
                       s, in, ies, ivs, ips, out, oes, ovs, ops, et, ot, m->n, m->h, m->rows,
                       m->cols_pad, nvec, total, ngroups);
  } else {
    set_error("transform: need unit element strides or unit vector strides on both sides");
    return NFGS_ERR_ARG;
  }
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

// Output blocks split into the fewest groups of at most nine, each group as
// even as the instantiated widths (5, 7, 9 blocks) allow.
template <typename TS>
int launch_wide(const PairedMatrix* m, int dir, const TS* in, long ies, long ivs, long ips,
                TS* out, long oes, long ovs, long ops, int nvec, long total, hipStream_t s) {
  const int hblocks = (m->h + 15) / 16;
  const int ngroups = (hblocks + kWNB - 1) / kWNB;
  const int need = (hblocks + ngroups - 1) / ngroups;
  if (need <= 5)
    return launch_wide_nb<TS, 5>(m, dir, in, ies, ivs, ips, out, oes, ovs, ops, nvec, total,
                                 ngroups, s);
  if (need <= 7)
    return launch_wide_nb<TS, 7>(m, dir, in, ies, ivs, ips, out, oes, ovs, ops, nvec, total,
                                 ngroups, s);
  return launch_wide_nb<TS, 9>(m, dir, in, ies, ivs, ips, out, oes, ovs, ops, nvec, total,
                               ngroups, s);
}

// Storage type TS (the plan dtype) with arithmetic type T.  FP32 plans run
// the products in FP64 (T = double): the matrices are FP64, elements are
// widened on their way into LDS and each output is rounded once when it is
// stored.  With FP32 MFMA every product-add rounds (about 3 ulp per length-
// 258 pass, three times an FFT's), and the modal solves amplify that noise
// into the derivatives of B and J.
template <typename T, typename TS>
int launch_t(const PairedMatrix* m, int dir, const void* in, long ies, long ivs, long ips,
             void* out, long oes, long ovs, long ops, int nvec, int nplanes, hipStream_t s) {
  const long total = (long)nvec * nplanes;
  if (total == 0) return NFGS_OK;
  const TS* ti = static_cast<const TS*>(in);
  TS* to = static_cast<TS*>(out);
  if constexpr (std::is_same<TS, float>::value) {
    // the digit kernels address both arrays with 32-bit byte offsets
    const long ispan = ((long)(nplanes - 1) * ips + (long)(nvec - 1) * ivs + (long)(m->n - 1) * ies + 1) * 4;
    const long ospan = ((long)(nplanes - 1) * ops + (long)(nvec - 1) * ovs + (long)(m->n - 1) * oes + 1) * 4;
    const bool pos = ies > 0 && ivs >= 0 && ips >= 0 && oes > 0 && ovs >= 0 && ops >= 0;
    if (m->hb > 0 && pos && total < (1L << 31) && ispan < (1L << 32) && ospan < (1L << 32)) {
      if (ies == 1 && oes == 1)
        return launch_i8<true>(m, dir, ti, ies, ivs, ips, to, oes, ovs, ops, nvec, (int)total, s);
      if (ivs == 1 && ovs == 1)
        return launch_i8<false>(m, dir, ti, ies, ivs, ips, to, oes, ovs, ops, nvec, (int)total, s);
    }
  }
  if (ies == 1 && oes == 1) {
    const int rc =
        launch_v2<T, true, TS>(m, dir, ti, ies, ivs, ips, to, oes, ovs, ops, nvec, total, s);
    if (rc >= 0) return rc;
  } else if (ivs == 1 && ovs == 1) {
    const int rc =
        launch_v2<T, false, TS>(m, dir, ti, ies, ivs, ips, to, oes, ovs, ops, nvec, total, s);
    if (rc >= 0) return rc;
  }
  // long vectors: 128 vectors per workgroup, nine 16-row output blocks per pass
  return launch_wide<TS>(m, dir, ti, ies, ivs, ips, to, oes, ovs, ops, nvec, total, s);
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
  if (p->d.dtype == NFGS_F32 && p->lateral_arith == NFGS_LATERAL_BF16X3) {
    const long total = (long)nvec * nplanes;
    if (total == 0) return NFGS_OK;
    return launch_bf16x3(m, direction, static_cast<const float*>(in), ies, ivs, ips,
                         static_cast<float*>(out), oes, ovs, ops, nvec, total, s);
  }
  if (p->d.dtype == NFGS_F32)
    return launch_t<double, float>(m, direction, in, ies, ivs, ips, out, oes, ovs, ops, nvec,
                                   nplanes, s);
  return launch_t<double, double>(m, direction, in, ies, ivs, ips, out, oes, ovs, ops, nvec,
                                  nplanes, s);
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
