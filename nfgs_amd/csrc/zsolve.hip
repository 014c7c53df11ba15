// Batched z-direction Thomas sweeps in modal space.
//
// One lane owns one horizontal mode and walks its column of nz planes; the
// lanes of a wavefront hold consecutive modes, so every plane access is one
// contiguous 256-byte (f32) run.  The forward elimination stores c' and d'
// into caller scratch (c') and the output (d'); the back substitution then
// overwrites the output with the solution.  Arithmetic follows the
// reference recurrences row by row (nfgs_solver.py:5579-5624, :567-608).

#include <type_traits>

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
    const T* __restrict__ kk, const T* __restrict__ bh, const T* __restrict__ qh,
    const T* __restrict__ gh, int nz, long plane, T inv_dz2, T two_inv_dz2, T dz, T dz2,
    T two_dz, T rho, T one_m_rho, T kb, T fu_rho, T rho_over_dz, double* zm_partial, double norm,
    bool div_src) {
  const long m = (long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
  const T ls = div_src ? l : T(1);  // the source is S / lambda, or S itself
  const T lo = inv_dz2, up = inv_dz2;
  const T diag = l - two_inv_dz2;
  const T b = bh[m];
  T cp, dp;
  {
    const T s1 = src[plane + m] / ls;
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
    const T s = sv / ls;
    const T den = diag - lo * cp;
    cp = up / den;
    dp = (s - lo * dp) / den;
    cs[(long)(r + 1) * plane + m] = cp;
    out[(long)(r + 1) * plane + m] = dp;
  });
  // Top Robin row with u_{N-2} eliminated through the PDE at plane N-1.
  {
    const T s = src[(long)(nz - 2) * plane + m] / ls;
    const T rl = (-l) * dz2 - T(2);
    const T rd = T(2) + two_dz * kk[m];
    // top Robin value g of (d_z + k) u (0 on the NFGS path)
    const T rr = (gh ? two_dz * gh[m] : T(0)) - s * dz2;
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
  const long m = (long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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

// ---------------------------------------------------------------------------
// Segmented sweeps: the column of a mode is split into segments of SEG rows,
// one per wavefront of the block (lane = mode, 64 modes per block), and the
// right-hand sides and pivots of a segment stay in registers, so HBM sees
// one read of the source and one write of the solution per node (the Thomas
// form above also streams c' and d' out and back).  Per segment:
//   1. downward elimination with the left neighbour x_L unknown, giving
//      x_last = Pd + Qd x_R + Rd x_L and the pivots 1/den;
//   2. upward elimination, giving x_first = Pu + Qu x_L + Ru x_R;
//   3. one lane per mode solves the interface system of all segments
//      (x_L of a segment is the last unknown of the one below, x_R the first
//      of the one above) through LDS;
//   4. forward and back substitution with the known x_L and x_R.
// For diagonally dominant rows (every mode but (0,0), which is not solved)
// this is as stable as the plain elimination.  Rows are interior rows
// (a = c = 1/dz^2, b per mode) except the first row (a = 0) and one special
// last row (c = 0); rows past the last one are interior rows with a zero
// right-hand side, which the last row (c = 0) does not see.
// ---------------------------------------------------------------------------
// Pivot reciprocals: the low modes' columns are ill-conditioned and amplify
// pivot rounding, so FP32 refines the hardware reciprocal (about 1 ulp) with
// one Newton step, r (2 - x r), to within rounding of the quotient; FP64
// divides.
template <typename T>
__device__ __forceinline__ T recip(T x) {
  return T(1) / x;
}
template <>
__device__ __forceinline__ float recip<float>(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}

constexpr int kSegMax = 8;  // segments (wavefronts) per block
template <typename T>
struct SegLds {
  T cf[kSegMax][6][64];
  T x[kSegMax][2][64];
  double zm[kSegMax];
};

// Rows of one segment: interior (inv, bg, inv); in segment 0 row 0 is
// (0, b1, c1); row `last` of this segment (when 0 <= last < SEG) is
// (al, bl, 0).
template <typename T>
struct SegRows {
  T inv, bg, b1, c1, al, bl;
  bool first;
  int last;
  template <int I>
  __device__ __forceinline__ void at(int lastv, T& a, T& b, T& c) const {
    a = inv;
    b = bg;
    c = inv;
    if (I == 0 && first) {
      a = T(0);
      b = b1;
      c = c1;
    }
    if (I == lastv) {
      a = (I == 0 && first) ? T(0) : al;
      b = bl;
      c = T(0);
    }
  }
};

// Unrolled loop over compile-time row indices.
template <int I, int N, typename F>
__device__ __forceinline__ void for_rows(F f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>());
    for_rows<I + 1, N>(f);
  }
}
template <int I, typename F>
__device__ __forceinline__ void for_rows_down(F f) {
  if constexpr (I >= 0) {
    f(std::integral_constant<int, I>());
    for_rows_down<I - 1>(f);
  }
}

// Solve rows [0, SEG) of this segment for every lane; d holds the
// right-hand sides on entry and the solution on exit.
template <typename T, int SEG>
__device__ __forceinline__ void seg_solve(SegLds<T>& L, T (&d)[SEG], const SegRows<T>& R, int ns,
                                          int k, int lane) {
  T rden[SEG];
  T Pd, Qd, Rd, Pu, Qu, Ru;
  // the special row's index is re-read per pass so that its compares are
  // not kept across passes
  int last = R.last;
  {
    T cp = T(0), dp = T(0), gp = T(0);
    for_rows<0, SEG>([&](auto I) {
      T a, b, c;
      R.template at<I>(last, a, b, c);
      const T r = recip(b - a * cp);
      rden[I] = r;
      cp = c * r;
      dp = (d[I] - a * dp) * r;
      gp = (I == 0 ? -a : -a * gp) * r;
    });
    Pd = dp;
    Qd = -cp;
    Rd = gp;
  }
  asm volatile("" : "+s"(last));
  {
    T an = T(0), dn = T(0), dl = T(0);
    for_rows_down<SEG - 1>([&](auto I) {
      T a, b, c;
      R.template at<I>(last, a, b, c);
      const T r = recip(b - c * an);
      an = a * r;
      dn = (d[I] - c * dn) * r;
      dl = (I == SEG - 1 ? -c : -c * dl) * r;
    });
    Pu = dn;
    Qu = -an;
    Ru = dl;
  }
  L.cf[k][0][lane] = Pd;
  L.cf[k][1][lane] = Qd;
  L.cf[k][2][lane] = Rd;
  L.cf[k][3][lane] = Pu;
  L.cf[k][4][lane] = Qu;
  L.cf[k][5][lane] = Ru;
  lds_barrier();
  if (k == 0) {
    // x_last of segment j-1 = al + be x_first of segment j; eliminate upward
    T al = T(0), be = T(0);
    for (int j = 0; j < ns; ++j) {
      const T pd = L.cf[j][0][lane], qd = L.cf[j][1][lane], rd = L.cf[j][2][lane];
      const T pu = L.cf[j][3][lane], qu = L.cf[j][4][lane], ru = L.cf[j][5][lane];
      const T den = T(1) - qu * be;
      const T sg = (pu + qu * al) / den, ta = ru / den;
      const T na = pd + rd * al + rd * be * sg, nb = rd * be * ta + qd;
      L.cf[j][0][lane] = sg;
      L.cf[j][1][lane] = ta;
      L.cf[j][2][lane] = na;
      L.cf[j][3][lane] = nb;
      al = na;
      be = nb;
    }
    T xn = T(0);
    for (int j = ns - 1; j >= 0; --j) {
      const T xe = L.cf[j][2][lane] + L.cf[j][3][lane] * xn;
      const T xs = L.cf[j][0][lane] + L.cf[j][1][lane] * xn;
      L.x[j][0][lane] = xs;
      L.x[j][1][lane] = xe;
      xn = xs;
    }
  }
  lds_barrier();
  const T xl = k > 0 ? L.x[k - 1][1][lane] : T(0);
  const T xr = k < ns - 1 ? L.x[k + 1][0][lane] : T(0);
  asm volatile("" : "+s"(last));
  {
    T dp = xl;
    for_rows<0, SEG>([&](auto I) {
      T a, b, c;
      R.template at<I>(last, a, b, c);
      dp = (d[I] - a * dp) * rden[I];
      d[I] = dp;
    });
  }
  asm volatile("" : "+s"(last));
  {
    T x = xr;
    for_rows_down<SEG - 1>([&](auto I) {
      T a, b, c;
      R.template at<I>(last, a, b, c);
      x = d[I] - c * rden[I] * x;
      d[I] = x;
    });
  }
}

// Robin problem of Phi_c (rows r = 0..nz-2 on planes 1..nz-1), the same
// rows as robin_sweep_kernel.
template <typename T, int SEG>
__global__ __launch_bounds__(64 * kSegMax) void robin_seg_kernel(
    const T* __restrict__ src, T* __restrict__ out, const T* __restrict__ lam,
    const T* __restrict__ kk, const T* __restrict__ bh, const T* __restrict__ qh,
    const T* __restrict__ gh, int nz, long plane, T inv_dz2, T two_inv_dz2, T dz, T dz2,
    T two_dz, T rho, T one_m_rho, T kb, T fu_rho, T rho_over_dz, double* zm_partial, double norm,
    bool div_src) {
  __shared__ SegLds<T> L;
  // the segment index is uniform per wavefront: keep it (and the row
  // addresses) scalar
  const int lane = threadIdx.x & 63, ns = blockDim.x >> 6;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // neighbouring mode blocks share the cache lines at their ends (a plane's
  // rows are not line aligned): keep them in one XCD's L2
  const long m = (long)xcd_tile(blockIdx.x, gridDim.x) * 64 + lane;
  const bool live = m < plane;
  const unsigned mo = (unsigned)((live ? m : 0) * (long)sizeof(T));  // byte offset in a plane
  const long pb = plane * (long)sizeof(T);                           // plane bytes
  const int nu = nz - 1;
  const int r0 = k * SEG;
  const int last = nu - 1 - r0;  // top row in this segment when 0 <= last < SEG
  auto row_ptr = [&](const T* base, int plane_index) {
    return reinterpret_cast<const char*>(base) + (long)plane_index * pb;
  };
  // right-hand sides: row r reads plane r+1 (rows past the top re-read the
  // top plane and are zeroed below; the top row itself uses plane nz-2)
  T d[SEG];
  {
    const char* rp = row_ptr(src, r0 + 1);
    const char* lastp = row_ptr(src, nz - 1);
    for_rows<0, SEG>([&](auto I) {
      const char* q = rp < lastp ? rp : lastp;
      d[I] = ld(reinterpret_cast<const T*>(q), mo);
      rp += pb;
    });
  }
  const bool has_top = last >= 0 && last < SEG;
  const T stop = has_top ? ld(reinterpret_cast<const T*>(row_ptr(src, nz - 2)), mo) : T(0);
  // top Robin value g of (d_z + k) u (0 on the NFGS path)
  const T gtop = (has_top && gh) ? two_dz * ld(gh, mo) : T(0);
  if (m == 0) {
    // (0,0) mode: no correction; sum of squares of the normalised source
    double acc = 0.0;
    for_rows<0, SEG>([&](auto I) {
      if (I <= last) {
        const double v = (double)d[I] / norm;
        acc += v * v;
      }
    });
    if (k == 0) {
      const double v = (double)src[0] / norm;
      acc += v * v;
    }
    L.zm[k] = acc;
  }
  const T l = ld(lam, mo);
  const T il = div_src ? T(1) / l : T(1);  // the source is S / lambda, or S itself
  const T b0 = ld(bh, mo);
  SegRows<T> R;
  R.inv = inv_dz2;
  R.bg = l - two_inv_dz2;
  R.first = k == 0;
  R.b1 = rho > T(0) ? one_m_rho * l - two_inv_dz2 : R.bg;
  R.c1 = rho > T(0) ? fu_rho : inv_dz2;
  R.al = (-l) * dz2 - T(2);
  R.bl = T(2) + two_dz * ld(kk, mo);
  R.last = last;
  T q0 = T(0);
  if (k == 0 && rho > T(0)) q0 = ld(qh, mo);
  for_rows<0, SEG>([&](auto I) {
    T v = d[I] * il;
    if (I == 0 && k == 0)
      v = rho > T(0) ? (one_m_rho * v - rho_over_dz * q0) - kb * b0 : v - inv_dz2 * b0;
    if (I == last) v = gtop - (stop * il) * dz2;
    if (I > last) v = T(0);
    d[I] = v;
  });
  seg_solve<T, SEG>(L, d, R, ns, k, lane);
  if (live) {
    char* wp = reinterpret_cast<char*>(out) + (long)(r0 + 1) * pb;
    for_rows<0, SEG>([&](auto I) {
      if (I <= last) st(reinterpret_cast<T*>(wp), mo, m == 0 ? T(0) : d[I]);
      wp += pb;
    });
    if (k == 0) out[m] = m == 0 ? T(0) : b0;
  }
  if (m == 0 && k == 0 && zm_partial) {
    double s = 0.0;
    for (int j = 0; j < ns; ++j) s += L.zm[j];
    *zm_partial = s;
  }
}

// Dirichlet problem of Az on the interior modes (rows i = 0..nz-3 on planes
// 1..nz-2), the same rows as dirichlet_sweep_kernel.
template <typename T, int SEG>
__global__ __launch_bounds__(64 * kSegMax) void dirichlet_seg_kernel(
    const T* __restrict__ src, T* __restrict__ out, const T* __restrict__ lamd,
    const T* __restrict__ bh, int nz, long plane, T inv_dz2, T two_inv_dz2) {
  __shared__ SegLds<T> L;
  const int lane = threadIdx.x & 63, ns = blockDim.x >> 6;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long m = (long)xcd_tile(blockIdx.x, gridDim.x) * 64 + lane;
  const bool live = m < plane;
  const unsigned mo = (unsigned)((live ? m : 0) * (long)sizeof(T));
  const long pb = plane * (long)sizeof(T);
  const int n = nz - 2;
  const int r0 = k * SEG;
  const int last = n - 1 - r0;
  T d[SEG];
  {
    const char* rp = reinterpret_cast<const char*>(src) + (long)(r0 + 1) * pb;
    const char* lastp = reinterpret_cast<const char*>(src) + (long)n * pb;
    for_rows<0, SEG>([&](auto I) {
      const char* q = rp < lastp ? rp : lastp;
      d[I] = ld(reinterpret_cast<const T*>(q), mo);
      rp += pb;
    });
  }
  const T l = ld(lamd, mo);
  const T il = T(1) / l;
  const T u0 = ld(bh, mo);
  SegRows<T> R;
  R.inv = inv_dz2;
  R.bg = l - two_inv_dz2;
  R.first = k == 0;
  R.b1 = R.bg;
  R.c1 = inv_dz2;
  R.al = inv_dz2;
  R.bl = R.bg;
  R.last = last;
  for_rows<0, SEG>([&](auto I) {
    T v = (-d[I]) * il;
    if (I == 0 && k == 0) v = v - inv_dz2 * u0;
    if (I > last) v = T(0);
    d[I] = v;
  });
  seg_solve<T, SEG>(L, d, R, ns, k, lane);
  if (live) {
    char* wp = reinterpret_cast<char*>(out) + (long)(r0 + 1) * pb;
    for_rows<0, SEG>([&](auto I) {
      if (I <= last) st(reinterpret_cast<T*>(wp), mo, d[I]);
      wp += pb;
    });
    if (k == 0) out[m] = u0;
    if (k == ns - 1) out[(long)(nz - 1) * plane + m] = T(0);
  }
}

template <typename T>
int robin_t(const nfgs_plan* p, const void* bh, const void* qh, const void* gh, const void* src,
            void* out, void* scr, double* zm, double rho, bool div_src, hipStream_t s) {
  const nfgs_grid_desc& d = p->d;
  const double dz = d.dz, dz2 = dz * dz;
  const double norm = 4.0 * (d.nx - 1) * (d.ny - 1);
  constexpr int SEG = sizeof(T) == 4 ? 64 : 32;
  const int ns = (d.nz - 1 + SEG - 1) / SEG;
  if (ns <= kSegMax) {
    hipLaunchKernelGGL((robin_seg_kernel<T, SEG>), dim3((unsigned)((p->plane + 63) / 64)),
                       dim3(64 * ns), 0, s, static_cast<const T*>(src), static_cast<T*>(out),
                       static_cast<const T*>(p->lam), static_cast<const T*>(p->kk),
                       static_cast<const T*>(bh), static_cast<const T*>(qh),
                       static_cast<const T*>(gh), d.nz, p->plane,
                       T(1.0 / dz2),
                       T(2.0 / dz2), T(dz), T(dz2), T(2.0 * dz), T(rho), T(1.0 - rho),
                       T((1.0 + 0.5 * rho) / dz2), T((1.0 - 0.5 * rho) / dz2), T(rho / dz), zm,
                       norm, div_src);
    NFGS_LAUNCH_CHECK();
    return NFGS_OK;
  }
  const unsigned nb = (unsigned)((p->plane + 255) / 256);
  hipLaunchKernelGGL((robin_sweep_kernel<T>), dim3(nb), dim3(256), 0, s,
                     static_cast<const T*>(src), static_cast<T*>(out), static_cast<T*>(scr),
                     static_cast<const T*>(p->lam), static_cast<const T*>(p->kk),
                     static_cast<const T*>(bh), static_cast<const T*>(qh),
                       static_cast<const T*>(gh), d.nz, p->plane,
                     T(1.0 / dz2), T(2.0 / dz2), T(dz), T(dz2), T(2.0 * dz), T(rho), T(1.0 - rho),
                     T((1.0 + 0.5 * rho) / dz2), T((1.0 - 0.5 * rho) / dz2), T(rho / dz), zm,
                     norm, div_src);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

template <typename T>
int dirichlet_t(const nfgs_plan* p, const nfgs_boundary* bnd, const void* src, void* out,
                void* scr, hipStream_t s) {
  const double dz2 = p->d.dz * p->d.dz;
  constexpr int SEG = sizeof(T) == 4 ? 64 : 32;
  const int ns = (p->d.nz - 2 + SEG - 1) / SEG;
  if (ns <= kSegMax) {
    hipLaunchKernelGGL((dirichlet_seg_kernel<T, SEG>), dim3((unsigned)((p->iplane + 63) / 64)),
                       dim3(64 * ns), 0, s, static_cast<const T*>(src), static_cast<T*>(out),
                       static_cast<const T*>(p->lamd), static_cast<const T*>(bnd->az_bottom_hat),
                       p->d.nz, p->iplane, T(1.0 / dz2), T(2.0 / dz2));
    NFGS_LAUNCH_CHECK();
    return NFGS_OK;
  }
  const unsigned nb = (unsigned)((p->iplane + 255) / 256);
  hipLaunchKernelGGL((dirichlet_sweep_kernel<T>), dim3(nb), dim3(256), 0, s,
                     static_cast<const T*>(src), static_cast<T*>(out), static_cast<T*>(scr),
                     static_cast<const T*>(p->lamd), static_cast<const T*>(bnd->az_bottom_hat),
                     p->d.nz, p->iplane, T(1.0 / dz2), T(2.0 / dz2));
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace

int robin_modal(const nfgs_plan* p, const void* bh, const void* qh, const void* src, void* out,
                void* scr, double* partials, bool div_src, hipStream_t s) {
  double* zm = partials ? partials + (long)NFGS_Q_ZERO_MODE * p->pblocks : nullptr;
  if (p->d.dtype == NFGS_F32)
    return robin_t<float>(p, bh, qh, nullptr, src, out, scr, zm, p->d.rho, div_src, s);
  return robin_t<double>(p, bh, qh, nullptr, src, out, scr, zm, p->d.rho, div_src, s);
}

}  // namespace nfgs

extern "C" int nfgs_robin_solve(const nfgs_plan* p, const nfgs_boundary* bnd, const void* src_hat,
                                void* out_hat, void* scratch, double* partials, void* stream) {
  NFGS_REQUIRE(p && bnd && src_hat && out_hat && scratch, NFGS_ERR_ARG,
               "robin_solve: null argument");
  NFGS_REQUIRE(bnd->phi_bottom_hat && (p->d.rho == 0.0 || bnd->q_hat), NFGS_ERR_STATE,
               "robin_solve: boundary tables are not set");
  NFGS_REQUIRE(src_hat != out_hat && out_hat != scratch, NFGS_ERR_ARG,
               "robin_solve: buffers must be distinct");
  return nfgs::robin_modal(p, bnd->phi_bottom_hat, bnd->q_hat, src_hat, out_hat, scratch, partials,
                           true, nfgs::as_stream(stream));
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

extern "C" int nfgs_modal_robin_solve(const nfgs_plan* p, const void* src_hat,
                                      const void* bottom_hat, const void* dtarget_hat,
                                      const void* top_hat, double rho, int divide_by_lambda,
                                      void* out_hat, void* scratch, double* zero_mode_sum,
                                      void* stream) {
  NFGS_REQUIRE(p && src_hat && bottom_hat && out_hat && scratch, NFGS_ERR_ARG,
               "modal_robin_solve: null argument");
  NFGS_REQUIRE(rho >= 0.0 && rho <= 1.0, NFGS_ERR_ARG,
               "bottom_robin_weight must be between 0 and 1");
  NFGS_REQUIRE(rho == 0.0 || dtarget_hat, NFGS_ERR_ARG,
               "bottom_derivative_hat is required for mixed bottom Robin");
  NFGS_REQUIRE(src_hat != out_hat && out_hat != scratch, NFGS_ERR_ARG,
               "modal_robin_solve: buffers must be distinct");
  NFGS_REQUIRE(p->d.nz >= 3, NFGS_ERR_ARG, "modal_robin_solve: nz must be >= 3");
  hipStream_t s = nfgs::as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    return nfgs::robin_t<float>(p, bottom_hat, dtarget_hat, top_hat, src_hat, out_hat, scratch,
                                zero_mode_sum, rho, divide_by_lambda != 0, s);
  return nfgs::robin_t<double>(p, bottom_hat, dtarget_hat, top_hat, src_hat, out_hat, scratch,
                               zero_mode_sum, rho, divide_by_lambda != 0, s);
}
