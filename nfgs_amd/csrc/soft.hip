// The soft_cfcf start (init_mode="soft_cfcf", nfgs_solver.py:6152-6611).
//
// Every horizontal mode of Phi_c (DCT-I, wavenumber k from the plan's k
// table) and of Az (interior DST-I, k = sqrt(-lambda_D)) gets the z profile
// that minimises the modal form of int |curl J_c|^2 + eps int |J_c|^2 with
// L_k = d_zz - k^2 on the interior planes:
//   Phi_c: |d_z L_k p|^2 + (k^2 + eps) |L_k p|^2
//   Az:    |L_k p|^2 + eps (|d_z p|^2 + k^2 |p|^2)
// The bottom value (and for Phi_c the bottom derivative k p_0) and the
// transparent top Robin row are eliminated exactly, so the free values are
// planes 2..N-2 (Phi_c) or 1..N-2 (Az).  The normal equations of one mode are
// a banded SPD system (bandwidth 4 for Phi_c, 2 for Az) solved by a banded
// Cholesky factorisation in FP64.
//
// One thread owns one mode.  The rows are generated and accumulated in the
// reference's order, with FMA contraction off, so the FP64 arithmetic is the
// reference's operation for operation; the band lives in a global scratch
// laid out mode-fastest, so the threads of a wavefront touch consecutive
// doubles at every step.

#include "util.h"

#include <cfloat>
#include <cmath>

namespace nfgs {
namespace {

constexpr int kWin = 8;  // coefficient window of one affine row

// An affine expression of the free values: sum_j c[j - lo] u_j + k.
struct Aff {
  int lo;
  double c[kWin];
  double k;
};

__device__ __forceinline__ void aff_clear(Aff& a, int lo) {
  a.lo = lo;
#pragma unroll
  for (int i = 0; i < kWin; ++i) a.c[i] = 0.0;
  a.k = 0.0;
}

// dst += s * src (coefficients by index, then the constant), the order of the
// reference's affine sums (nfgs_solver.py:6152-6172).
__device__ __forceinline__ void aff_axpy(Aff& dst, double s, const Aff& src) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < kWin; ++i) {
    const int j = src.lo + i - dst.lo;
    if (j >= 0 && j < kWin) dst.c[j] = dst.c[j] + s * src.c[i];
  }
  dst.k = dst.k + s * src.k;
}

struct SoftMode {
  int kind;  // 0 Phi_c, 1 Az
  int nz;
  double dz, k, b;
};

// The profile value of plane z as an affine expression over the window lo
// (nfgs_solver.py:6240-6275).
__device__ void plane_expr(const SoftMode& m, int z, int lo, Aff& out) {
#pragma clang fp contract(off)
  aff_clear(out, lo);
  const int N = m.nz;
  const int first = m.kind == 0 ? 2 : 1;
  auto put = [&](int j, double v) {
    const int i = j - lo;
    if (i >= 0 && i < kWin) out.c[i] = out.c[i] + v;
  };
  if (z == 0) {
    out.k = m.b;
  } else if (m.kind == 0 && z == 1) {
    put(0, 0.25);
    out.k = (0.75 + (0.5 * m.dz) * m.k) * m.b;
  } else if (z < N - 1) {
    put(z - first, 1.0);
  } else {
    // (3 u_N - 4 u_{N-1} + u_{N-2}) / (2 dz) + k u_N = 0
    const double den = 3.0 + (2.0 * m.dz) * m.k;
    const double a = 4.0 / den, c = -1.0 / den;
    put(N - 2 - first, a * 1.0);
    if (N - 3 >= first) put(N - 3 - first, c * 1.0);
    out.k = (0.0 + a * 0.0) + c * 0.0;
  }
}

// L_k p at interior plane z: (p_{z-1} - 2 p_z + p_{z+1}) / dz^2 - k^2 p_z
// (nfgs_solver.py:6203-6238, :6331-6337).
__device__ void lap_expr(const SoftMode& m, int z, int lo, Aff& out) {
#pragma clang fp contract(off)
  const double inv = 1.0 / (m.dz * m.dz), m2 = -2.0 * inv;
  Aff p;
  Aff d2;
  aff_clear(d2, lo);
  plane_expr(m, z - 1, lo, p);
  aff_axpy(d2, inv, p);
  plane_expr(m, z, lo, p);
  aff_axpy(d2, m2, p);
  Aff pz = p;
  plane_expr(m, z + 1, lo, p);
  aff_axpy(d2, inv, p);
  aff_clear(out, lo);
  aff_axpy(out, 1.0, d2);
  aff_axpy(out, -(m.k * m.k), pz);
}

// Accumulate one weighted row into the band (lower[d][i] = H[i][i-d]) and
// the gradient (nfgs_solver.py:6371-6381).
__device__ __forceinline__ void add_row(const Aff& r, double w, int n, int bw, double* band,
                                        double* grad, long M, long m) {
#pragma clang fp contract(off)
  for (int a = 0; a < kWin; ++a) {
    const int i = r.lo + a;
    if (i < 0 || i >= n) continue;
    const double wi = w * r.c[a];
    grad[(long)i * M + m] += wi * r.k;
    for (int b = 0; b <= a; ++b) {
      const int j = r.lo + b;
      if (j < 0 || i - j > bw) continue;
      band[((long)(i - j) * n + i) * M + m] += wi * r.c[b];
    }
  }
}

// One thread per mode: assemble, factor, solve and write the profile
// (nfgs_solver.py:6274-6402).  kt holds k (kind 0) or lambda_D (kind 1) in
// the plan dtype; modes with k <= 0 get a zero profile.  part[m] and
// part[M + m] receive the mode's curl and current energies (the modal
// quadrature of :6486-6540, weighted by the DCT-I end factors for Phi_c);
// part[2M + m] is 1 where the solve was not finite.
template <typename T>
__global__ __launch_bounds__(256) void soft_modes_kernel(
    int kind, const T* __restrict__ kt, const T* __restrict__ bottom, long M, int nxm, int nym,
    int nz, double dz, double eps, double ridge, double* __restrict__ band,
    double* __restrict__ grad, T* __restrict__ out, long out_plane, double* __restrict__ part) {
#pragma clang fp contract(off)
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int N = nz;
  const int first = kind == 0 ? 2 : 1;
  const int n = N - 1 - first;
  const int bw = kind == 0 ? 4 : 2;
  SoftMode md;
  md.kind = kind;
  md.nz = N;
  md.dz = dz;
  if (kind == 0) {
    md.k = (double)kt[m];
  } else {
    const T neg = -kt[m];
    md.k = (double)sqrt(neg > T(0) ? neg : T(0));
  }
  md.b = (double)bottom[m];
  part[m] = 0.0;
  part[M + m] = 0.0;
  part[2 * M + m] = 0.0;
  if (!(md.k > 0.0)) {
    for (int z = 0; z < N; ++z) out[(long)z * out_plane + m] = T(0);
    return;
  }
  for (int d = 0; d <= bw; ++d)
    for (int i = 0; i < n; ++i) band[((long)d * n + i) * M + m] = 0.0;
  for (int i = 0; i < n; ++i) grad[(long)i * M + m] = 0.0;

  const double k2 = md.k * md.k;
  const double inv2 = 1.0 / (2.0 * dz);
  if (kind == 0) {
    // rows of the interior Laplacian sequence L_m = L_k p at plane m + 1:
    // its first derivative (one-sided ends) with weight 1, then L_m itself
    // with weight k^2 + eps
    const int cnt = N - 2;
    for (int q = 0; q < cnt; ++q) {
      const int lo = q - 4 > 0 ? q - 4 : 0;
      Aff d, l;
      aff_clear(d, lo);
      if (q == 0) {
        lap_expr(md, 1, lo, l);
        aff_axpy(d, -3.0 * inv2, l);
        lap_expr(md, 2, lo, l);
        aff_axpy(d, 4.0 * inv2, l);
        lap_expr(md, 3, lo, l);
        aff_axpy(d, -inv2, l);
      } else if (q == cnt - 1) {
        lap_expr(md, cnt, lo, l);
        aff_axpy(d, 3.0 * inv2, l);
        lap_expr(md, cnt - 1, lo, l);
        aff_axpy(d, -4.0 * inv2, l);
        lap_expr(md, cnt - 2, lo, l);
        aff_axpy(d, inv2, l);
      } else {
        lap_expr(md, q, lo, l);
        aff_axpy(d, -inv2, l);
        lap_expr(md, q + 2, lo, l);
        aff_axpy(d, inv2, l);
      }
      add_row(d, 1.0, n, bw, band, grad, M, m);
      lap_expr(md, q + 1, lo, l);
      add_row(l, k2 + eps, n, bw, band, grad, M, m);
    }
  } else {
    for (int z = 1; z < N - 1; ++z) {
      const int lo = z - 3 > 0 ? z - 3 : 0;
      Aff r;
      lap_expr(md, z, lo, r);
      add_row(r, 1.0, n, bw, band, grad, M, m);
      if (eps > 0.0) {
        Aff p;
        aff_clear(r, lo);
        plane_expr(md, z - 1, lo, p);
        aff_axpy(r, -inv2, p);
        plane_expr(md, z + 1, lo, p);
        aff_axpy(r, inv2, p);
        add_row(r, 1.0 * eps, n, bw, band, grad, M, m);
        plane_expr(md, z, lo, p);
        add_row(p, (1.0 * eps) * k2, n, bw, band, grad, M, m);
      }
    }
  }

  auto B = [&](int d, int i) -> double& { return band[((long)d * n + i) * M + m]; };
  // ridge on the diagonal, scaled by its largest entry (at least 1)
  double dmax = 1.0;
  for (int i = 0; i < n; ++i) dmax = fmax(dmax, fabs(B(0, i)));
  if (ridge > 0.0)
    for (int i = 0; i < n; ++i) B(0, i) = B(0, i) + ridge * dmax;
  double smax = 1.0;
  for (int i = 0; i < n; ++i) smax = fmax(smax, fabs(B(0, i)));
  const double floor_v = DBL_EPSILON * smax;
  // banded Cholesky in place: each entry's original value is read just
  // before it is overwritten by the factor
  for (int i = 0; i < n; ++i) {
    double diag = B(0, i);
    for (int o = 1; o <= (bw < i ? bw : i); ++o) {
      const double f = B(o, i);
      diag -= f * f;
    }
    const double piv = sqrt(diag > floor_v ? diag : floor_v);
    B(0, i) = piv;
    const int last = i + bw < n - 1 ? i + bw : n - 1;
    for (int row = i + 1; row <= last; ++row) {
      double v = B(row - i, row);
      int c0 = i - bw > 0 ? i - bw : 0;
      if (row - bw > c0) c0 = row - bw;
      for (int col = c0; col < i; ++col) v -= B(row - col, row) * B(i - col, i);
      B(row - i, row) = v / piv;
    }
  }
  // forward then backward substitution on -gradient, in place
  auto G = [&](int i) -> double& { return grad[(long)i * M + m]; };
  for (int i = 0; i < n; ++i) {
    double v = -G(i);
    for (int o = 1; o <= (bw < i ? bw : i); ++o) v -= B(o, i) * G(i - o);
    G(i) = v / B(0, i);
  }
  bool finite = true;
  for (int i = n - 1; i >= 0; --i) {
    double v = G(i);
    const int last = i + bw < n - 1 ? i + bw : n - 1;
    for (int row = i + 1; row <= last; ++row) v -= B(row - i, row) * G(row);
    const double u = v / B(0, i);
    G(i) = u;
    finite = finite && isfinite(u);
  }
  part[2 * M + m] = finite ? 0.0 : 1.0;

  // the profile: constant plus the free values in the expression's order
  // (the top plane takes its upper free value first), in the plan dtype
  auto prof = [&](int z) -> double {
    if (z == 0) return md.b;
    if (kind == 0 && z == 1) return (0.75 + (0.5 * dz) * md.k) * md.b + 0.25 * G(0);
    if (z < N - 1) return 0.0 + 1.0 * G(z - first);
    const double den = 3.0 + (2.0 * dz) * md.k;
    double v = 0.0 + (4.0 / den * 1.0) * G(N - 2 - first);
    if (N - 3 >= first) v = v + (-1.0 / den * 1.0) * G(N - 3 - first);
    return v;
  };
  for (int z = 0; z < N; ++z) out[(long)z * out_plane + m] = (T)prof(z);

  // modal energies of the stored profile (nfgs_solver.py:6486-6540)
  double cur = 0.0, curl = 0.0;
  auto pv = [&](int z) -> double { return (double)out[(long)z * out_plane + m]; };
  const double dz2 = dz * dz;
  if (kind == 0) {
    const int cnt = N - 2;
    auto qv = [&](int q) { return ((pv(q + 2) - 2.0 * pv(q + 1)) + pv(q)) / dz2 - k2 * pv(q + 1); };
    for (int q = 0; q < cnt; ++q) {
      const double qq = qv(q);
      double dq;
      if (q == 0)
        dq = ((-3.0 * qv(0) + 4.0 * qv(1)) - qv(2)) / (2.0 * dz);
      else if (q == cnt - 1)
        dq = ((3.0 * qv(cnt - 1) - 4.0 * qv(cnt - 2)) + qv(cnt - 3)) / (2.0 * dz);
      else
        dq = (qv(q + 1) - qv(q - 1)) / (2.0 * dz);
      cur += qq * qq;
      curl += dq * dq + k2 * (qq * qq);
    }
    const int x = (int)(m % nxm), y = (int)(m / nxm);
    const double w = (x == 0 || x == nxm - 1 ? 0.5 : 1.0) * (y == 0 || y == nym - 1 ? 0.5 : 1.0);
    part[m] = w * (k2 * curl);
    part[M + m] = w * (k2 * cur);
  } else {
    for (int z = 1; z < N - 1; ++z) {
      const double a = pv(z);
      const double qq = ((pv(z + 1) - 2.0 * a) + pv(z - 1)) / dz2 - k2 * a;
      const double da = (pv(z + 1) - pv(z - 1)) / (2.0 * dz);
      curl += qq * qq;
      cur += da * da + k2 * (a * a);
    }
    part[m] = k2 * curl;
    part[M + m] = k2 * cur;
  }
}

template <typename T>
struct SoftWork {
  double* band;
  double* grad;
  double* part;
  T* hat;
  T* tmp;
};

template <typename T>
SoftWork<T> soft_work(const nfgs_plan* p, void* work) {
  const long M = p->plane;
  const long n = p->d.nz;
  SoftWork<T> w;
  char* c = static_cast<char*>(work);
  w.band = reinterpret_cast<double*>(c);
  c += sizeof(double) * 5 * n * M;
  w.grad = reinterpret_cast<double*>(c);
  c += sizeof(double) * n * M;
  w.part = reinterpret_cast<double*>(c);
  c += sizeof(double) * (3 * M + kRedBlocks + 8);
  w.hat = reinterpret_cast<T*>(c);
  c += sizeof(T) * (p->N > 8 * M ? p->N : 8 * M);
  w.tmp = reinterpret_cast<T*>(c);
  return w;
}

template <typename T>
int modes(const nfgs_plan* p, int kind, const void* kt, const void* bottom, long M, int nxm,
          int nym, double eps, double ridge, const SoftWork<T>& w, T* out, long out_plane,
          double* energies, hipStream_t s) {
  hipLaunchKernelGGL(soft_modes_kernel<T>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s,
                     kind, static_cast<const T*>(kt), static_cast<const T*>(bottom), M, nxm, nym,
                     p->d.nz, p->d.dz, eps, ridge, w.band, w.grad, out, out_plane, w.part);
  NFGS_LAUNCH_CHECK();
  double* part = w.part;
  double* red = w.part + 3 * M;
  for (int q = 0; q < 3; ++q)
    NFGS_TRY(reduce_to(M, [=] __device__(long i) { return part[q * M + i]; }, red,
                       energies + q, 0, s));
  return NFGS_OK;
}

template <typename T>
int start_t(const nfgs_plan* p, const nfgs_boundary* bnd, double eps, double ridge, T* phi,
            T* az, void* work, double* out, hipStream_t s) {
  const SoftWork<T> w = soft_work<T>(p, work);
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
  const int nxi = nx - 2, nyi = ny - 2;
  const long pl = p->plane, ip = p->iplane;
  const Ctx<T> c{p, s, nx, ny, nz, pl};
  // Phi_c: modal profiles of every DCT-I mode, back to space on planes
  // 1..nz-1 (plane 0 is the bottom potential itself)
  NFGS_TRY(modes<T>(p, 0, p->kk, bnd->phi_bottom_hat, pl, nx, ny, eps, ridge, w, w.hat, pl, out,
                    s));
  NFGS_TRY(xy_pair<T>(c, NFGS_INVERSE, w.hat + pl, phi + pl, w.tmp, nz - 1));
  NFGS_HIP_CHECK(hipMemcpyAsync(phi, bnd->phi_bottom, pl * sizeof(T), hipMemcpyDeviceToDevice, s));
  // Az: modal profiles of every interior DST-I mode, back to the interior of
  // planes 1..nz-1; the side walls are zero and plane 0 is Az_bottom
  NFGS_TRY(modes<T>(p, 1, p->lamd, bnd->az_bottom_hat, ip, nxi, nyi, eps, ridge, w, w.hat, ip,
                    out + 3, s));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_INVERSE, w.hat + ip, 1, nxi, ip, w.tmp + ip, 1,
                            nxi, ip, nyi, nz - 1, s));
  NFGS_TRY(fill_zero(p, az + pl, (long)(nz - 1) * pl, s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_INVERSE, w.tmp + ip, nxi, 1, ip, az + pl + nx + 1,
                            nx, 1, pl, nxi, nz - 1, s));
  NFGS_HIP_CHECK(hipMemcpyAsync(az, bnd->az_bottom, pl * sizeof(T), hipMemcpyDeviceToDevice, s));
  return NFGS_OK;
}

// The diagnostics of the start (nfgs_solver.py:6542-6611) as sums of squares
// (the host forms the RMS values and energies):
//  0 (Phi_c[0] - Phi_above)^2       1 (dPhi_c/dz[0] - target)^2
//  2 (Az[0] - Az_bottom)^2          3 (B_z[0] - Bz_obs)^2
//  4 (J_z[0] - J_z,T)^2             5 |B_c,h[0] - expected|^2
//  6 |B_h[0] - B_h,obs|^2           7 top Robin row of Phi_c, in space
//  8 top Robin row of Az, in space  9 |curl J|^2 (FP64 sum)
// 10 |J|^2
// target = idct2(k dct2(Phi_above)); expected = -grad_h target + (dAz_b/dy,
// -dAz_b/dx).
template <typename T>
int diag_t(const nfgs_plan* p, const nfgs_boundary* bnd, const T* phi, const T* az, const T* b,
           const T* j, const T* bxo, const T* byo, void* work, double* out, hipStream_t s) {
  const SoftWork<T> w = soft_work<T>(p, work);
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
  const int nxi = nx - 2, nyi = ny - 2;
  const long pl = p->plane, ip = p->iplane, N = p->N;
  const Ctx<T> c{p, s, nx, ny, nz, pl};
  const T dx = (T)p->d.dx, dy = (T)p->d.dy, dz = (T)p->d.dz;
  const T* kk = static_cast<const T*>(p->kk);
  const T* lamd = static_cast<const T*>(p->lamd);
  const T* pab = static_cast<const T*>(bnd->phi_bottom);
  const T* pbh = static_cast<const T*>(bnd->phi_bottom_hat);
  const T* azb = static_cast<const T*>(bnd->az_bottom);
  const T* bzo = static_cast<const T*>(bnd->bz_obs);
  const T* jzt = static_cast<const T*>(bnd->jz_t);
  T* hat = w.hat;   // 3 planes of modal data, then the combinations
  T* tmp = w.tmp;   // transform scratch
  T* target = hat + 4 * pl;
  T* top_phi = hat + 5 * pl;
  T* top_az = hat + 6 * pl;
  double* red = w.part;
  // target = idct2(k * dct2(Phi_above)); the stored dct2 has its (0,0) mode
  // zeroed, which k = 0 multiplies anyway
  T* kh = hat + 3 * pl;
  NFGS_TRY(launch_map(pl, [=] __device__(long i) { kh[i] = kk[i] * pbh[i]; }, s));
  NFGS_TRY(xy_pair<T>(c, NFGS_INVERSE, kh, target, tmp, 1));
  // top rows: dct2 of the top three planes of Phi_c, the Robin combination,
  // idct2; the same with the interior DST-I of Az
  NFGS_TRY(xy_pair<T>(c, NFGS_FORWARD, phi + (long)(nz - 3) * pl, hat, tmp, 3));
  const T i2dz = T(2) * dz;
  NFGS_TRY(launch_map(pl, [=] __device__(long i) {
    const T h0 = hat[i], h1 = hat[pl + i], h2 = hat[2 * pl + i];
    kh[i] = (T(3) * h2 - T(4) * h1 + h0) / i2dz + kk[i] * h2;
  }, s));
  NFGS_TRY(xy_pair<T>(c, NFGS_INVERSE, kh, top_phi, tmp, 1));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_FORWARD, az + (long)(nz - 3) * pl + nx + 1, 1, nx,
                            pl, tmp, 1, nxi, ip, nyi, 3, s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_FORWARD, tmp, nxi, 1, ip, hat, nxi, 1, ip, nxi, 3,
                            s));
  NFGS_TRY(launch_map(ip, [=] __device__(long i) {
    const T neg = -lamd[i];
    const T ka = sqrt(neg > T(0) ? neg : T(0));
    const T h0 = hat[i], h1 = hat[ip + i], h2 = hat[2 * ip + i];
    kh[i] = (T(3) * h2 - T(4) * h1 + h0) / i2dz + ka * h2;
  }, s));
  NFGS_TRY(launch_transform(p, &p->dst_x, NFGS_INVERSE, kh, 1, nxi, ip, tmp, 1, nxi, ip, nyi, 1,
                            s));
  NFGS_TRY(launch_transform(p, &p->dst_y, NFGS_INVERSE, tmp, nxi, 1, ip, top_az, nxi, 1, ip, nxi,
                            1, s));
  auto sq = [](double v) { return v * v; };
  // bottom rows
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return sq((double)(phi[i] - pab[i])); }, red,
                     out + 0, 0, s));
  auto dphi0 = [=] __device__(long i) -> T {
    return (T(-3) * phi[i] + T(4) * phi[pl + i] - phi[2 * pl + i]) / (T(2) * dz);
  };
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return sq((double)(dphi0(i) - target[i])); },
                     red, out + 1, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return sq((double)(az[i] - azb[i])); }, red,
                     out + 2, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return sq((double)(b[2 * N + i] - bzo[i])); },
                     red, out + 3, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return sq((double)(j[2 * N + i] - jzt[i])); },
                     red, out + 4, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    const int y = (int)(i / nx), x = (int)(i - (long)y * nx);
    auto dp = [&](int o) { return dphi0(i + o); };
    const T ddx = fd1g<T>([&](int o) { return dp(o); }, x, nx, dx);
    const T ddy = fd1g<T>([&](int o) { return dp(o * nx); }, y, ny, dy);
    const T ax_ = fd1g<T>([&](int o) { return az[i + o]; }, x, nx, dx);
    const T ay_ = fd1g<T>([&](int o) { return az[i + (long)o * nx]; }, y, ny, dy);
    const T tx = fd1g<T>([&](int o) { return target[i + o]; }, x, nx, dx);
    const T ty = fd1g<T>([&](int o) { return target[i + (long)o * nx]; }, y, ny, dy);
    const T bx_ = fd1g<T>([&](int o) { return azb[i + o]; }, x, nx, dx);
    const T by_ = fd1g<T>([&](int o) { return azb[i + (long)o * nx]; }, y, ny, dy);
    const T ex = -tx + by_, ey = -ty - bx_;
    const T cx = -ddx + ay_, cy = -ddy - ax_;
    return sq((double)(cx - ex)) + sq((double)(cy - ey));
  }, red, out + 5, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) {
    return sq((double)(b[i] - bxo[i])) + sq((double)(b[N + i] - byo[i]));
  }, red, out + 6, 0, s));
  NFGS_TRY(reduce_to(pl, [=] __device__(long i) { return sq((double)top_phi[i]); }, red, out + 7,
                     0, s));
  NFGS_TRY(reduce_to(ip, [=] __device__(long i) { return sq((double)top_az[i]); }, red, out + 8,
                     0, s));
  // the reconstructed state's |curl J|^2 and |J|^2 (centred derivatives with
  // one-sided ends, in the plan dtype)
  NFGS_TRY(reduce_to(N, [=] __device__(long i) {
    const int z = (int)(i / pl);
    const long r = i - (long)z * pl;
    const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
    const T* jx = j;
    const T* jy = j + N;
    const T* jz = j + 2 * N;
    auto ddx = [&](const T* f) { return fd1g<T>([&](int o) { return f[i + o]; }, x, nx, dx); };
    auto ddy = [&](const T* f) {
      return fd1g<T>([&](int o) { return f[i + (long)o * nx]; }, y, ny, dy);
    };
    auto ddz = [&](const T* f) {
      return fd1g<T>([&](int o) { return f[i + (long)o * pl]; }, z, nz, dz);
    };
    const T c0 = ddy(jz) - ddz(jy), c1 = ddz(jx) - ddx(jz), c2 = ddx(jy) - ddy(jx);
    return sq((double)c0) + sq((double)c1) + sq((double)c2);
  }, red, out + 9, 0, s));
  NFGS_TRY(reduce_to(N, [=] __device__(long i) {
    return sq((double)j[i]) + sq((double)j[N + i]) + sq((double)j[2 * N + i]);
  }, red, out + 10, 0, s));
  return NFGS_OK;
}

}  // namespace
}  // namespace nfgs

using namespace nfgs;

extern "C" {

long nfgs_soft_workspace_bytes(const nfgs_plan* p) {
  if (!p) return 0;
  const long M = p->plane, n = p->d.nz;
  const long hat = p->N > 8 * M ? p->N : 8 * M;
  return (long)sizeof(double) * (6 * n * M + 3 * M + kRedBlocks + 8) + (hat + p->N) * p->esize;
}

int nfgs_soft_start(const nfgs_plan* p, const nfgs_boundary* bnd, double eps, double ridge,
                    void* phi_out, void* az_out, void* work, double* out, void* stream) {
  NFGS_REQUIRE(p && bnd && phi_out && az_out && work && out, NFGS_ERR_ARG,
               "soft start: null argument");
  NFGS_REQUIRE(p->d.nz >= 5 && p->d.nx >= 3 && p->d.ny >= 3, NFGS_ERR_ARG,
               "soft start: needs Nz >= 5 and Nx, Ny >= 3");
  NFGS_REQUIRE(eps >= 0.0 && ridge >= 0.0, NFGS_ERR_ARG,
               "soft start: epsilon and ridge must be non-negative");
  NFGS_REQUIRE(bnd->phi_bottom && bnd->az_bottom && bnd->phi_bottom_hat && bnd->az_bottom_hat,
               NFGS_ERR_STATE, "soft start: boundary is not set");
  hipStream_t s = as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    return start_t<float>(p, bnd, eps, ridge, static_cast<float*>(phi_out),
                          static_cast<float*>(az_out), work, out, s);
  return start_t<double>(p, bnd, eps, ridge, static_cast<double*>(phi_out),
                         static_cast<double*>(az_out), work, out, s);
}

int nfgs_soft_diagnostics(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi,
                          const void* az, const void* b, const void* j, const void* bx_obs,
                          const void* by_obs, void* work, double* out, void* stream) {
  NFGS_REQUIRE(p && bnd && phi && az && b && j && bx_obs && by_obs && work && out, NFGS_ERR_ARG,
               "soft diagnostics: null argument");
  NFGS_REQUIRE(p->d.nz >= 5 && p->d.nx >= 3 && p->d.ny >= 3, NFGS_ERR_ARG,
               "soft diagnostics: needs Nz >= 5 and Nx, Ny >= 3");
  NFGS_REQUIRE(bnd->phi_bottom && bnd->az_bottom && bnd->phi_bottom_hat && bnd->bz_obs &&
                   bnd->jz_t,
               NFGS_ERR_STATE, "soft diagnostics: boundary is not set");
  hipStream_t s = as_stream(stream);
  if (p->d.dtype == NFGS_F32)
    return diag_t<float>(p, bnd, static_cast<const float*>(phi), static_cast<const float*>(az),
                         static_cast<const float*>(b), static_cast<const float*>(j),
                         static_cast<const float*>(bx_obs), static_cast<const float*>(by_obs),
                         work, out, s);
  return diag_t<double>(p, bnd, static_cast<const double*>(phi), static_cast<const double*>(az),
                        static_cast<const double*>(b), static_cast<const double*>(j),
                        static_cast<const double*>(bx_obs), static_cast<const double*>(by_obs),
                        work, out, s);
}

}  // extern "C"
