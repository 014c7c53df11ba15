// Coupled horizontal-harmonic closures (horizontal_zero_mode_policy
// "coupled_k0" and "constrained_spline_harmonic", nfgs_solver.py:5058-5398,
// :5952-6067, :6760-6791).  The horizontally uniform field H(z) of the state
// follows, every step, the horizontal mean of the parallel current that the
// PT scalars do not carry: with the constant DCT-I mode of each plane (its
// trapezoid mean) want_c(z) = mean(J_par,c) - mean(J_pt*,c), the profile
// solves D_z H_x = want_y and D_z H_y = -want_x in a least-squares sense
// with H(0) fixed to the bottom target, and H relaxes towards it with
// k0_update_mu.
//
// Every profile solve is linear in (q, h0): H = M q + v h0.  The host builds
// M and v once per grid in FP64 (the closure's normal equations by Cholesky,
// or the spline's saddle-point system by pivoted elimination), the same role
// as the plan's transform tables; the device applies them every step.

#include "util.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nfgs {
namespace {

using Mat = std::vector<double>;  // row-major

// --------------------------------------------------------------------------
// Host linear algebra on small dense matrices.
// --------------------------------------------------------------------------
// Cholesky factor L (lower) of an SPD n x n matrix; false if not SPD.
bool cholesky(const Mat& a, int n, Mat& l) {
  l.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = a[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= l[(size_t)j * n + k] * l[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    l[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double s = a[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= l[(size_t)i * n + k] * l[(size_t)j * n + k];
      l[(size_t)i * n + j] = s / d;
    }
  }
  return true;
}

// x = (L L^T)^{-1} b in place.
void cholesky_solve(const Mat& l, int n, double* b) {
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= l[(size_t)i * n + k] * b[k];
    b[i] = s / l[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= l[(size_t)k * n + i] * b[k];
    b[i] = s / l[(size_t)i * n + i];
  }
}

// Inverse of a general n x n matrix by elimination with partial pivoting.
bool inverse(Mat a, int n, Mat& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (a[(size_t)piv * n + c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double d = a[(size_t)c * n + c];
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c] / d;
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  for (int r = 0; r < n; ++r) {
    const double d = a[(size_t)r * n + r];
    for (int k = 0; k < n; ++k) inv[(size_t)r * n + k] /= d;
  }
  return true;
}

// Eigenvalues of a symmetric n x n matrix by cyclic Jacobi rotations.
std::vector<double> sym_eigenvalues(Mat a, int n) {
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double th = 0.5 * (aqq - app) / apq;
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
      }
  }
  std::vector<double> ev(n);
  for (int i = 0; i < n; ++i) ev[i] = a[(size_t)i * n + i];
  return ev;
}

// Clamped open-uniform cubic B-spline collocation at n_points equispaced
// parameters in [0, 1] (Cox-de Boor recursion; the last point takes the last
// basis function, nfgs_solver.py:5098-5147).
Mat bspline_basis(int np, int nb) {
  const int deg = 3, inner = nb - deg - 1;
  std::vector<double> knot;
  for (int i = 0; i <= deg; ++i) knot.push_back(0.0);
  for (int i = 1; i <= inner; ++i) knot.push_back((double)i / (double)(inner + 1));
  for (int i = 0; i <= deg; ++i) knot.push_back(1.0);
  std::vector<double> t(np);
  for (int i = 0; i < np; ++i) t[i] = np > 1 ? (double)i / (double)(np - 1) : 0.0;
  int width = (int)knot.size() - 1;
  Mat b((size_t)np * width, 0.0);
  for (int i = 0; i < width; ++i)
    for (int r = 0; r < np; ++r) b[(size_t)r * width + i] = (t[r] >= knot[i] && t[r] < knot[i + 1]);
  b[(size_t)(np - 1) * width + width - 1] = 1.0;
  for (int order = 1; order <= deg; ++order) {
    Mat nx((size_t)np * (width - 1), 0.0);
    for (int i = 0; i < width - 1; ++i) {
      const double ld = knot[i + order] - knot[i], rd = knot[i + order + 1] - knot[i + 1];
      for (int r = 0; r < np; ++r) {
        double v = 0.0;
        if (ld > 0.0) v += (t[r] - knot[i]) / ld * b[(size_t)r * width + i];
        if (rd > 0.0) v += (knot[i + order + 1] - t[r]) / rd * b[(size_t)r * width + i + 1];
        nx[(size_t)r * (width - 1) + i] = v;
      }
    }
    b.swap(nx);
    width -= 1;
    for (int i = 0; i < width; ++i) b[(size_t)(np - 1) * width + i] = 0.0;
    b[(size_t)(np - 1) * width + width - 1] = 1.0;
  }
  return b;  // np x nb
}

// M (nz x nz) and v (nz) of H = M q + v h0.
int build_tables(int nz, double dz, int kind, int disc, double smooth, int nb, Mat& M,
                 std::vector<double>& v, double* condition) {
  M.assign((size_t)nz * nz, 0.0);
  v.assign(nz, 0.0);
  const double ss = std::sqrt(smooth);
  // face derivative D ((nz-1) x nz) and true three-point curvature C ((nz-2) x nz)
  const int nd = nz - 1, nc = nz >= 3 ? nz - 2 : 0;
  auto D = [&](int r, int c) { return c == r ? -1.0 / dz : (c == r + 1 ? 1.0 / dz : 0.0); };
  auto Cv = [&](int r, int c) {
    return c == r ? 1.0 / dz : (c == r + 1 ? -2.0 / dz : (c == r + 2 ? 1.0 / dz : 0.0));
  };
  if (kind == NFGS_K0_SPLINE) {
    NFGS_REQUIRE(nb >= 4 && nb <= nz, NFGS_ERR_ARG, "k0 tables: basis size must be in [4, Nz]");
    const Mat basis = bspline_basis(nz, nb);
    const int rows = nd + ((smooth > 0.0 && nc > 0) ? nc : 0);
    Mat design((size_t)rows * nb, 0.0);
    for (int r = 0; r < nd; ++r)
      for (int k = 0; k < nb; ++k) {
        double s = 0.0;
        for (int c = 0; c < nz; ++c) s += D(r, c) * basis[(size_t)c * nb + k];
        design[(size_t)r * nb + k] = s;
      }
    if (rows > nd)
      for (int r = 0; r < nc; ++r)
        for (int k = 0; k < nb; ++k) {
          double s = 0.0;
          for (int c = 0; c < nz; ++c) s += Cv(r, c) * basis[(size_t)c * nb + k];
          design[(size_t)(nd + r) * nb + k] = ss * s;
        }
    Mat normal((size_t)nb * nb, 0.0);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        double s = 0.0;
        for (int r = 0; r < rows; ++r) s += design[(size_t)r * nb + i] * design[(size_t)r * nb + j];
        normal[(size_t)i * nb + j] = s;
      }
    double dmax = 0.0;
    for (int i = 0; i < nb; ++i) dmax = std::fmax(dmax, std::fabs(normal[(size_t)i * nb + i]));
    const double ridge = std::fmax(dmax, 1.0) * 1.0e-12;
    for (int i = 0; i < nb; ++i) normal[(size_t)i * nb + i] += ridge;
    // constraints: H(0) = h0 and the one-sided second-order dH/dz at the top is 0
    std::vector<double> top(nz, 0.0);
    top[nz - 1] = 3.0 / (2.0 * dz);
    top[nz - 2] = -4.0 / (2.0 * dz);
    top[nz - 3] = 1.0 / (2.0 * dz);
    const int nk = nb + 2;
    Mat kkt((size_t)nk * nk, 0.0);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) kkt[(size_t)i * nk + j] = normal[(size_t)i * nb + j];
    for (int k = 0; k < nb; ++k) {
      double c1 = 0.0;
      for (int c = 0; c < nz; ++c) c1 += top[c] * basis[(size_t)c * nb + k];
      const double c0 = basis[k];
      kkt[(size_t)nb * nk + k] = c0;
      kkt[(size_t)(nb + 1) * nk + k] = c1;
      kkt[(size_t)k * nk + nb] = c0;
      kkt[(size_t)k * nk + nb + 1] = c1;
    }
    Mat inv;
    NFGS_REQUIRE(inverse(kkt, nk, inv), NFGS_ERR_ARG, "k0 tables: singular spline system");
    if (condition) {
      const std::vector<double> ev = sym_eigenvalues(kkt, nk);
      double lo = 1e300, hi = 0.0;
      for (double e : ev) {
        lo = std::fmin(lo, std::fabs(e));
        hi = std::fmax(hi, std::fabs(e));
      }
      *condition = hi / lo;
    }
    // rhs of the normal block: design^T [Q q; 0] with Q the face average
    // (Q q)_r = (q_r + q_{r+1}) / 2; the constraint rows carry (h0, 0)
    Mat G((size_t)nb * nz, 0.0);  // design^T [Q; 0]
    for (int k = 0; k < nb; ++k)
      for (int r = 0; r < nd; ++r) {
        const double w = 0.5 * design[(size_t)r * nb + k];
        G[(size_t)k * nz + r] += w;
        G[(size_t)k * nz + r + 1] += w;
      }
    Mat coefM((size_t)nb * nz, 0.0);
    std::vector<double> coefv(nb, 0.0);
    for (int i = 0; i < nb; ++i) {
      for (int j = 0; j < nz; ++j) {
        double s = 0.0;
        for (int k = 0; k < nb; ++k) s += inv[(size_t)i * nk + k] * G[(size_t)k * nz + j];
        coefM[(size_t)i * nz + j] = s;
      }
      coefv[i] = inv[(size_t)i * nk + nb];
    }
    for (int z = 0; z < nz; ++z) {
      for (int j = 0; j < nz; ++j) {
        double s = 0.0;
        for (int k = 0; k < nb; ++k) s += basis[(size_t)z * nb + k] * coefM[(size_t)k * nz + j];
        M[(size_t)z * nz + j] = s;
      }
      double s = 0.0;
      for (int k = 0; k < nb; ++k) s += basis[(size_t)z * nb + k] * coefv[k];
      v[z] = s;
    }
    return NFGS_OK;
  }
  // coupled_k0: least squares over the interior values with H(0) = h0
  const bool legacy = disc == NFGS_K0_LEGACY;
  Mat Dl, Cl;
  if (legacy) {
    Dl.assign((size_t)nz * nz, 0.0);
    for (int r = 1; r < nz - 1; ++r) {
      Dl[(size_t)r * nz + r - 1] = -0.5 / dz;
      Dl[(size_t)r * nz + r + 1] = 0.5 / dz;
    }
    Dl[0] = -1.5 / dz;
    Dl[1] = 2.0 / dz;
    Dl[2] = -0.5 / dz;
    Dl[(size_t)(nz - 1) * nz + nz - 1] = 1.5 / dz;
    Dl[(size_t)(nz - 1) * nz + nz - 2] = -2.0 / dz;
    Dl[(size_t)(nz - 1) * nz + nz - 3] = 0.5 / dz;
    Cl.assign((size_t)nz * nz, 0.0);
    for (int i = 0; i < nz; ++i)
      for (int k = 0; k < nz; ++k) {
        const double a = Dl[(size_t)i * nz + k];
        if (a == 0.0) continue;
        for (int j = 0; j < nz; ++j) Cl[(size_t)i * nz + j] += a * Dl[(size_t)k * nz + j];
      }
  }
  // rows of the operator A over unknowns 1..nz-1, and the columns that carry
  // q (Aq) and h0 (Ah) on the right-hand side
  const int nu = nz - 1;
  std::vector<std::vector<double>> A, Aq;
  std::vector<double> Ah;
  if (legacy) {
    for (int r = 1; r < nz; ++r) {
      std::vector<double> row(nu), rq(nz, 0.0);
      for (int c = 1; c < nz; ++c) row[c - 1] = Dl[(size_t)r * nz + c];
      rq[r] = 1.0;
      A.push_back(row);
      Aq.push_back(rq);
      Ah.push_back(-Dl[(size_t)r * nz]);
    }
    if (smooth > 0.0)
      for (int r = 1; r < nz; ++r) {
        std::vector<double> row(nu), rq(nz, 0.0);
        for (int c = 1; c < nz; ++c) row[c - 1] = ss * Cl[(size_t)r * nz + c];
        A.push_back(row);
        Aq.push_back(rq);
        Ah.push_back(-ss * Cl[(size_t)r * nz]);
      }
  } else {
    for (int r = 0; r < nd; ++r) {
      std::vector<double> row(nu), rq(nz, 0.0);
      for (int c = 1; c < nz; ++c) row[c - 1] = D(r, c);
      rq[r] = 0.5;
      rq[r + 1] = 0.5;
      A.push_back(row);
      Aq.push_back(rq);
      Ah.push_back(-D(r, 0));
    }
    if (nc > 0 && smooth > 0.0)
      for (int r = 0; r < nc; ++r) {
        std::vector<double> row(nu), rq(nz, 0.0);
        for (int c = 1; c < nz; ++c) row[c - 1] = ss * Cv(r, c);
        A.push_back(row);
        Aq.push_back(rq);
        Ah.push_back(-ss * Cv(r, 0));
      }
  }
  const int rows = (int)A.size();
  Mat normal((size_t)nu * nu, 0.0);
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nu; ++j) {
      double s = 0.0;
      for (int r = 0; r < rows; ++r) s += A[r][i] * A[r][j];
      normal[(size_t)i * nu + j] = s;
    }
  Mat l;
  NFGS_REQUIRE(cholesky(normal, nu, l), NFGS_ERR_ARG, "k0 tables: normal matrix not positive");
  // columns of N^{-1} A^T [Aq | Ah]
  std::vector<double> col(nu);
  for (int j = 0; j <= nz; ++j) {
    for (int i = 0; i < nu; ++i) {
      double s = 0.0;
      for (int r = 0; r < rows; ++r) s += A[r][i] * (j < nz ? Aq[r][j] : Ah[r]);
      col[i] = s;
    }
    cholesky_solve(l, nu, col.data());
    for (int i = 0; i < nu; ++i) {
      if (j < nz)
        M[(size_t)(i + 1) * nz + j] = col[i];
      else
        v[i + 1] = col[i];
    }
  }
  v[0] = 1.0;
  if (condition) *condition = 0.0;
  return NFGS_OK;
}

// --------------------------------------------------------------------------
// Device kernels.
// --------------------------------------------------------------------------
// Second z derivative at plane z (nfgs_solver.py:1717-1727).
template <typename T, typename G>
__device__ __forceinline__ T d2z_at(G g, int z, int nz, T dz2) {
  if (z > 0 && z < nz - 1) return (g(1) - T(2) * g(0) + g(-1)) / dz2;
  if (z == 0) return nz >= 4 ? (T(2) * g(0) - T(5) * g(1) + T(4) * g(2) - g(3)) / dz2
                             : (g(1) - g(0)) / dz2;
  return (T(-2) * g(0) + T(2) * g(-1)) / dz2;
}

// Trapezoid means per plane (the constant DCT-I mode) of J_par x, y and of
// the PT current x, y of (phi, az) (nfgs_solver.py:3311-3341 with the
// NFGS boundary values: zero on the side walls and the top).  One block per
// plane, FP64 sums in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k0_means_kernel(const T* __restrict__ jp,
                                                       const T* __restrict__ phi,
                                                       const T* __restrict__ az, int nx, int ny,
                                                       int nz, T dx, T dy, T dz,
                                                       double* __restrict__ means) {
  __shared__ double red[4][256];
  const int z = blockIdx.x;
  const long pl = (long)nx * ny, N = pl * nz;
  const T dz2 = dz * dz, idx2 = T(1) / (dx * dx), idy2 = T(1) / (dy * dy);
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (long r = threadIdx.x; r < pl; r += 256) {
    const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
    const double w = ((x == 0 || x == nx - 1) ? 1.0 : 2.0) * ((y == 0 || y == ny - 1) ? 1.0 : 2.0);
    const long i = (long)z * pl + r;
    s[0] += w * (double)jp[i];
    s[1] += w * (double)jp[N + i];
    if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z < nz - 1) {
      auto ddz = [&](const T* f, long c) {
        return fd1g<T>([&](int o) { return f[(long)(z + o) * pl + c]; }, z, nz, dz);
      };
      auto lphi = [&](long c) {
        const int yy = (int)(c / nx), xx = (int)(c - (long)yy * nx);
        const T bcz = lap_xy_at(phi + (long)z * pl, yy, xx, ny, nx, (long)nx, idx2, idy2);
        return bcz + d2z_at<T>([&](int o) { return phi[(long)(z + o) * pl + c]; }, z, nz, dz2);
      };
      const T dx_dzaz = fd1g<T>([&](int o) { return ddz(az, r + o); }, x, nx, dx);
      const T dy_dzaz = fd1g<T>([&](int o) { return ddz(az, r + (long)o * nx); }, y, ny, dy);
      const T dy_l = fd1g<T>([&](int o) { return lphi(r + (long)o * nx); }, y, ny, dy);
      const T dx_l = fd1g<T>([&](int o) { return lphi(r + o); }, x, nx, dx);
      s[2] += w * (double)(dx_dzaz + dy_l);
      s[3] += w * (double)(dy_dzaz - dx_l);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double norm = 4.0 * (double)(nx - 1) * (double)(ny - 1);
#pragma unroll
    for (int c = 0; c < 4; ++c) means[(long)c * nz + z] = red[c][0] / norm;
  }
}

// Centre derivative of a profile (nfgs_solver.py:5400-5431): the staggered
// form averages the face differences, the legacy form applies the centred
// matrix; one-sided second-order ends in both.
template <typename T>
__device__ __forceinline__ T centre_derivative(const T* p, int z, int nz, T dz, bool legacy) {
  if (z == 0) return legacy ? T(-1.5) / dz * p[0] + T(2) / dz * p[1] + T(-0.5) / dz * p[2]
                            : (T(-3) * p[0] + T(4) * p[1] - p[2]) / (T(2) * dz);
  if (z == nz - 1)
    return legacy ? T(0.5) / dz * p[nz - 3] + T(-2) / dz * p[nz - 2] + T(1.5) / dz * p[nz - 1]
                  : (T(3) * p[nz - 1] - T(4) * p[nz - 2] + p[nz - 3]) / (T(2) * dz);
  if (legacy) return T(-0.5) / dz * p[z - 1] + T(0.5) / dz * p[z + 1];
  return T(0.5) * ((p[z] - p[z - 1]) / dz + (p[z + 1] - p[z]) / dz);
}

// One workgroup: H* = M q + v h0 per component, the relaxed profile into
// h_new, and the diagnostics (want, H*, relative ODE residual).
template <typename T>
__global__ __launch_bounds__(256) void k0_profile_kernel(
    const double* __restrict__ means, const double* __restrict__ tables, int nz, T dz,
    bool legacy, const T* __restrict__ h_bottom, const T* __restrict__ h_old, double mu,
    T* __restrict__ h_new, double* __restrict__ diag, double* __restrict__ residual) {
  extern __shared__ double sm[];
  double* q = sm;                  // 2 nz: q_x = want_y, q_y = -want_x
  T* hs = reinterpret_cast<T*>(sm + 2 * nz);  // 2 nz, the solver dtype
  __shared__ double red[2][256];
  const double* M = tables;
  const double* v = tables + (long)nz * nz;
  for (int z = threadIdx.x; z < nz; z += 256) {
    const double wx = means[z] - means[2L * nz + z];
    const double wy = means[nz + z] - means[3L * nz + z];
    q[z] = wy;
    q[nz + z] = -wx;
    if (diag) {
      diag[z] = wx;
      diag[nz + z] = wy;
    }
  }
  __syncthreads();
  const double h0[2] = {(double)h_bottom[0], (double)h_bottom[1]};
  for (int z = threadIdx.x; z < nz; z += 256) {
    // row z of M is column z of the stored (column-major) table
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < nz; ++j) {
      const double m = M[(long)j * nz + z];
      sx += m * q[j];
      sy += m * q[nz + j];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      double s = c == 0 ? sx : sy;
      s += v[z] * h0[c];
      const T ht = (T)s;
      hs[(long)c * nz + z] = ht;
      if (diag) diag[(2L + c) * nz + z] = (double)ht;
      h_new[(long)c * nz + z] =
          mu < 1.0 ? (T)((T)mu * ht + (T)(1.0 - mu) * h_old[(long)c * nz + z]) : ht;
    }
  }
  __syncthreads();
  if (!residual) return;
  // the current of H*, (-dH_y/dz, dH_x/dz), against want
  double r2 = 0.0, w2 = 0.0;
  for (int z = threadIdx.x; z < nz; z += 256) {
    const double ax = -(double)centre_derivative<T>(hs + nz, z, nz, dz, legacy);
    const double ay = (double)centre_derivative<T>(hs, z, nz, dz, legacy);
    const double wx = -q[nz + z], wy = q[z];
    r2 += (ax - wx) * (ax - wx) + (ay - wy) * (ay - wy);
    w2 += wx * wx + wy * wy;
  }
  red[0][threadIdx.x] = r2;
  red[1][threadIdx.x] = w2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n2 = 2.0 * nz;
    const double w = sqrt(red[1][0] / n2);
    *residual = sqrt(red[0][0] / n2) / (w > 1e-30 ? w : 1e-30);
  }
}

// The spline closure's cell current (nfgs_solver.py:5919-5941) differs from
// the frozen rule's in two places, which the state pass takes as inputs:
// J_z is the cell projection of the face average of the PT J_z (zero on the
// side walls and the top, the toroidal J_z on the bottom), and the side walls
// carry the harmonic current (-dH_y/dz, dH_x/dz) instead of zero.  One thread
// per node forms J_z; the first 2*nz threads of the launch also write the
// wall current.
template <typename T>
__global__ __launch_bounds__(256) void spline_prep_kernel(
    const T* __restrict__ az, const T* __restrict__ jzt, const T* __restrict__ h, int nx, int ny,
    int nz, T idx2, T idy2, T dz, T* __restrict__ jz, T* __restrict__ jh) {
  const long pl = (long)nx * ny;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < 2 * nz) {
    const int z = (int)(i % nz);
    const bool ycomp = i >= nz;
    T v = T(0);
    if (z < nz - 1)
      v = ycomp ? centre_derivative<T>(h, z, nz, dz, false)
                : -centre_derivative<T>(h + nz, z, nz, dz, false);
    jh[i] = v;
  }
  if (i >= pl * nz) return;
  // one node per thread: the PT J_z of its plane and the two beside it
  const int z = (int)(i / pl);
  const long r = i - (long)z * pl;
  const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
  if (z == 0) {
    jz[i] = jzt[r];
    return;
  }
  if (z == nz - 1 || x == 0 || x == nx - 1 || y == 0 || y == ny - 1) {
    jz[i] = T(0);
    return;
  }
  // PT J_z = -lap_xy(Az) on the interior planes, J_z,T on the bottom, 0 on top
  auto pt = [&](int k) -> T {
    if (k == 0) return jzt[r];
    if (k == nz - 1) return T(0);
    const T* c = az + (long)k * pl + r;
    T acc = T(0);
    acc += (c[1] - T(2) * c[0] + c[-1]) * idx2;
    acc += (c[nx] - T(2) * c[0] + c[-nx]) * idy2;
    return -acc;
  };
  const T lo = pt(z - 1), mid = pt(z), up = pt(z + 1);
  jz[i] = T(0.5) * (T(0.5) * (lo + mid) + T(0.5) * (mid + up));
}

// The two fields the PT current of the means is formed from, once per node
// of planes 0..nz-2: lphi = lap_xy(phi) + d2z(phi) (B_c,z + d2 Phi/dz2, whose
// lateral gradient is the poloidal part) and dzaz = dAz/dz, with the
// arithmetic of k0_means_kernel.
template <typename T>
__global__ __launch_bounds__(256) void k0_fields_kernel(const T* __restrict__ phi,
                                                        const T* __restrict__ az, int nx, int ny,
                                                        int nz, T dz, T idx2, T idy2,
                                                        T* __restrict__ lphi,
                                                        T* __restrict__ dzaz) {
  const long pl = (long)nx * ny;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= pl * (nz - 1)) return;
  const int z = (int)(i / pl);
  const long c = i - (long)z * pl;
  const int y = (int)(c / nx), x = (int)(c - (long)y * nx);
  const T dz2 = dz * dz;
  const T bcz = lap_xy_at(phi + (long)z * pl, y, x, ny, nx, (long)nx, idx2, idy2);
  lphi[i] = bcz + d2z_at<T>([&](int o) { return phi[(long)(z + o) * pl + c]; }, z, nz, dz2);
  dzaz[i] = fd1g<T>([&](int o) { return az[(long)(z + o) * pl + c]; }, z, nz, dz);
}

// The trapezoid plane sums of k0_means_kernel from the precomputed fields,
// split over gridDim.x chunks of every plane (blockIdx.y): block partials
// part[((z * chunks) + chunk) * 4 + q].
template <typename T>
__global__ __launch_bounds__(256) void k0_sums_kernel(const T* __restrict__ jp,
                                                      const T* __restrict__ lphi,
                                                      const T* __restrict__ dzaz, int nx, int ny,
                                                      int nz, T dx, T dy,
                                                      double* __restrict__ part) {
  __shared__ double red[4][256];
  const int z = blockIdx.y, chunk = blockIdx.x, chunks = gridDim.x;
  const long pl = (long)nx * ny, N = pl * nz;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (long r = (long)chunk * 256 + threadIdx.x; r < pl; r += 256L * chunks) {
    const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
    const double w = ((x == 0 || x == nx - 1) ? 1.0 : 2.0) * ((y == 0 || y == ny - 1) ? 1.0 : 2.0);
    const long i = (long)z * pl + r;
    s[0] += w * (double)jp[i];
    s[1] += w * (double)jp[N + i];
    if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z < nz - 1) {
      const T dx_dzaz = fd1g<T>([&](int o) { return dzaz[i + o]; }, x, nx, dx);
      const T dy_dzaz = fd1g<T>([&](int o) { return dzaz[i + (long)o * nx]; }, y, ny, dy);
      const T dy_l = fd1g<T>([&](int o) { return lphi[i + (long)o * nx]; }, y, ny, dy);
      const T dx_l = fd1g<T>([&](int o) { return lphi[i + o]; }, x, nx, dx);
      s[2] += w * (double)(dx_dzaz + dy_l);
      s[3] += w * (double)(dy_dzaz - dx_l);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[((long)z * chunks + chunk) * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// means[q * nz + z] = sum over the chunks of plane z, in chunk order.
__global__ void k0_sums_final_kernel(const double* __restrict__ part, int chunks, int nx, int ny,
                                     int nz, double* __restrict__ means) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  const double norm = 4.0 * (double)(nx - 1) * (double)(ny - 1);
  for (int q = 0; q < 4; ++q) {
    double v = 0.0;
    for (int c = 0; c < chunks; ++c) v += part[((long)z * chunks + c) * 4 + q];
    means[(long)q * nz + z] = v / norm;
  }
}

// The means, then the profile.  With the scratch of a step (lphi and dzaz:
// N values each of the plan dtype, part: room for nz * 16 * 4 doubles) the
// means come from the two-pass kernels, else from the one-pass kernel.
template <typename T>
int update_t(const nfgs_plan* p, const nfgs_k0_step* k, const void* jp, const void* phi,
             const void* az, void* lphi, void* dzaz, double* part, int part_doubles,
             hipStream_t s) {
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
  if (lphi && dzaz && part) {
    const long n = p->plane * (nz - 1);
    hipLaunchKernelGGL(k0_fields_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       static_cast<const T*>(phi), static_cast<const T*>(az), nx, ny, nz,
                       (T)p->d.dz, (T)(1.0 / (p->d.dx * p->d.dx)),
                       (T)(1.0 / (p->d.dy * p->d.dy)), static_cast<T*>(lphi),
                       static_cast<T*>(dzaz));
    NFGS_LAUNCH_CHECK();
    int chunks = part_doubles / (4 * nz);
    chunks = chunks > 16 ? 16 : (chunks < 1 ? 1 : chunks);
    hipLaunchKernelGGL(k0_sums_kernel<T>, dim3(chunks, nz), dim3(256), 0, s,
                       static_cast<const T*>(jp), static_cast<const T*>(lphi),
                       static_cast<const T*>(dzaz), nx, ny, nz, (T)p->d.dx, (T)p->d.dy, part);
    NFGS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k0_sums_final_kernel, dim3((nz + 255) / 256), dim3(256), 0, s, part, chunks,
                       nx, ny, nz, k->work);
    NFGS_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(k0_means_kernel<T>, dim3(nz), dim3(256), 0, s, static_cast<const T*>(jp),
                       static_cast<const T*>(phi), static_cast<const T*>(az), nx, ny, nz,
                       (T)p->d.dx, (T)p->d.dy, (T)p->d.dz, k->work);
    NFGS_LAUNCH_CHECK();
  }
  const size_t lds = (size_t)2 * nz * sizeof(double) + (size_t)2 * nz * sizeof(T);
  hipLaunchKernelGGL(k0_profile_kernel<T>, dim3(1), dim3(256), lds, s, k->work, k->tables, nz,
                     (T)p->d.dz, k->discretization == NFGS_K0_LEGACY,
                     static_cast<const T*>(k->h_bottom), static_cast<const T*>(k->h_old), k->mu,
                     static_cast<T*>(k->h_new), k->diag, k->residual);
  NFGS_LAUNCH_CHECK();
  return NFGS_OK;
}

}  // namespace

int k0_update_scratch(const nfgs_plan* p, const nfgs_k0_step* k, const void* jp, const void* phi,
                      const void* az, void* lphi, void* dzaz, void* part, hipStream_t s) {
  NFGS_REQUIRE(p && k && jp && phi && az && k->tables && k->h_bottom && k->h_old && k->h_new &&
                   k->work,
               NFGS_ERR_ARG, "k0 update: null argument");
  NFGS_REQUIRE(p->d.nz >= 4 && p->d.nz <= 1024, NFGS_ERR_ARG,
               "k0 update: Nz must be in [4, 1024]");
  NFGS_REQUIRE(k->mu > 0.0 && k->mu <= 1.0, NFGS_ERR_ARG, "k0 update: mu must be in (0, 1]");
  // part holds the N plan-dtype values of a workspace volume
  const long cap = p->N * p->esize / (long)sizeof(double);
  const int pd = (int)(cap > (1L << 30) ? (1L << 30) : cap);
  if (p->d.dtype == NFGS_F32)
    return update_t<float>(p, k, jp, phi, az, lphi, dzaz, static_cast<double*>(part), pd, s);
  return update_t<double>(p, k, jp, phi, az, lphi, dzaz, static_cast<double*>(part), pd, s);
}

// The state pass of a closure's state: the spline closure's J_z and wall
// current go to the pass through jz_work (N values) and jh_work (2*nz).
int closure_state_pass(const nfgs_plan* p, const nfgs_boundary* bnd, int kind, const void* phi,
                       const void* az, void* b, void* jp, void* jout, double* partials,
                       void* jz_work, void* jh_work, hipStream_t s) {
  if (kind != NFGS_K0_SPLINE)
    return state_pass_closure(p, bnd, phi, az, b, jp, jout, partials, nullptr, nullptr, s);
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
  const long n = std::max<long>(p->N, 2L * nz);
  const dim3 grid((unsigned)((n + 255) / 256));
  if (p->d.dtype == NFGS_F32)
    hipLaunchKernelGGL(spline_prep_kernel<float>, grid, dim3(256), 0, s,
                       static_cast<const float*>(az), static_cast<const float*>(bnd->jz_t),
                       static_cast<const float*>(bnd->h_profile), nx, ny, nz,
                       (float)(1.0 / (p->d.dx * p->d.dx)), (float)(1.0 / (p->d.dy * p->d.dy)),
                       (float)p->d.dz, static_cast<float*>(jz_work),
                       static_cast<float*>(jh_work));
  else
    hipLaunchKernelGGL(spline_prep_kernel<double>, grid, dim3(256), 0, s,
                       static_cast<const double*>(az), static_cast<const double*>(bnd->jz_t),
                       static_cast<const double*>(bnd->h_profile), nx, ny, nz,
                       1.0 / (p->d.dx * p->d.dx), 1.0 / (p->d.dy * p->d.dy), p->d.dz,
                       static_cast<double*>(jz_work), static_cast<double*>(jh_work));
  NFGS_LAUNCH_CHECK();
  return state_pass_closure(p, bnd, phi, az, b, jp, jout, partials, jz_work, jh_work, s);
}

}  // namespace nfgs

using namespace nfgs;

extern "C" {

long nfgs_k0_table_count(int nz) { return nz > 0 ? (long)nz * nz + nz : 0; }

int nfgs_k0_tables(int nz, double dz, int kind, int discretization, double smoothness,
                   int basis_size, double* tables, double* condition) {
  NFGS_REQUIRE(tables && nz >= 4 && nz <= 4096 && dz > 0.0 && smoothness >= 0.0, NFGS_ERR_ARG,
               "k0 tables: need a device table, 4 <= Nz <= 4096, dz > 0, smoothness >= 0");
  NFGS_REQUIRE(kind == NFGS_K0_COUPLED || kind == NFGS_K0_SPLINE, NFGS_ERR_ARG,
               "k0 tables: kind must be NFGS_K0_COUPLED or NFGS_K0_SPLINE");
  NFGS_REQUIRE(discretization == NFGS_K0_STAGGERED || discretization == NFGS_K0_LEGACY,
               NFGS_ERR_ARG, "k0 tables: unknown discretization");
  Mat M;
  std::vector<double> v;
  if (const int rc = build_tables(nz, dz, kind, discretization, smoothness, basis_size, M, v,
                                  condition))
    return rc;
  // the device holds M column by column, so the profile kernel's threads
  // (one output row each) read consecutive doubles
  Mat Mt(M.size());
  for (int i = 0; i < nz; ++i)
    for (int j = 0; j < nz; ++j) Mt[(size_t)j * nz + i] = M[(size_t)i * nz + j];
  NFGS_HIP_CHECK(hipMemcpy(tables, Mt.data(), Mt.size() * sizeof(double), hipMemcpyHostToDevice));
  NFGS_HIP_CHECK(hipMemcpy(tables + M.size(), v.data(), v.size() * sizeof(double),
                           hipMemcpyHostToDevice));
  return NFGS_OK;
}

int nfgs_k0_update(const nfgs_plan* p, const nfgs_k0_step* k, const void* jp, const void* phi,
                   const void* az, void* stream) {
  NFGS_REQUIRE(p && k && jp && phi && az && k->tables && k->h_bottom && k->h_old && k->h_new &&
                   k->work,
               NFGS_ERR_ARG, "k0 update: null argument");
  NFGS_REQUIRE(p->d.nz >= 4 && p->d.nz <= 1024, NFGS_ERR_ARG,
               "k0 update: Nz must be in [4, 1024]");
  NFGS_REQUIRE(k->mu > 0.0 && k->mu <= 1.0, NFGS_ERR_ARG, "k0 update: mu must be in (0, 1]");
  if (p->d.dtype == NFGS_F32)
    return update_t<float>(p, k, jp, phi, az, nullptr, nullptr, nullptr, 0, as_stream(stream));
  return update_t<double>(p, k, jp, phi, az, nullptr, nullptr, nullptr, 0, as_stream(stream));
}

int nfgs_k0_state_pass(const nfgs_plan* p, const nfgs_boundary* bnd, int kind, const void* phi,
                       const void* az, void* b_out, void* jp_out, void* j_out, double* partials,
                       void* work, void* stream) {
  NFGS_REQUIRE(p && bnd && phi && az, NFGS_ERR_ARG, "k0 state pass: null argument");
  NFGS_REQUIRE(kind == NFGS_K0_COUPLED || kind == NFGS_K0_SPLINE, NFGS_ERR_ARG,
               "k0 state pass: kind must be NFGS_K0_COUPLED or NFGS_K0_SPLINE");
  NFGS_REQUIRE(p->d.nz >= 4, NFGS_ERR_ARG, "k0 state pass: Nz must be >= 4");
  NFGS_REQUIRE(bnd->b_bg && bnd->bz_obs && bnd->h_profile && bnd->jz_t, NFGS_ERR_STATE,
               "k0 state pass: boundary is not set");
  NFGS_REQUIRE(kind != NFGS_K0_SPLINE || work, NFGS_ERR_ARG,
               "k0 state pass: the spline closure needs a work buffer");
  void* jh = work ? static_cast<char*>(work) + p->N * p->esize : nullptr;
  return closure_state_pass(p, bnd, kind, phi, az, b_out, jp_out, j_out, partials, work, jh,
                            as_stream(stream));
}

}  // extern "C"
