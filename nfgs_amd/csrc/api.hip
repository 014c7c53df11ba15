// Kernels behind the drop-in's per-stage methods that the reference's
// verification harnesses call (they are not on the iteration's path, which
// uses the fused kernels of stencil.hip): the collocated reconstruction of the
// endpoint-Neumann PT subspace and the explicit k_h = 0 harmonic field.

#include "util.h"

namespace nfgs {
namespace {

// Collocated PT state from the scalars (nfgs_solver.py:3311-3341 with
// ccu_pure_current, NFGS's B_pot = B_background; :1711-1715, :1717-1727,
// :1760-1768, :3255-3257, :6069-6078):
//   Bc = (-Dx Dz Phi + Dy Az, -Dy Dz Phi - Dx Az, Bc_z),
//   J  = (Dx Dz Az + Dy L, Dy Dz Az - Dx L, -lap_xy Az),  L = Bc_z + D2z Phi,
//   B  = B_bg + Bc;  J = 0 on the side walls and the top, B_z(0) = Bz_obs,
//   J_z(0) = Jz_T.
// The nested derivatives are evaluated per node from the scalars, in the
// reference's order (the inner z derivative first).
template <typename T>
int reconstruct_pt_t(const nfgs_plan* p, const nfgs_boundary* bnd, const T* phi, const T* az,
                     const T* bcz, T* b, T* j, T* bc, hipStream_t s) {
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
  const long pl = p->plane, N = p->N;
  const T dx = T(p->d.dx), dy = T(p->d.dy), dz = T(p->d.dz);
  const T dz2 = dz * dz, idx2 = T(1) / (dx * dx), idy2 = T(1) / (dy * dy);
  const T* bg = static_cast<const T*>(bnd->b_bg);
  const T* bzo = static_cast<const T*>(bnd->bz_obs);
  const T* jzt = static_cast<const T*>(bnd->jz_t);
  return launch_map(N, [=] __device__(long i) {
    const int z = (int)(i / pl);
    const long r = i - (long)z * pl;
    const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
    const long col = r;  // (y, x) offset in a plane
    // d/dz of a volume at (z, yy, xx)
    auto ddz = [&](const T* f, long c) {
      return fd1g<T>([&](int o) { return f[(long)(z + o) * pl + c]; }, z, nz, dz);
    };
    // second z derivative (nfgs_solver.py:1717-1727)
    auto d2z = [&](const T* f, long c) -> T {
      auto g = [&](int o) { return f[(long)(z + o) * pl + c]; };
      if (z > 0 && z < nz - 1) return (g(1) - T(2) * g(0) + g(-1)) / dz2;
      if (z == 0) return nz >= 4 ? (T(2) * g(0) - T(5) * g(1) + T(4) * g(2) - g(3)) / dz2
                                 : (g(1) - g(0)) / dz2;
      return (T(-2) * g(0) + T(2) * g(-1)) / dz2;
    };
    const long c0 = col;
    // horizontal derivatives of z-derivative fields
    const T dx_dzphi = fd1g<T>([&](int o) { return ddz(phi, c0 + o); }, x, nx, dx);
    const T dy_dzphi = fd1g<T>([&](int o) { return ddz(phi, c0 + (long)o * nx); }, y, ny, dy);
    const T dx_az = fd1g<T>([&](int o) { return az[i + o]; }, x, nx, dx);
    const T dy_az = fd1g<T>([&](int o) { return az[i + (long)o * nx]; }, y, ny, dy);
    const T bcx = -dx_dzphi + dy_az;
    const T bcy = -dy_dzphi - dx_az;
    const T bcz0 = bcz[i];
    auto lphi = [&](long c) { return bcz[(long)z * pl + c] + d2z(phi, c); };
    const T dx_dzaz = fd1g<T>([&](int o) { return ddz(az, c0 + o); }, x, nx, dx);
    const T dy_dzaz = fd1g<T>([&](int o) { return ddz(az, c0 + (long)o * nx); }, y, ny, dy);
    const T dy_l = fd1g<T>([&](int o) { return lphi(c0 + (long)o * nx); }, y, ny, dy);
    const T dx_l = fd1g<T>([&](int o) { return lphi(c0 + o); }, x, nx, dx);
    T jx = dx_dzaz + dy_l;
    T jy = dy_dzaz - dx_l;
    T jz = -lap_xy_at(az + (long)z * pl, y, x, ny, nx, (long)nx, idx2, idy2);
    T bz = bg[2 * N + i] + bcz0;
    const bool wall = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == nz - 1;
    if (wall) jx = jy = jz = T(0);
    if (z == 0) {
      bz = bzo[col];
      jz = jzt[col];
    }
    if (bc) {
      bc[i] = bcx;
      bc[N + i] = bcy;
      bc[2 * N + i] = bcz0;
    }
    if (b) {
      b[i] = bg[i] + bcx;
      b[N + i] = bg[N + i] + bcy;
      b[2 * N + i] = bz;
    }
    if (j) {
      j[i] = jx;
      j[N + i] = jy;
      j[2 * N + i] = jz;
    }
  }, s);
}

// k_h = 0 field and current of a profile H (2 x nz): b = (H_x, H_y, 0) and
// j = (-dH_y/dz, dH_x/dz, 0) at the cell centres, with the derivative of the
// staggered-face discretization (face differences averaged to the centres,
// second-order one-sided at the ends; nfgs_solver.py:5400-5455).
template <typename T>
int k0_field_t(const nfgs_plan* p, const T* h, T* b, T* j, hipStream_t s) {
  const int nz = p->d.nz;
  const long pl = p->plane, N = p->N;
  const T dz = T(p->d.dz);
  return launch_map(N, [=] __device__(long i) {
    const int z = (int)(i / pl);
    T d[2];
    for (int c = 0; c < 2; ++c) {
      const T* f = h + (long)c * nz;
      if (nz < 2) {
        d[c] = T(0);
      } else if (nz == 2) {
        d[c] = (f[1] - f[0]) / dz;
      } else if (z == 0) {
        d[c] = (T(-3) * f[0] + T(4) * f[1] - f[2]) / (T(2) * dz);
      } else if (z == nz - 1) {
        d[c] = (T(3) * f[nz - 1] - T(4) * f[nz - 2] + f[nz - 3]) / (T(2) * dz);
      } else {
        d[c] = T(0.5) * ((f[z] - f[z - 1]) / dz + (f[z + 1] - f[z]) / dz);
      }
    }
    if (b) {
      b[i] = h[z];
      b[N + i] = h[nz + z];
      b[2 * N + i] = T(0);
    }
    if (j) {
      j[i] = -d[1];
      j[N + i] = d[0];
      j[2 * N + i] = T(0);
    }
  }, s);
}

}  // namespace
}  // namespace nfgs

using namespace nfgs;

extern "C" int nfgs_reconstruct_pt(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi,
                                   const void* az, const void* bcz, void* b_out, void* j_out,
                                   void* bc_out, void* stream) {
  NFGS_REQUIRE(p && bnd && phi && az && bcz && (b_out || j_out || bc_out), NFGS_ERR_ARG,
               "reconstruct_pt: null argument");
  NFGS_REQUIRE(p->d.nx >= 3 && p->d.ny >= 3 && p->d.nz >= 3, NFGS_ERR_ARG,
               "reconstruct_pt: every axis needs at least 3 nodes");
  NFGS_REQUIRE(bnd->b_bg && bnd->bz_obs && bnd->jz_t, NFGS_ERR_STATE,
               "reconstruct_pt: boundary is not set");
  if (p->d.dtype == NFGS_F32)
    return reconstruct_pt_t<float>(p, bnd, static_cast<const float*>(phi),
                                   static_cast<const float*>(az), static_cast<const float*>(bcz),
                                   static_cast<float*>(b_out), static_cast<float*>(j_out),
                                   static_cast<float*>(bc_out), as_stream(stream));
  return reconstruct_pt_t<double>(p, bnd, static_cast<const double*>(phi),
                                  static_cast<const double*>(az), static_cast<const double*>(bcz),
                                  static_cast<double*>(b_out), static_cast<double*>(j_out),
                                  static_cast<double*>(bc_out), as_stream(stream));
}

extern "C" int nfgs_k0_field(const nfgs_plan* p, const void* profile, void* b_out, void* j_out,
                             void* stream) {
  NFGS_REQUIRE(p && profile && (b_out || j_out), NFGS_ERR_ARG, "k0_field: null argument");
  if (p->d.dtype == NFGS_F32)
    return k0_field_t<float>(p, static_cast<const float*>(profile), static_cast<float*>(b_out),
                             static_cast<float*>(j_out), as_stream(stream));
  return k0_field_t<double>(p, static_cast<const double*>(profile), static_cast<double*>(b_out),
                            static_cast<double*>(j_out), as_stream(stream));
}
