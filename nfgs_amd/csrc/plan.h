// Internal plan layout shared by the NFGS kernels' launchers.
#pragma once

#include "launch_shape.h"
#include "nfgs_common.h"

constexpr int kRedScratch = 4096;

// Half-size operands of a type-I transform matrix T (n x n).  T has the
// pairing symmetry T[n-1-k][j] = (-1)^j T[k][j], so with E[k][j'] = T[k][2j']
// and O[k][j'] = T[k][2j'+1] the outputs are out[k] = (E x_even + O x_odd)[k]
// and out[n-1-k] = (E x_even - O x_odd)[k] for k < h = ceil(n/2).  E and O are
// stored transposed (row j', column k) and zero-padded to rows_pad x cols_pad.
struct PairedMatrix {
  int n = 0, h = 0, rows = 0, rows_pad = 0, cols_pad = 0;
  void* et[2] = {nullptr, nullptr};  // [direction] E^T
  void* ot[2] = {nullptr, nullptr};  // [direction] O^T
  // Signed base-256 digit form of E and O for the exact-integer FP32 kernel
  // (FP32 plans with n <= 258): K chunks of 64 paired rows, an optional tail
  // row kept in FP64, at most 8 16-row output blocks plus an optional
  // remainder output row 16*hb (h = 16*hb + 1) computed in FP64.  dig is
  // laid out as the MFMA operand registers of each block; entries are
  // round(T * 2^shift).
  int nch = 0, tail = 0, hb = 0, rem = 0;
  int blocks_per_cu = 0;  // launch override of the FP32 kernel (0: as many as fit)
  int shift[2] = {0, 0};
  void* dig[2] = {nullptr, nullptr};    // [direction] int8 [half][digit][chunk][block][lane][16]
  void* mtail[2] = {nullptr, nullptr};  // [direction] double [half][16*hb] tail row
  void* mrem[2] = {nullptr, nullptr};   // [direction] double [n] row 16*hb of T
  // Split-bf16 form of E and O for the BF16x3 kernel (FP32 plans): each
  // entry as hi + lo bf16, laid out as the 16x16x16 MFMA operand of every
  // (16-row K chunk, 16-column output block): bfr[dir][part][chunk][block]
  // [lane] holds 4 bf16 (paired rows 4*(lane/16)..+3 of output lane%16);
  // part 0..3 = E hi, E lo, O hi, O lo.
  int bchunks = 0, bblocks = 0;
  void* bfr[2] = {nullptr, nullptr};
};

struct nfgs_plan {
  nfgs_grid_desc d;
  long N = 0;        // nz*ny*nx
  long plane = 0;    // ny*nx
  long iplane = 0;   // (ny-2)*(nx-2)
  int esize = 4;     // element bytes
  PairedMatrix dct_x, dct_y, dst_x, dst_y;
  void* lam = nullptr;   // ny*nx endpoint-Neumann eigenvalues
  void* kk = nullptr;    // ny*nx sqrt(-lambda)
  void* lamd = nullptr;  // (ny-2)*(nx-2) endpoint-Dirichlet eigenvalues
  void* ax = nullptr;    // nx   2 sin(pi k / (2(nx-1))) / dx
  void* ay = nullptr;    // ny   2 sin(pi k / (2(ny-1))) / dy
  int zchunks = 0;       // z-chunk count of the tiled stencils (0: the cost model)
  int lateral_arith = 0; // NFGS_LATERAL_* arithmetic of the FP32 lateral transforms
  int pblocks = 0;       // blocks of the reducing kernels (partials per slot)
  double* red = nullptr; // kRedScratch doubles of reduction scratch (not re-entrant)
};

namespace nfgs {
const PairedMatrix* find_matrix(const nfgs_plan* p, int kind, int n);
int fill_zero(const nfgs_plan* p, void* ptr, long n, hipStream_t s);
// CUs of the current device (cached on first use)
int cu_count_stencil();
// Workgroups of a tiled stencil launch with the plan's current z-chunking.
long tiled_blocks(const nfgs_plan* p);
int fill_zero_f64(double* ptr, long n, hipStream_t s);
// The split-bf16 x3 lateral transform (FP32 data and accumulation).
int launch_bf16x3(const PairedMatrix* m, int dir, const float* in, long ies, long ivs, long ips,
                  float* out, long oes, long ovs, long ops, int nvec, long total, hipStream_t s);
int launch_transform(const nfgs_plan* p, const PairedMatrix* m, int direction, const void* in,
                     long ies, long ivs, long ips, void* out, long oes, long ovs, long ops,
                     int nvec, int nplanes, hipStream_t s);
// Modal Robin z-solve with explicit bottom tables: (d_zz + lambda) u = S/lambda
// (div_src) or = S, u(0) = bh, the rho-mixed first row with target qh.
int robin_modal(const nfgs_plan* p, const void* bh, const void* qh, const void* src, void* out,
                void* scr, double* partials, bool div_src, hipStream_t s);
// Two-stage Phi_c update helpers: out[0:plane] = lambda * phi_bottom_hat and
// out[plane:2 plane] = lambda * q_hat (0 without q_hat); hat[q][j] /= lambda[j]
// on nplanes planes, the constant mode set to 0.
int two_stage_tables(const nfgs_plan* p, const nfgs_boundary* bnd, void* out, hipStream_t s);
int divide_by_lambda(const nfgs_plan* p, void* hat, int nplanes, hipStream_t s);
// The state pass with a closure's cell J_z (jzc: a volume, boundary values
// included) and side-wall current (jh: 2 x nz, J_x then J_y by height) in
// place of the PT J_z and zero walls; both null is nfgs_state_pass.
int state_pass_closure(const nfgs_plan* p, const nfgs_boundary* bnd, const void* phi,
                       const void* az, void* b, void* jp, void* jout, double* partials,
                       const void* jzc, const void* jh, hipStream_t s);
// The closure update of a step with workspace scratch: lphi and dzaz take N
// values each of the plan dtype, part a workspace volume (as doubles).
int k0_update_scratch(const nfgs_plan* p, const nfgs_k0_step* k, const void* jp, const void* phi,
                      const void* az, void* lphi, void* dzaz, void* part, hipStream_t s);
// The state pass of a closure (kind NFGS_K0_COUPLED or NFGS_K0_SPLINE);
// the spline closure's inputs to state_pass_closure are built in jz_work
// (N values) and jh_work (2*nz values), both of the plan dtype.
int closure_state_pass(const nfgs_plan* p, const nfgs_boundary* bnd, int kind, const void* phi,
                       const void* az, void* b, void* jp, void* jout, double* partials,
                       void* jz_work, void* jh_work, hipStream_t s);
}  // namespace nfgs
