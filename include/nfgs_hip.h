/*
 * nfgs_hip.h — C-ABI of the MI355X (gfx950) NFGS plain-iterate library.
 *
 * The reference exposes this path only as the Python class NFGSSolverGPU
 * (nfgs_solver.py:3905; public API __all__ at :29).  Each entry point below
 * replaces one stage of that class, cited per function.  All buffers are
 * caller-owned device pointers (plain pointers, element counts, strides in
 * elements), arrays are C-contiguous in (component, z, y, x) order with x
 * fastest, and the element type is the plan's dtype (float or double).
 * Work is enqueued on the given HIP stream (NULL = default stream) and is
 * asynchronous; nothing here allocates device memory after plan creation.
 *
 * Every function returns NFGS_OK (0) or a negative error code; the text of
 * the last error on the calling thread is returned by nfgs_last_error().
 * A plan is not re-entrant; distinct plans may be used from distinct threads.
 */
#ifndef NFGS_HIP_H
#define NFGS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NFGS_OK = 0,
  NFGS_ERR_ARG = -1,   /* invalid argument (maps to ValueError)            */
  NFGS_ERR_HIP = -2,   /* HIP runtime failure                              */
  NFGS_ERR_STATE = -3, /* call-order violation (maps to RuntimeError)      */
  NFGS_ERR_ALLOC = -4  /* plan table allocation failed                     */
};

enum { NFGS_F32 = 0, NFGS_F64 = 1 };
enum { NFGS_DCT1 = 0, NFGS_DST1 = 1 };
enum { NFGS_FORWARD = 0, NFGS_INVERSE = 1 };

/* Metric slots written by nfgs_reduce() (doubles). */
enum {
  NFGS_Q_SIGMA_NUM = 0,  /* sum |J_f x B_f| / max(|B_f|,1e-10) over the face window */
  NFGS_Q_SIGMA_DEN = 1,  /* sum |J_f| over the face window                          */
  NFGS_Q_DIV_ABS = 2,    /* sum |div B| over all nodes                              */
  NFGS_Q_BMAG = 3,       /* sum |B| over all nodes                                  */
  NFGS_Q_B2 = 4,         /* sum |B|^2 over all nodes                                */
  NFGS_Q_SRC2 = 5,       /* sum S_Bcz^2 (source of the Phi_c solve)                 */
  NFGS_Q_ZERO_MODE = 6,  /* sum over heights of (DCT2(S_Bcz)[z,0,0] / norm)^2        */
  NFGS_Q_FACE_J2 = 7,    /* sum of squared face currents (final J pass only)        */
  NFGS_Q_CELL_J2 = 8,    /* sum of squared centre currents (final J pass only)      */
  NFGS_NQ = 9
};

typedef struct nfgs_plan nfgs_plan;

/* Grid and constants of one solve; mirrors NFGSSolverGPU.__init__
 * (nfgs_solver.py:3925-4349, :736-937) after calculation-grid padding. */
typedef struct {
  int nx, ny, nz;                 /* calculation grid (each >= 3)                     */
  double dx, dy, dz;
  int dtype;                      /* NFGS_F32 or NFGS_F64                             */
  double rho;                     /* bottom_divergence_robin_weight in [0,1]          */
  double alpha_max;               /* alpha clamp, alpha_limit_factor/min(dx,dy,dz)    */
  double weak_field_floor_b2;     /* floor of |B_f|^2 in the alpha projection         */
  int win_z0, win_z1;             /* sigma window on vertical faces, half-open        */
  int win_y0, win_y1;             /* (nfgs_solver.py:2656-2676 on the face grid)      */
  int win_x0, win_x1;
} nfgs_grid_desc;

/* Fixed per-boundary device arrays produced by the setup stage
 * (set_bottom_vector_boundary, nfgs_solver.py:4548-4919). */
typedef struct {
  const void* b_bg;           /* 3*N  B_background = B_below + B_open              */
  const void* bz_obs;         /* ny*nx observed Bz on the calculation grid          */
  const void* jz_t;           /* ny*nx toroidal bottom current -lap_xy(Az_bottom)    */
  const void* phi_bottom;     /* ny*nx Phi_above_bottom                              */
  const void* az_bottom;      /* ny*nx Az_bottom                                     */
  const void* phi_bottom_hat; /* ny*nx DCT2(Phi_above_bottom), (0,0) zeroed          */
  const void* q_hat;          /* ny*nx DCT2(bottom dPhi_c/dz target), (0,0) zeroed   */
  const void* az_bottom_hat;  /* (ny-2)*(nx-2) DST2 of the interior of Az_bottom     */
  const void* h_profile;      /* 2*nz frozen linear harmonic H_k0                    */
} nfgs_boundary;

/* Caller-owned work buffers of one solve. */
typedef struct {
  void* jp;         /* 3*N  parallel current J_par of the current state          */
  void* s_bcz;      /* N    source of the Bc_z / Phi_c solve                      */
  void* s_jz;       /* nz*(ny-2)*(nx-2) interior source of the Az solve           */
  void* hat0;       /* N    transform work                                        */
  void* hat1;       /* N    transform work                                        */
  void* cscr;       /* N    Thomas c' scratch                                      */
  double* partials; /* nfgs_partials_count(plan) doubles                          */
  double* metrics;  /* NFGS_NQ doubles                                            */
} nfgs_workspace;

const char* nfgs_last_error(void);
/* SHA-256 (hex) of the sources this library was compiled from; the Python
 * loader refuses a library whose digest differs from the sources beside it.
 * No counterpart in the reference (a build-consistency check). */
const char* nfgs_source_digest(void);

/* Plan: transform matrices (paired DCT-I/DST-I halves), lambda/k tables.
 * Replaces the constructor's table caches (nfgs_solver.py:1083-1118,
 * :4357-4383). */
int nfgs_plan_create(const nfgs_grid_desc* desc, nfgs_plan** out);
int nfgs_plan_destroy(nfgs_plan* plan);
/* Number of doubles the partials buffer must hold. */
long nfgs_partials_count(const nfgs_plan* plan);
/* Device pointer of a plan table in the plan dtype: 0 endpoint-Neumann
 * lambda (ny*nx), 1 k = sqrt(-lambda) (ny*nx), 2 endpoint-Dirichlet lambda
 * ((ny-2)*(nx-2)), 3 alpha_x (nx), 4 alpha_y (ny). */
const void* nfgs_plan_table(const nfgs_plan* plan, int which);
/* Element count of plan table `which` (0 if none). */
long nfgs_plan_table_size(const nfgs_plan* plan, int which);
/* Copy plan table `which` (in the plan dtype) to the caller's device buffer
 * dst.  Table 0 replaces _endpoint_neumann_lam_xy (nfgs_solver.py:1098-1109),
 * table 2 _dirichlet_lam_xy_interior (:1083-1096). */
int nfgs_copy_plan_table(const nfgs_plan* plan, int which, void* dst, void* stream);

/* Launch tuning of a plan (a project autotuner picks the values per grid;
 * the reference has no counterpart).  NFGS_TUNE_ZCHUNKS: number of z-chunks
 * the tiled stencils split the grid into (0: the built-in cost model; a
 * count whose last chunk would hold fewer than two planes is rejected).
 * Chunking changes only the order in which the monitor partials are summed.
 * NFGS_TUNE_TRANSFORM_BLOCKS_PER_CU: workgroups per CU of the FP32 transform
 * (0: as many as fit; results do not depend on it). */
enum { NFGS_TUNE_ZCHUNKS = 0, NFGS_TUNE_TRANSFORM_BLOCKS_PER_CU = 1 };
int nfgs_plan_set_tuning(nfgs_plan* plan, int knob, int value);
/* Current value of a knob (-1 for an unknown knob). */
int nfgs_plan_get_tuning(const nfgs_plan* plan, int knob);

/* Arithmetic of an FP32 plan's lateral transforms.  NFGS_LATERAL_EXACT (the
 * default): exact integer digit products (lengths up to 258) or FP64 products,
 * one FP32 rounding per output.  NFGS_LATERAL_BF16X3: the precision form,
 * every input and matrix entry split into two bf16 terms, three bf16 MFMA
 * products per term, FP32 accumulation (relative error near 2^-16 per
 * product).  FP64 plans take only NFGS_LATERAL_EXACT.  No counterpart in the
 * reference (its transforms are FFTs in the solver dtype). */
enum { NFGS_LATERAL_EXACT = 0, NFGS_LATERAL_BF16X3 = 1 };
int nfgs_plan_set_lateral_arithmetic(nfgs_plan* plan, int arith);
/* The current NFGS_LATERAL_* value (-1 for a null plan). */
int nfgs_plan_get_lateral_arithmetic(const nfgs_plan* plan);

/* Unnormalised type-I transform (forward) or its inverse along one axis of a
 * batch of vectors: vector g = p*nvec + q starts at p*plane_stride +
 * q*vec_stride and element j sits at j*elem_stride.  Either both element
 * strides or both vector strides must be 1.  n must be nx or ny for DCT-I and
 * nx-2 or ny-2 for DST-I.  Replaces _dct1/_idct1/_dst1/_idst1
 * (nfgs_solver.py:1019-1081). */
int nfgs_transform(const nfgs_plan* plan, int kind, int direction, const void* in,
                   long in_elem_stride, long in_vec_stride, long in_plane_stride, void* out,
                   long out_elem_stride, long out_vec_stride, long out_plane_stride, int n,
                   int nvec, int nplanes, void* stream);

/* Modal Robin z-solve of (d_zz + lambda) Phi = S/lambda per endpoint-Neumann
 * mode, rho-mixed first row and eliminated top Robin row.  src_hat is the
 * raw DCT2 of the source (nz*ny*nx); out_hat receives the modal solution
 * (plane 0 = phi_bottom_hat).  Optionally accumulates the zero-mode sum into
 * partials.  Replaces _solve_phi_c_robin/_solve_modal_robin_from_hat
 * (nfgs_solver.py:5627-5672, :5512-5625). */
int nfgs_robin_solve(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* src_hat,
                     void* out_hat, void* scratch, double* partials, void* stream);

/* The modal Robin solve with every input explicit: per endpoint-Neumann
 * mode (d_zz + lambda) u = S (or S / lambda with divide_by_lambda), u(0) =
 * bottom_hat, the first interior row blended with the bottom derivative
 * condition by rho (dtarget_hat, required when rho > 0), and the top Robin
 * row (d_z + k) u = g with g = top_hat (NULL: 0) after eliminating u_{N-2}.
 * src_hat, out_hat and scratch are nz planes of ny*nx; the (0,0) mode of
 * out_hat is 0.  zero_mode_sum (one device double, may be NULL) receives the
 * sum over heights of (src_hat[z,0,0] / (4 (nx-1)(ny-1)))^2.  Replaces
 * _solve_modal_robin_from_hat (nfgs_solver.py:5512-5625) up to its final
 * inverse DCT2. */
int nfgs_modal_robin_solve(const nfgs_plan* plan, const void* src_hat, const void* bottom_hat,
                           const void* dtarget_hat, const void* top_hat, double rho,
                           int divide_by_lambda, void* out_hat, void* scratch,
                           double* zero_mode_sum, void* stream);

/* Dirichlet z-solve (d_zz + lambda_D) Az = -S/lambda_D per interior DST-I
 * mode, Az(0) = az_bottom_hat, Az(top) = 0, on compact (nz,ny-2,nx-2)
 * arrays.  Replaces _solve_az_from_jz_poisson_source's modal part and
 * _thomas_solve_batched_lowmem (nfgs_solver.py:3263-3306, :535-608). */
int nfgs_dirichlet_solve(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* src_hat,
                         void* out_hat, void* scratch, void* stream);

/* B = B_bg + (-Dx Dz Phi + Dy Az, -Dy Dz Phi - Dx Az, lap_xy Phi) + H, with
 * Bz(0) = Bz_obs.  Replaces the B half of _reconstruct_state_from_scalars
 * (nfgs_solver.py:5901-5942, :3311-3341, :1711-1715). */
int nfgs_reconstruct_b(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* phi,
                       const void* az, void* b_out, void* stream);

/* Vertical-face pass over B: face current of the carrier B - B_bg, alpha
 * projection, faces-to-centres parallel current J_par with the bottom row
 * taken from the state current (jp_out, may be NULL); the state current J
 * (j_out, may be NULL); and block partials of sigma, divB and energy
 * (partials, may be NULL).  Jz is -lap_xy(Az) with its boundary values.
 * Replaces compute_parallel_currents (:5811-5839), the staggered J of
 * reconstruction (:5917-5941), compute_sigma (:5882-5895), compute_divB
 * and compute_energy (:2689-2697). */
int nfgs_face_pass(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* b,
                   const void* az, void* jp_out, void* j_out, double* partials, void* stream);

/* nfgs_face_pass with the current taken from a caller's J (3*N) instead of
 * Az: the centre Jz of the face state is J[2] and the bottom row of J_par is
 * J[:, 0], as the reference's iteration_step takes them from its J argument
 * (compute_parallel_currents(J, B) and J_para[:, 0] = J[:, 0],
 * nfgs_solver.py:6690-6711, :5811-5839).  For the state's own current the
 * result equals nfgs_face_pass's.  jp_out or partials may be NULL (not both);
 * the partials give compute_sigma, compute_divB and compute_energy
 * (nfgs_solver.py:5882-5895, :2689-2697) through nfgs_reduce. */
int nfgs_face_pass_j(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* b,
                     const void* j, void* jp_out, double* partials, void* stream);

/* Collocated reconstruction of the endpoint-Neumann PT subspace from the
 * scalars (Phi_c, Az, Bc_z; N each): Bc = (-Dx Dz Phi + Dy Az,
 * -Dy Dz Phi - Dx Az, Bc_z), J = (Dx Dz Az + Dy L, Dy Dz Az - Dx L,
 * -lap_xy Az) with L = Bc_z + D2z Phi, B = B_bg + Bc, and the boundary
 * values (J = 0 on the side walls and the top, B_z(0) = Bz_obs,
 * J_z(0) = Jz_T).  Any of b_out, j_out, bc_out (3*N each) may be NULL.
 * Replaces _reconstruct_pt_state_from_scalars (nfgs_solver.py:5897-5899,
 * :3311-3341). */
int nfgs_reconstruct_pt(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* phi,
                        const void* az, const void* bcz, void* b_out, void* j_out, void* bc_out,
                        void* stream);

/* The explicit k_h = 0 field and current of a profile (2*nz: H_x then H_y):
 * b = (H_x, H_y, 0), j = (-dH_y/dz, dH_x/dz, 0) with the staggered-face
 * centre derivative.  b_out or j_out (3*N) may be NULL.  Replaces
 * _k0_field_and_current and _k0_center_derivative
 * (nfgs_solver.py:5400-5455). */
int nfgs_k0_field(const nfgs_plan* plan, const void* profile, void* b_out, void* j_out,
                  void* stream);

/* Sources S_Bcz = Dy Jpx - Dx Jpy (full grid) and S_Jz = lap_xy Jpz -
 * D2xz Jpx - D2yz Jpy (interior, compact), plus the S_Bcz^2 partial.
 * Replaces nfgs_solver.py:6726-6745. */
int nfgs_sources(const nfgs_plan* plan, const void* jp, void* s_bcz, void* s_jz,
                 double* partials, void* stream);

/* Fused reconstruction + face pass: B of (phi, az) as nfgs_reconstruct_b
 * builds it, consumed plane by plane by the face pass of nfgs_face_pass
 * without a round trip through memory.  Outputs as nfgs_face_pass; b_out
 * (3*N, may be NULL) receives B.  Results equal nfgs_reconstruct_b followed
 * by nfgs_face_pass.  Replaces _reconstruct_state_from_scalars plus
 * _staggered_face_current and the monitors (nfgs_solver.py:5901-5942,
 * :5738-5839, :5882-5895, :2656-2697). */
int nfgs_state_pass(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* phi,
                    const void* az, void* b_out, void* jp_out, void* j_out, double* partials,
                    void* stream);

/* Deterministic final reduction of the partials into NFGS_NQ metrics. */
int nfgs_reduce(const nfgs_plan* plan, const double* partials, double* metrics, void* stream);

/* One full plain iteration at mu = 1 from ws->jp (the parallel current of the
 * current state): sources, Phi_c and Az solves into phi_out/az_out (whose
 * fixed planes must already hold Phi_above_bottom, Az_bottom and zeros),
 * then nfgs_state_pass, which leaves the next J_par in ws->jp, the metrics
 * of the new state in ws->metrics and, when b_out is not NULL, B in b_out.
 * Replaces NFGSSolverGPU.iteration_step (nfgs_solver.py:6647-6935) plus the
 * per-iteration monitor calls of the driver (:3694-3702).  With
 * zero_source != 0 the sources are zero (the gs_cfcf start, :6613-6645). */
int nfgs_advance(const nfgs_plan* plan, const nfgs_boundary* bnd, const nfgs_workspace* ws,
                 void* phi_out, void* az_out, void* b_out, int zero_source, void* stream);

/* nfgs_advance with the poloidal update selected (poloidal_update_mode,
 * nfgs_solver.py:5727-5733): NFGS_POLOIDAL_DIRECT solves Phi_c directly
 * (:5627-5672, what nfgs_advance does); NFGS_POLOIDAL_TWO_STAGE solves B_c,z
 * with the same rows and recovers Phi_c_hat = B_c,z_hat / lambda plane by
 * plane (:5674-5725; four more transform passes).  The two agree up to
 * rounding.  s_bcz also carries the two lambda-scaled bottom planes. */
enum { NFGS_POLOIDAL_DIRECT = 0, NFGS_POLOIDAL_TWO_STAGE = 1 };
int nfgs_advance_ex(const nfgs_plan* plan, const nfgs_boundary* bnd, const nfgs_workspace* ws,
                    void* phi_out, void* az_out, void* b_out, int zero_source, int poloidal_mode,
                    void* stream);

/* --- boundary setup (set_bottom_vector_boundary, once per boundary) --- */

/* Fortran readin+botbal bottom Jz (nfgs_solver.py:1498-1567 legacy helper
 * grid when helper dims exceed the calculation grid; :4519-4546 otherwise).
 * bx, by: calculation-grid planes.  The helper grid is (hny, hnx) with the
 * calculation grid placed at (bfy, bfx).  legacy selects the helper-grid
 * edge zeroing.  ijzbcor in {0..4}. */
int nfgs_setup_jz(const nfgs_plan* plan, const void* bx, const void* by, void* jz, int hny,
                  int hnx, int bfy, int bfx, int ijzbcor, int legacy, void* stream);

/* nfgs_setup_jz without a plan, for callers that hold only the grid
 * (NFGSSolverGPU._derive_bottom_jz_fortran called on a plain namespace, as
 * the reference's tests do: tests/test_bottom_current_padding.py:11-41):
 * dtype NFGS_F32 or NFGS_F64, spacings dx, dy, the rest as there
 * (nfgs_solver.py:4489-4546, :1498-1567). */
int nfgs_bottom_jz(const void* bx, const void* by, void* jz, int ny, int nx, double dx, double dy,
                   int dtype, int hny, int hnx, int bfy, int bfx, int ijzbcor, int legacy,
                   void* stream);

/* Mirror-Neumann 5-point lateral Laplacian of nplanes stacked (ny, nx)
 * planes (every node; ends take the mirrored neighbour).  Replaces _lap_xy
 * (nfgs_solver.py:1741-1749). */
int nfgs_lap_xy(const nfgs_plan* plan, const void* in, void* out, int nplanes, void* stream);

/* Endpoint-Dirichlet 2-D Poisson of scale*src on one plane: interior solve,
 * zero edges (nfgs_solver.py:1132-1144).  work: 2*(ny-2)*(nx-2) elements. */
int nfgs_setup_poisson2d_dirichlet(const nfgs_plan* plan, const void* src, double scale,
                                   void* out, void* work, void* stream);

/* out = scale * lap_xy(in) on one plane, edges zeroed when zero_edges != 0
 * (nfgs_solver.py:3039-3046). */
int nfgs_setup_plane_lap(const nfgs_plan* plan, const void* in, double scale, int zero_edges,
                         void* out, void* stream);

/* Workspace bytes needed by nfgs_setup_gaussian. */
long nfgs_gaussian_workspace_bytes(const nfgs_plan* plan);

/* Gaussian separation with the spectral horizontal operator, uniform-open
 * mean Bz, frozen linear harmonic closure (nfgs_solver.py:4671-4919,
 * :4990-5056, :5069-5096).  Writes B_bg (3*N), bz_above, phi_bottom
 * (Phi_above_bottom), dphi_target, d0, phi_above_hat (planes), h_target
 * (2 values, the constant bottom harmonic residual) and diag
 * (NFGS_GDIAG_COUNT device doubles: DCT constant modes, sums and sums of
 * squares from which the reference's diagnostics follow). */
enum {
  NFGS_GD_BX_MEAN = 0,   /* DCT2(Bx)[0,0] / (4(nx-1)(ny-1))                      */
  NFGS_GD_BY_MEAN,       /* same for By                                          */
  NFGS_GD_BZ_MEAN,       /* same for Bz = Bz_open_mean                            */
  NFGS_GD_D0_MEAN,       /* constant mode of the spectral divergence             */
  NFGS_GD_D0_SQ,         /* sum D0^2                                             */
  NFGS_GD_RECON_SQ,      /* sum (B_bg,z(0) + Bz_above - Bz_obs)^2                 */
  NFGS_GD_RECON_MAX,     /* max |B_bg,z(0) + Bz_above - Bz_obs|                   */
  NFGS_GD_BG_INV_SQ,     /* sum (lap_xy Phi_below + open - B_bg,z)^2              */
  NFGS_GD_HRES_SQ,       /* sum of the squared bottom horizontal residual         */
  NFGS_GD_HOBS_SQ,       /* sum Bx_obs^2 + By_obs^2                               */
  NFGS_GD_HCONST_X,      /* constant residual mode, x (= H_x)                     */
  NFGS_GD_HCONST_Y,      /* constant residual mode, y (= H_y)                     */
  NFGS_GD_HNONCONST_SQ,  /* sum of squared non-constant residual                  */
  NFGS_GD_HDIV_SQ,       /* sum of squared spectral divergence of the residual    */
  NFGS_GD_HCURL_SQ,      /* sum of squared FD curl of the residual                */
  NFGS_GD_BZ_SUM,        /* sum Bz_obs                                            */
  NFGS_GD_BZ_ABS_SUM,    /* sum |Bz_obs|                                          */
  NFGS_GD_BX_SUM,        /* sum Bx_obs                                            */
  NFGS_GD_BY_SUM,        /* sum By_obs                                            */
  NFGS_GDIAG_COUNT
};
int nfgs_setup_gaussian(const nfgs_plan* plan, const void* bx, const void* by, const void* bz,
                        const void* az_bottom, void* b_bg, void* bz_above, void* phi_bottom,
                        void* dphi_target, void* d0, void* phi_above_hat, void* h_target,
                        double* diag, void* workspace, void* stream);

/* ---- initial state from a caller's field (init_mode="field") ----------- */

/* Workspace bytes needed by nfgs_project_field (three volumes). */
long nfgs_project_workspace_bytes(const nfgs_plan* plan);

/* Projects a field into the NFGS scalars, as the cfcf handoff does with the
 * Stable solver's B and Az (nfgs_solver.py:6080-6147; the Neumann solve
 * :1120-1130 and :1146-1180): phi[0] = Phi_above_bottom, phi[1:] solves
 * lap_xy phi = B_z - B_bg,z per plane with every plane's constant mode set to
 * zero (its value per height, :4466-4468, is written to mean_bcz: nz device
 * doubles); az = az_in with Az_bottom on the bottom plane and zero on the top
 * plane and the lateral walls.  b is 3*N, az_in, phi and az are N.  The
 * projected state's B and J come from nfgs_state_pass.  Not in the
 * reference as an entry point: the reference builds its handoff state inside
 * _build_stable_cfcf_initial_state. */
int nfgs_project_field(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* b,
                       const void* az_in, void* phi, void* az, double* mean_bcz, void* workspace,
                       void* stream);

/* ---- the outward Gaussian field on z = 0 -------------------------------- */

/* Workspace bytes needed by nfgs_gaussian_above_field (five planes). */
long nfgs_above_field_workspace_bytes(const nfgs_plan* plan);

/* Replaces NFGSSolver.gaussian_above_bottom_field (nfgs_solver.py:4921-4943):
 * out (3 planes, component-first x, y, z) = (-d/dx, -d/dy) of
 * dPhi/dz = idct2(k * phi_above_hat) with the alpha-consistent spectral
 * gradient (:4385-4406: DST_x x DCT_y and DCT_x x DST_y, zero on the walls
 * each sine transform leaves out), and B_z = bz_above.  phi_above_hat and
 * bz_above are the planes nfgs_setup_gaussian wrote. */
int nfgs_gaussian_above_field(const nfgs_plan* plan, const void* phi_above_hat,
                              const void* bz_above, void* out, void* workspace, void* stream);

/* ---- post-selection diagnostics ---------------------------------------- */

/* out[0] = sum (a-b)^2, out[1] = sum b^2 over n elements (FP64 device
 * doubles).  The relative rms differences of the toroidal boundary
 * diagnostics (nfgs_solver.py:3047-3055) and of final_vs_best_B_relative
 * (nfgs_solver.py:3820-3835). */
int nfgs_sq_diff(const nfgs_plan* plan, const void* a, const void* b, long n, double* out,
                 void* stream);

/* Workspace bytes needed by nfgs_source_diagnostics (24 planes). */
long nfgs_source_diagnostics_workspace_bytes(const nfgs_plan* plan);

/* compute_source_diagnostics of a selected B (nfgs_solver.py:7014-7196).
 * Writes NFGS_SD_COUNT device doubles (sums over one plane; the host takes
 * rms = sqrt(sum / (nx*ny))).  Needs nz >= 3. */
enum {
  NFGS_SD_BOTTOM_RECON_SQ = 0, /* sum (B_bg,z(0) + Bz_above - Bz_obs)^2          :7030-7033 */
  NFGS_SD_BOTTOM_RECON_MAX,    /* max |B_bg,z(0) + Bz_above - Bz_obs|                       */
  NFGS_SD_CORR_BOTTOM_SQ,      /* sum (Bc_z(0) - Bz_above)^2                          :7033 */
  NFGS_SD_TOP_ROBIN_SQ,        /* top Robin residual of Bc_z                     :7035-7042 */
  NFGS_SD_TOP_PHI_ROBIN_SQ,    /* top Robin residual of Phi_c                    :7044-7053 */
  NFGS_SD_TOP_UPWARD_SQ,       /* upward-going part at the top                   :7054-7059 */
  NFGS_SD_TOP_BCZ_SQ,          /* sum Bc_z(top)^2 (ratio denominator)                      */
  NFGS_SD_BOTTOM_DIV_SQ,       /* sum (div_xy B(0) - D0)^2                       :7067-7081 */
  NFGS_SD_D0_SQ,               /* sum D0^2                                                 */
  NFGS_SD_DPHI_RES_SQ,         /* sum (dPhi_c/dz(0) - dphi_target)^2             :7082-7090 */
  NFGS_SD_COUNT
};
int nfgs_source_diagnostics(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* b,
                            const void* d0, const void* dphi_target, const void* bz_above,
                            void* workspace, double* out, void* stream);

/* ---- per-iteration diagnostics (record_iteration_diagnostics) ---------- *
 * The reference's iteration_step diagnostics at mu = 1 (nfgs_solver.py:
 * 6660-6680, :6806-6931).  Each function writes FP64 sums on the device; the
 * host forms relative RMS values from them (sqrt(num / n) / max(sqrt(den / n),
 * 1e-30), n the element count of the slice) in the reference's terms:
 *   phi/az map and update residuals (whole volumes),
 *   B_c,z residual (d_zz + lap_xy) lap_xy(Phi_c) - S_Bcz on planes 2..nz-2
 *     (1..nz-2 when nz <= 4), J_z residual -(d_zz + lap_xy) lap_xy(Az) - S_Jz
 *     on planes 1..nz-2,
 *   poloidal (d_zz + lap_xy) Phi_c - source_phi over the B_c,z planes,
 *     source_phi = IDCT2(DCT2(S_Bcz) / lambda), zero mode 0; its scale is the
 *     largest RMS of the three terms,
 *   toroidal (d_zz + lap_xy) Az - IDST2(-DST2(S_Jz) / lambda_D) inside,
 *   the rho-mixed first row (1-rho) poloidal[1] - rho (dPhi_c/dz(0) -
 *     target) / dz against its two terms and source_phi[1]. */
enum {
  NFGS_DR_PHI_MAP = 0, NFGS_DR_PHI_REF = 1, NFGS_DR_AZ_MAP = 2, NFGS_DR_AZ_REF = 3,
  NFGS_DR_BCZ_RES = 4, NFGS_DR_BCZ_SRC = 5, NFGS_DR_JZ_RES = 6, NFGS_DR_JZ_SRC = 7,
  NFGS_DR_POL_RES = 8, NFGS_DR_POL_D2Z = 9, NFGS_DR_POL_LAP = 10, NFGS_DR_POL_SRC = 11,
  NFGS_DR_TOR_RES = 12, NFGS_DR_TOR_D2Z = 13, NFGS_DR_TOR_LAP = 14, NFGS_DR_TOR_SRC = 15,
  NFGS_DR_MIX_RES = 16, NFGS_DR_MIX_PDE = 17, NFGS_DR_MIX_DER = 18, NFGS_DR_MIX_SRC = 19,
  NFGS_DR_COUNT = 20
};
/* Face fractions of a state and the bottom imprint of its B: counts of face
 * points with |B_f|^2 <= weak_field_floor_b2 and with |J_f.B_f| /
 * max(|B_f|^2, floor) > alpha_max over the (nz-1)*ny*nx vertical faces
 * (the state's own J_z), and sum |B(z=0) - B_obs|^2, sum |B_obs|^2. */
enum { NFGS_DF_WEAK = 0, NFGS_DF_CLIP = 1, NFGS_DF_IMPRINT = 2, NFGS_DF_OBS = 3,
       NFGS_DF_COUNT = 4 };
/* Bytes of the workspace of nfgs_iteration_residuals / nfgs_face_fractions
 * (five volumes, the interior volume of S_Jz and the FP64 partials). */
long nfgs_iteration_diagnostics_workspace_bytes(const nfgs_plan* plan);
/* S_Jz on the lateral walls of every plane (a volume; only wall nodes are
 * written) from the parallel current a step starts from (ws.jp before
 * nfgs_advance overwrites it). */
int nfgs_diag_source_edges(const nfgs_plan* plan, const void* jp, void* sj_walls, void* stream);
/* The NFGS_DR_* sums of one step from (phi_old, az_old) to (phi_new, az_new)
 * with the step's sources: s_bcz (volume) and s_jz (interior planes) as
 * nfgs_advance left them in ws, sj_walls from nfgs_diag_source_edges. */
int nfgs_iteration_residuals(const nfgs_plan* plan, const void* phi_old, const void* az_old,
                             const void* phi_new, const void* az_new, const void* s_bcz,
                             const void* s_jz, const void* sj_walls, const void* dphi_target,
                             void* workspace, double* out, void* stream);
/* The NFGS_DF_* sums of the state (b, az), or of (b, jz) with a caller's
 * centre J_z volume jz (az is then not read); bx_obs, by_obs the observed
 * planes on the calculation grid (bz_obs is the boundary's). */
int nfgs_face_fractions(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* b,
                        const void* az, const void* jz, const void* bx_obs, const void* by_obs,
                        void* workspace, double* out, void* stream);

/* ---- common scoring of output fields (FP64, any grid, no plan) --------- */

/* Second-order curl and divergence of b (3*nz*ny*nx, component-first z-y-x):
 * centred differences inside, the second-order one-sided forms on the end
 * planes, forward differences on an axis of 2 samples, 0 on an axis of 1
 * (nfgs/diagnostics/common_fd2.py:23-61).  j (3 volumes) or div (1 volume)
 * may be NULL. */
int nfgs_fd2_curl_div(const double* b, int nz, int ny, int nx, double dx, double dy, double dz,
                      double* j, double* div, void* stream);

/* Sums over the box [z0,z1) x [y0,y1) x [x0,x1) for the agreement metrics of
 * util/metrics.py:128-176 (b_ref against b), sigma_J (j against b,
 * common_fd2.py:64-83 and util/metrics.py:66-92 with |B| floored at b2min)
 * and the normalised divergence of the controlled evaluation (div and |b|,
 * nfgs/_backend/scripts/recompute_td_referee_diagnostics.py:344-400).
 * b_ref, j or div may be NULL (their sums are then 0).  out: NFGS_SC_COUNT
 * device doubles; workspace: nfgs_scores_workspace_bytes() bytes of device
 * memory. */
enum {
  NFGS_SC_DOT = 0,   /* sum b.B                      */
  NFGS_SC_REF2,      /* sum |B|^2                    */
  NFGS_SC_SOL2,      /* sum |b|^2                    */
  NFGS_SC_COS,       /* sum b.B / (|B||b| + 1e-12)   */
  NFGS_SC_DIFF,      /* sum |b - B|                  */
  NFGS_SC_REFN,      /* sum |B|                      */
  NFGS_SC_DIFF_REL,  /* sum |b - B| / (|B| + 1e-12)  */
  NFGS_SC_JXB,       /* sum |j x b| / max(|b|, b2min) */
  NFGS_SC_JN,        /* sum |j|                      */
  NFGS_SC_BN,        /* sum |b|                      */
  NFGS_SC_DIVABS,    /* sum |div|                    */
  NFGS_SC_COUNT
};
long nfgs_scores_workspace_bytes(void);
int nfgs_field_scores(const double* b_ref, const double* b, const double* j, const double* div,
                      int nz, int ny, int nx, int z0, int z1, int y0, int y1, int x0, int x1,
                      double b2min, double* workspace, double* out, void* stream);

/* nfgs_field_scores without a workspace argument: the per-block partials
 * live in a device array of the library (nfgs_field_scores_static_capacity()
 * doubles, 88 KiB, held on each GPU it runs on until the process ends), so
 * the call allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  Results equal nfgs_field_scores'.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU race. */
long nfgs_field_scores_static_capacity(void);
int nfgs_field_scores_static(const double* b_ref, const double* b, const double* j,
                             const double* div, int nz, int ny, int nx, int z0, int z1, int y0,
                             int y1, int x0, int x1, double b2min, double* out, void* stream);

/* Relative error of an outward bottom field against a target, for the CCI
 * comparison.  b (target) and c (candidate) are three planes (x, y, z
 * components) of ny*nx FP64.  Over the window [y0,y1) x [x0,x1):
 *   out[0] = sum |c-b|^2, out[1] = sum |b|^2, out[2] = max |b|,
 *   out[3] = number of pixels in the window;
 * and for each of the ncut (<= 16) threshold fractions frac[k] (host array),
 * with t_k = frac[k] * (max|b| + eps):
 *   out[4+2k] = number of pixels with |b| > t_k,
 *   out[5+2k] = sum over those pixels of |c-b| / |b|.
 * out holds 4 + 2*ncut device doubles.  Replaces the host arithmetic of
 * cci_errors (nfgs/controlled.py:55-68; thresholds frac*max|b|, eps = 0) and
 * of bottom_bgt_relative_error
 * (nfgs/_backend/scripts/plot_nfgs_td_robin_metrics_vs_er.py:299-326;
 * floor_frac*(max|b| + 1e-30)). */
int nfgs_outward_errors(const double* b, const double* c, int ny, int nx, int y0, int y1, int x0,
                        int x1, const double* frac, int ncut, double eps, double* out,
                        void* stream);

/* ------------------------------------------------------------------------
 * Coupled horizontal-harmonic closures (horizontal_zero_mode_policy
 * "coupled_k0" and "constrained_spline_harmonic"; nfgs_solver.py:5058-5398,
 * :5952-6067, :6760-6791).  The state's uniform horizontal field H(z)
 * (2 x nz, the plan dtype; nfgs_boundary.h_profile) follows the horizontal
 * mean of the parallel current the PT scalars do not carry.
 * ------------------------------------------------------------------------ */
enum { NFGS_K0_COUPLED = 1, NFGS_K0_SPLINE = 2 };
enum { NFGS_K0_STAGGERED = 0, NFGS_K0_LEGACY = 1 };

/* Doubles of the closure table of nz planes (nz*nz + nz). */
long nfgs_k0_table_count(int nz);
/* Build the table (device, FP64): the profile of one component is
 * H = M q + v h0 for a current target q (nz) and bottom value h0.  kind
 * NFGS_K0_COUPLED takes the least-squares solve of nfgs_solver.py:5262-5398
 * (discretization NFGS_K0_STAGGERED: face D1 and three-point curvature;
 * NFGS_K0_LEGACY: plane-centred, :5308-5353); NFGS_K0_SPLINE the
 * constrained cubic B-spline of :5149-5261 (basis_size functions).  Replaces
 * _initialize_coupled_k0 / _initialize_constrained_spline_harmonic.  With
 * condition non-NULL the spline's KKT 2-norm condition number is written
 * there (host double).  Synchronous; builds on the host, copies to tables. */
int nfgs_k0_tables(int nz, double dz, int kind, int discretization, double smoothness,
                   int basis_size, double* tables, double* condition);

/* One closure update (nfgs_solver.py:5952-6067, :6760-6791). */
typedef struct {
  int kind, discretization; /* as for nfgs_k0_tables                              */
  double mu;                /* k0_update_mu in (0, 1]                             */
  const double* tables;     /* nfgs_k0_table_count(nz) doubles                    */
  const void* h_bottom;     /* 2: the bottom harmonic target (plan dtype)        */
  const void* h_old;        /* 2*nz profile of the step's starting state          */
  void* h_new;              /* 2*nz profile of the new state (may equal h_old)    */
  double* work;             /* 4*nz doubles: trapezoid means by height            */
  double* diag;             /* NULL or 4*nz doubles: want_x, want_y, H*_x, H*_y    */
  double* residual;         /* NULL or 1 double: relative ODE residual of H*      */
} nfgs_k0_step;
/* h_new = mu H* + (1 - mu) h_old with H* the profile of want = mean(J_par) -
 * mean(J_pt) per height (the constant DCT-I mode of each plane), J_par the
 * step's parallel current jp (3*N) and J_pt the PT current of (phi, az). */
int nfgs_k0_update(const nfgs_plan* plan, const nfgs_k0_step* k0, const void* jp,
                   const void* phi, const void* az, void* stream);
/* nfgs_advance_ex with the closure: the update runs between the Poisson
 * solves and the closure's state pass (nfgs_k0_state_pass, built in the
 * workspace's s_bcz and hat0), which takes bnd->h_profile == k0->h_new.
 * With zero_source (the gs_cfcf start) h_new is a copy of h_old. */
int nfgs_advance_k0(const nfgs_plan* plan, const nfgs_boundary* bnd, const nfgs_workspace* ws,
                    void* phi_out, void* az_out, void* b_out, int zero_source, int poloidal_mode,
                    const nfgs_k0_step* k0, void* stream);
/* Polarity masks of the bottom parallel current (nfgs_solver.py:6694-6701). */
enum { NFGS_POLARITY_BOTH = 0, NFGS_POLARITY_POSITIVE = 1, NFGS_POLARITY_NEGATIVE = 2 };
/* One step from the state (phi_old, az_old) with relaxation and the polarity
 * mask (iteration_step(mu, polarity_mask), nfgs_solver.py:6694-6701,
 * :6776-6790): the bottom row of ws->jp (the old state's J_par) is kept only
 * where the observed B_z is positive (NFGS_POLARITY_POSITIVE) or negative
 * (NFGS_POLARITY_NEGATIVE); the sources and solves follow; with k0 (NULL for
 * the frozen closure) the closure update; then, away from the fixed planes,
 * phi_out = mu phi* + (1 - mu) phi_old and az_out = mu az* + (1 - mu) az_old
 * (0 < mu <= 1), and the state pass of the relaxed state.  The fixed planes
 * of phi_out and az_out (bottom planes, Az's top plane and side walls) must
 * hold their values, as the state slots of nfgs_advance_ex do. */
int nfgs_advance_relaxed(const nfgs_plan* plan, const nfgs_boundary* bnd,
                         const nfgs_workspace* ws, const void* phi_old, const void* az_old,
                         double mu, int polarity, void* phi_out, void* az_out, void* b_out,
                         int poloidal_mode, const nfgs_k0_step* k0, void* stream);
/* nfgs_state_pass of a closure's state (bnd->h_profile: the state's
 * profile).  The coupled closure's current is the frozen rule's; the spline
 * closure's (nfgs_solver.py:5919-5941) has J_z the cell projection of the
 * face average of the PT J_z and the harmonic current (-dH_y/dz, dH_x/dz) on
 * the side walls, which enter J_par, the monitors and j_out.  work: N + 2*nz
 * values of the plan dtype for the spline closure, else unused (may be NULL). */
int nfgs_k0_state_pass(const nfgs_plan* plan, const nfgs_boundary* bnd, int kind, const void* phi,
                       const void* az, void* b_out, void* jp_out, void* j_out, double* partials,
                       void* work, void* stream);

/* ------------------------------------------------------------------------
 * soft_cfcf start (init_mode="soft_cfcf", nfgs_solver.py:6152-6611).
 * ------------------------------------------------------------------------ */
/* Bytes of the work buffer of nfgs_soft_start / nfgs_soft_diagnostics. */
long nfgs_soft_workspace_bytes(const nfgs_plan* plan);
/* The start's scalars (replaces _soft_cfcf_initial_state's modal solves and
 * transforms, :6423-6470): every DCT-I mode of Phi_c and interior DST-I mode
 * of Az gets the profile minimising the modal curl-current form with weight
 * eps (soft_cfcf_epsilon) on the current and a diagonal ridge (:6322-6402),
 * by a banded Cholesky factorisation per mode in FP64.  phi_out, az_out: N
 * values each; plane 0 is the bottom potential / Az_bottom, the side walls of
 * Az are zero and its top plane is the profile's (not zero).  out: 6 doubles,
 * the Phi_c modal curl and current energies (DCT-I end weights applied), the
 * count of Phi_c modes whose solve was not finite, then the same for Az. */
int nfgs_soft_start(const nfgs_plan* plan, const nfgs_boundary* bnd, double eps, double ridge,
                    void* phi_out, void* az_out, void* work, double* out, void* stream);
/* Sums of squares behind the start's diagnostics (:6542-6611) for a state
 * (phi, az) with its B and J (3*N each) and the observed B_x, B_y (ny*nx):
 * out[0..10] = bottom Phi_c, bottom dPhi_c/dz against idct2(k dct2(Phi_above)),
 * bottom Az, bottom B_z, bottom J_z, bottom current-carrying B_h, bottom
 * observed B_h, the transparent top rows of Phi_c and Az (in space), |curl J|^2
 * and |J|^2 over the volume. */
int nfgs_soft_diagnostics(const nfgs_plan* plan, const nfgs_boundary* bnd, const void* phi,
                          const void* az, const void* b, const void* j, const void* bx_obs,
                          const void* by_obs, void* work, double* out, void* stream);

/* ------------------------------------------------------------------------
 * Learned warm start.  No counterpart in the reference: its iterate() starts
 * only from gs_cfcf (nfgs_solver.py:6613-6645).  These are the kernels of a
 * small cosine-basis spectral neural operator that maps the gs_cfcf state
 * (B and the parallel current J_par) and rho to a correction of J_par; one
 * solver step (nfgs_advance_ex) from the corrected current gives the start.
 * The host (nfgs_amd.learned) composes them into the network's forward and
 * backward passes and its optimiser step.  FP32 activations (C, nz, ny, nx)
 * with x fastest; bfloat16 weights (uint16_t bit patterns); every grid axis
 * >= 16.
 * ------------------------------------------------------------------------ */
#define NFGS_FNO_WIDTH 32
#define NFGS_FNO_MODES 16
#define NFGS_FNO_FEATURES 8
#define NFGS_FNO_HIDDEN 64
#define NFGS_FNO_OUTPUTS 3
#define NFGS_FNO_IDENTITY 0
#define NFGS_FNO_GELU 1
#define NFGS_FNO_PROJECT 0
#define NFGS_FNO_EXPAND 1

/* out[o] = act(W in + b + E_x spec), W (cout, cin) bfloat16, spec (cout,
 * nz*ny, 16) or NULL, basis_x (nx, 16).  Shapes (cin, cout, act, spec):
 * (8,32,id,-) lift, (32,32,gelu|id,spec) spectral layer, (32,64,gelu,-) and
 * (64,3,id,-) head. */
int nfgs_fno_pointwise(const float* in, int cin, const uint16_t* weight, const float* bias,
                       const float* spec, const float* basis_x, int cout, int activation,
                       int nx, int ny, int nz, float* out, void* stream);
/* Backward of nfgs_fno_pointwise: g = grad_out + E_x grad_out_spec,
 * grad_z = g act'(z), grad_in = W^T grad_z (either output may be NULL where
 * the layer allows it). */
int nfgs_fno_pointwise_backward(const float* in, int cin, const uint16_t* weight,
                                const float* bias, const float* spec, const float* basis_x,
                                int cout, int activation, const float* grad_out,
                                const float* grad_out_spec, int nx, int ny, int nz,
                                float* grad_z, float* grad_in, void* stream);
/* grad_w[o][i] (+)= sum_p a[o][p] b[i][p], grad_b[o] (+)= sum_p a[o][p];
 * (ca, cb) in {(32,8), (32,32), (64,32), (3,64)}. */
long nfgs_fno_outer_sum_workspace_bytes(int ca, int cb);
int nfgs_fno_outer_sum(const float* a, int ca, const float* b, int cb, long npoints,
                       float* grad_w, float* grad_b, int accumulate, void* workspace,
                       void* stream);
/* Truncated DCT-II along the middle axis of an (outer, n, inner) array with
 * the (n, 16) basis: PROJECT maps n -> 16 (basis^T), EXPAND 16 -> n. */
int nfgs_fno_axis(const float* in, float* out, const float* basis, long outer, int n, long inner,
                  int direction, void* stream);
/* Per-mode channel mix with R (nmodes, 32, 32) bfloat16: out = R in, or R^T in. */
int nfgs_fno_mix(const uint16_t* weight, const float* in, float* out, int nmodes, int transpose,
                 void* stream);
/* grad_weight[m][o][i] += grad_out[o][m] in[i][m]. */
int nfgs_fno_mix_grad(const float* grad_out, const float* in, float* grad_weight, int nmodes,
                      void* stream);
/* Mean squared error of the n output values and its gradient; loss is one
 * device double. */
long nfgs_fno_loss_workspace_bytes(void);
int nfgs_fno_loss(const float* out, const float* target, long n, float* grad_out, double* loss,
                  void* workspace, void* stream);
/* Adam step on FP32 master weights; writes the bfloat16 copy and clears grad. */
int nfgs_fno_adam(float* master, uint16_t* weight, float* m, float* v, float* grad, long n,
                  double lr, double beta1, double beta2, double eps, int step, double grad_scale,
                  void* stream);
int nfgs_fno_to_bf16(const float* src, uint16_t* dst, long n, void* stream);
/* The eight input channels: B (3*N) times c_b, J_par (3*N) times c_j, rho,
 * z / (nz - 1).  dtype is the solver's (NFGS_F32 or NFGS_F64). */
int nfgs_fno_features(int dtype, const void* b, const void* jp, int nx, int ny, int nz, double c_b,
                      double c_j, double rho, float* feat, void* stream);
/* Training target (jp - jp_gs) inv_scale over n values. */
int nfgs_fno_targets(int dtype, const void* jp_gs, const void* jp, long n, double inv_scale,
                     float* target, void* stream);
/* Start current jp = jp_gs + scale * net over n values (jp may be jp_gs). */
int nfgs_fno_apply(int dtype, const void* jp_gs, const float* net, long n, double scale, void* jp,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NFGS_HIP_H */
