/*
 * mhm_hip.h - C ABI of the MI355X (gfx950) MHM local-problem stage.
 *
 * One library, libmhm_hip.so, built from the .hip sources in pymhm_amd/csrc.  Every
 * entry point returns an int status (MHM_OK on success); mhm_last_error()
 * returns the message of the last failure on the calling thread.  Array
 * arguments named d_* are device pointers on the handle's device; the caller
 * owns them.  Opaque handles own only their internal device workspaces.
 * Kernels are enqueued on the given HIP stream (NULL = default stream) and
 * are asynchronous unless the comment says the call synchronises.
 *
 * The reference (PyMHM) is pure Python and has no FFI; each entry point is
 * the native replacement of the Python function cited beside it (paths
 * relative to the reference repository root).  The reference side binds it
 * through ctypes, as pymhm_amd/_lib.py does.
 *
 * Layout conventions
 *   n    local DOFs of one macroelement, k retained modes (0 or 1), N = n + k
 *   m    multiplier (skeleton) DOFs of the macroelement
 *   R    standard response columns: 1 (source) + m (lifts) + k (retained action)
 *   Rows of per-macro vectors use the caller's ORIGINAL local numbering; the
 *   layout's band permutation is internal to the library.
 *   Column-major per macro: X[macro][column][row], row count N.
 */
#ifndef MHM_HIP_H
#define MHM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MHM_OK = 0,
  MHM_EINVAL = 1,    /* invalid argument -> ValueError                      */
  MHM_ESINGULAR = 2, /* non-positive / rank-deficient pivot -> LinearSolveError */
  MHM_ERESIDUAL = 3, /* residual criterion not met -> LinearSolveError      */
  MHM_ENOMEM = 4,    /* device allocation failed                            */
  MHM_EDEVICE = 5    /* HIP runtime error                                   */
};

/* Precision policy of factorisation and triangular solves.  Residuals,
 * refinement, the 2x2 bordered correction, condensation and reconstruction
 * are always binary64. */
enum { MHM_FP64 = 0, MHM_MIXED = 1 };

/* Per-macro status bits (mhm_batch_status). */
enum {
  MHM_STATUS_PIVOT = 1,    /* non-positive or tiny pivot in the band factor      */
  MHM_STATUS_RESIDUAL = 2, /* some column missed the residual criterion          */
  MHM_STATUS_NONFINITE = 4, /* non-finite operator, pivot or solution            */
  MHM_STATUS_COEFFICIENT = 8 /* last assembly: diffusion not finite and positive, or
                                source not finite, at some quadrature point        */
};

typedef struct mhm_layout mhm_layout;
typedef struct mhm_batch mhm_batch;

/* Mesh-only description of one local problem shape (a "layout").  All arrays
 * are HOST arrays, copied by mhm_layout_create.  Index arrays use the band
 * (factorisation) numbering; perm maps band row -> original local row. */
typedef struct mhm_layout_desc {
  int32_t n;              /* local DOFs                                         */
  int32_t k;              /* retained modes: 0 (SPD operator) or 1              */
  int32_t m;              /* multiplier columns                                 */
  int32_t nnz;            /* stored entries of the full symmetric pattern        */
  int32_t bandwidth;      /* max |i - j| over the pattern in band numbering      */
  int32_t pin;            /* band row receiving sigma = A_pp (k == 1), else -1   */
  const int32_t *perm;    /* [n] band row -> original row                        */
  const int32_t *rowptr;  /* [n+1] CSR of the pattern, band numbering            */
  const int32_t *cols;    /* [nnz] sorted column indices per row                 */
  /* template coupling B (may be NULL when every batch supplies its own) */
  const double *coupling; /* [m][n] column-major, band rows                      */
  const int32_t *coupling_rows; /* [ncoupling_rows] band rows where B may be nonzero */
  int32_t ncoupling_rows;
  /* template retained mode Z and constraint C (NULL: Z = 1; C required if k) */
  const double *kernel;     /* [n] band rows                                     */
  const double *constraint; /* [n] band rows                                     */
  /* optional assembly description (ne == 0: operators are supplied assembled) */
  int32_t ne;             /* local elements                                     */
  int32_t nb;             /* basis functions per element (3, 6 or 10)            */
  int32_t dim;            /* 2 or 3                                             */
  int32_t nq;             /* quadrature points per element                      */
  const int32_t *slot_ptr;   /* [nnz+1] contributions of each CSR slot           */
  const int32_t *slot_code;  /* [ne*nb*nb] (i*nb + j)*ne + e, fixed order         */
  const int32_t *load_ptr;   /* [n+1] contributions of each load row             */
  const int32_t *load_code;  /* [ne*nb] b*ne + e, fixed order                      */
  const double *grad_lambda; /* [ne][dim+1][dim] physical barycentric gradients  */
  const double *measure;     /* [ne] element areas / volumes                     */
  const double *qweights;    /* [nq] reference weights, unit sum                 */
  const double *dtable;      /* [nq][nb][dim+1] d phi_b / d lambda_i              */
  const double *vtable;      /* [nq][nb] phi_b                                    */
} mhm_layout_desc;

/* Macroface skeleton of a 2D layout (HOST arrays, copied): the boundary edges
 * of the local mesh cut into pieces that each lie inside one face segment, in
 * the side's counterclockwise fraction f in [0, 1], and the nonzero coupling
 * entries (band row, side, segment along f) with the pieces and local edge
 * basis functions (0 start, 1 end, 2 middle node of a P2 edge) that sum into
 * each.  The face space has nseg segments of one Legendre degree with breaks
 * symmetric about 1/2 (the second owner of a face sees them mirrored). */
typedef struct mhm_skeleton_desc {
  int32_t nsides, nseg, degree, edeg, ng;
  const double *breaks;      /* [nseg+1]                                          */
  const double *gx, *gw;     /* [ng] Gauss-Legendre nodes and weights on [-1, 1]  */
  int32_t npieces;
  const double *piece_f;     /* [npieces][4]: f0, f1 of the piece, e0, e1 of its edge */
  int32_t nentries;
  const int32_t *entry_row, *entry_side, *entry_seg; /* [nentries]                 */
  const int32_t *entry_ptr;  /* [nentries+1]                                      */
  const int32_t *contrib;    /* [entry_ptr[nentries]] piece * 3 + local basis      */
} mhm_skeleton_desc;

/* Diagnostics. */
int mhm_last_error(char *buf, size_t len);
int mhm_device_count(int32_t *count);
int mhm_version(int32_t *major, int32_t *minor);

/* Layout (mesh-only data, shared by every macro of one shape).
 * Replaces the per-cell mesh/DOF/trace preparation of the providers:
 * src/pymhm/_legacy/models/darcy/primal.py:44-78 (trace_coupling, nodal
 * space, constraints = M Z / Z^T M Z), fem/scalar/triangle.py:65-92,160-205. */
int mhm_layout_create(int32_t device, const mhm_layout_desc *desc, mhm_layout **out);
int mhm_layout_destroy(mhm_layout *layout);
int mhm_layout_info(const mhm_layout *layout, int32_t *n, int32_t *bandwidth, int32_t *nnz);

/* Batch of `macros` local problems of one layout with resident factors.
 * Replaces BatchedFactorization.__init__ (src/pymhm/execution/cuda.py:178-215)
 * and the grouping of condense_batched (cuda.py:291-312).  Allocates its
 * workspaces on the layout's device. */
int mhm_batch_create(const mhm_layout *layout, int32_t macros, int32_t precision,
                     int32_t max_columns, mhm_batch **out);
int mhm_batch_destroy(mhm_batch *batch);

/* Local assembly from coefficients tabulated at quadrature points:
 * d_kq[macro][element][q] (scalar diffusion), d_fq[macro][element][q] source
 * (NULL: constant f_const).  Produces the stiffness in the layout pattern, the
 * load vector and A Z (accumulated in double-double).
 * Replaces scalar_operators/p1_operators + the Numba Gram kernel:
 * src/pymhm/fem/scalar/triangle.py:208-257, fem/scalar/operators.py:154-197,
 * fem/scalar/_integration.py:43-88, and LocalProblem._retained_action
 * (core/contracts.py:201-211). */
int mhm_batch_assemble(mhm_batch *batch, const double *d_kq, const double *d_fq, double f_const,
                       void *stream);

/* Assembled operators instead of assembly: d_values[macro][nnz] in the
 * layout's band-numbered CSR slot order, d_loads[macro][n] ORIGINAL rows.
 * Replaces the dense upload of condense_batched (cuda.py:304-309). */
int mhm_batch_set_operator(mhm_batch *batch, const double *d_values, const double *d_loads,
                           void *stream);

/* Copy the assembled operators out: d_values[macro][nnz] (layout slot
 * order), d_loads[macro][n] and d_az[macro][n] (A Z), ORIGINAL rows; any may
 * be NULL.  Mirror of the (matrix, load, _retained_action) fields of
 * LocalProblem (core/contracts.py:92-211) for inspection and tests. */
int mhm_batch_operator(mhm_batch *batch, double *d_values, double *d_loads, double *d_az,
                       void *stream);

/* Per-macro coupling data (all optional, NULL keeps the layout default):
 * d_colmap[macro][m] (int32) and d_colscale[macro][m] give column a of the
 * macro's coupling as colscale[a] * (layout coupling column colmap[a]) - a
 * signed column permutation that carries macroface orientation (segment
 * order reversal and odd Legendre modes, fem/traces/interval.py:54-89, with
 * the orientation signs of meshes/triangle.py:53-74); d_couplings[macro][m][n]
 * full per-macro couplings (ORIGINAL rows); d_constraints[macro][n] and
 * d_kernels[macro][n] (ORIGINAL rows). */
int mhm_batch_set_coupling(mhm_batch *batch, const int32_t *d_colmap, const double *d_colscale,
                           const double *d_couplings, const double *d_constraints,
                           const double *d_kernels, void *stream);

/* Scale the constraint, equilibrate, pin the retained mode and factor every
 * macro's operator (band Cholesky of D (A + sigma e_p e_p^T) D).  With the
 * side stream on (mhm_batch_set_side_stream) and the LDS-ring panel factor,
 * the first pass of the next mhm_batch_condense (its scaled right-hand
 * sides) is queued on the batch's side stream beside the factor; the next
 * solving call on the batch orders itself after it.
 * Replaces local_condensation_matrix (core/condensation.py:150-174),
 * _equilibrated_lu and the pivot check (linalg/linear.py:400-422),
 * _factor_batch (execution/cuda.py:114-136). */
int mhm_batch_factor(mhm_batch *batch, void *stream);

/* Solve the standard columns [f, B, A Z] of the augmented system with up to
 * `max_refine` binary64 refinement steps (residual criterion rtol, per column,
 * as linalg/linear.py:132-152), then condense into
 * d_blocks[macro][(m+k)*(m+k)] (row-major) and d_loads[macro][m+k].
 * Replaces condense_local (core/condensation.py:118-122,125-147),
 * LinearFactorization.solve (linalg/linear.py:262-310),
 * local_response_from_solution (:177-191), local_global_contribution
 * (core/contributions.py:14-74) and local_condensed_load
 * (core/reconstruction.py:91-127). */
int mhm_batch_condense(mhm_batch *batch, int32_t max_refine, double rtol, double *d_blocks,
                       double *d_loads, void *stream);

/* Whether mhm_batch_condense condenses the macros that converge early on the
 * batch's own side stream while the others finish refining, and
 * mhm_batch_factor builds the condensation's first pass there (1, the
 * default), or everything runs on the caller's stream (0).  No counterpart in the reference
 * (its condensation is one host loop); for callers that run many batches
 * concurrently on their own streams, such as topology buckets of uneven
 * macros: a side stream shares the device's few hardware queues with those
 * streams, and its wait for the batch's settled pass holds up every kernel
 * queued behind it. */
int mhm_batch_set_side_stream(mhm_batch *batch, int32_t enable);

/* Band solve kernel family of the first pass of a solve (`first`) and of its
 * refinement passes (`refine`), each one of MHM_SOLVE_*; MHM_SOLVE_AUTO picks
 * by band width and column count.  A family that cannot serve the batch's
 * band and column count makes the solving call fail with MHM_EINVAL.  No
 * counterpart in the reference (SciPy picks its kernels); set by the tile
 * autotuner (pymhm_amd/tuning.py), which times the families on the batch. */
#define MHM_SOLVE_AUTO 0
#define MHM_SOLVE_MFMA 1       /* binary64 matrix cores, one wavefront per macro */
#define MHM_SOLVE_MFMA_TILES 2 /* binary64 matrix cores, one wavefront per 16-column tile */
#define MHM_SOLVE_WAVE 3       /* one wavefront per macro, lane = column (half-bandwidth <= 95) */
#define MHM_SOLVE_BLOCKED 4    /* four wavefronts per macro over column groups (half-bandwidth <= 255) */
int mhm_batch_set_solve_variants(mhm_batch *batch, int32_t first, int32_t refine);

/* Copy the responses of the last mhm_batch_condense in ORIGINAL row order:
 * d_source[macro][n], d_lifts[macro][n][m] (row-major like numpy (n, m)),
 * d_retained[macro][n] (E = Z - R A Z, or Z when A Z == 0).  Any may be NULL.
 * Replaces LocalResponse (core/contracts.py:321-333). */
int mhm_batch_responses(mhm_batch *batch, double *d_source, double *d_lifts, double *d_retained,
                        void *stream);

/* Solve `ncols` new right-hand sides of the augmented system against the
 * resident factors: d_rhs[macro][col][N] -> d_x[macro][col][N], ORIGINAL rows,
 * row n (k == 1) is the constraint row.  ncols <= max_columns.
 * Replaces BatchedFactorization.solve (execution/cuda.py:223-255) and
 * OfflineLocalProblem.response (core/offline.py:121-127). */
int mhm_batch_solve(mhm_batch *batch, const double *d_rhs, double *d_x, int32_t ncols,
                    int32_t max_refine, double rtol, void *stream);

/* Per-macro element geometry for macros that share a local topology but not
 * a shape (uneven polygonal macros): d_grad_lambda[macro][element][dim+1][dim]
 * (barycentric gradients) and d_measure[macro][element] replace the layout's
 * for mhm_batch_assemble and mhm_batch_sources.  Replaces the per-cell
 * geometry of physical_simplex_tabulation (fem/reference.py:763-820) and
 * TriangleMesh.areas (meshes/triangle.py:122-147) for cells of
 * PolygonMesh.submesh (meshes/polygonal.py:217-265). */
int mhm_batch_set_geometry(mhm_batch *batch, const double *d_grad_lambda, const double *d_measure,
                           void *stream);

/* New source fields against the resident factors, `ncols` at once (the online
 * step: OfflineLocalProblem.response, core/offline.py:121-127, with
 * LocalProblem.with_load, core/contracts.py:275-298).  For each field c the
 * loads are assembled from d_fq[c][macro][element][q] (NULL: f_const
 * everywhere) with the operations of mhm_batch_assemble, [f; 0] is solved with
 * the refinement of mhm_batch_condense, and the condensed loads
 * [B_t^T s ; (A Z)^T s - Z^T f] (core/reconstruction.py:91-127) go to
 * d_loads[c][macro][m+k]; d_source[c][macro][n] (original rows) receives the
 * source responses s when not NULL.  A field's results do not depend on the
 * other fields of the call.  Requires a layout with an assembly description
 * and ncols <= max_columns. */
int mhm_batch_sources(mhm_batch *batch, const double *d_fq, double f_const, int32_t ncols,
                      int32_t max_refine, double rtol, double *d_loads, double *d_source, void *stream);

/* Fields u = s - L lambda + E c for `ncases` trace/coarse vectors:
 * d_trace[macro][case][m] (local trace order), d_coarse[macro][case][k],
 * d_fields[macro][case][n] ORIGINAL rows.
 * Replaces reconstruct_response (core/reconstruction.py:130-141). */
int mhm_batch_reconstruct(mhm_batch *batch, const double *d_trace, const double *d_coarse,
                          double *d_fields, int32_t ncases, void *stream);

/* mhm_batch_reconstruct with the source responses of new loads:
 * d_source[case][macro][n] (original rows, as mhm_batch_sources returns them)
 * in place of the standard source response.  The online reconstruction of
 * OfflineHybridSystem.solve (core/offline.py:234-307, per cell
 * reconstruct_response with the new LocalResponse, core/offline.py:121-127). */
int mhm_batch_reconstruct_sources(mhm_batch *batch, const double *d_source, const double *d_trace,
                                  const double *d_coarse, double *d_fields, int32_t ncases, void *stream);

/* Synchronise the stream and copy the per-macro status bits to host memory
 * (host_status[macros]); also returns the bitwise OR in *any.  The bits are
 * those of the current factors (PIVOT, NONFINITE operator; kept until the next
 * mhm_batch_factor), of the last solving call (RESIDUAL, NONFINITE data; each
 * mhm_batch_condense/_solve/_sources starts from none) and of the last
 * assembly (COEFFICIENT). */
int mhm_batch_status(mhm_batch *batch, int32_t *host_status, int32_t *any, void *stream);

/* Per-macro refinement counts of the last solve (host_counts[macros]); synchronises. */
int mhm_batch_iterations(mhm_batch *batch, int32_t *host_counts, void *stream);

/* Row-major COO coordinates of `count` blocks of width w:
 * d_rows[c][i*w+j] = d_indices[c][i], d_cols[c][i*w+j] = d_indices[c][j].
 * Replaces contribution_coordinates (core/_assembly.py:9-25). */
int mhm_scatter_coordinates(const int64_t *d_indices, int32_t width, int32_t count,
                            int64_t *d_rows, int64_t *d_cols, void *stream);

/* Ordered gather of contributions into global vectors/values: for every
 * output g, out[g] = sum over t in [ptr[g], ptr[g+1]) of src[code[t]] in the
 * stored (cell) order; when d_abs_out != NULL it also receives sum |src|.
 * Replaces accumulate_load (core/_assembly.py:28-42) and the duplicate
 * summation of the COO -> CSC conversion (core/contributions.py:175-185). */
int mhm_ordered_gather(const int64_t *d_ptr, const int64_t *d_code, const double *d_src,
                       int64_t count, double *d_out, double *d_abs_out, void *stream);

/* ---------------------------------------------------------------- dense batched LU
 * General real matrices, the reference's resident factor object
 * BatchedFactorization(matrices, rtol, atol) (src/pymhm/execution/cuda.py:165-288):
 * row then column max-abs equilibration (:198-207), LU with partial pivoting
 * and the rank test of _factor_batch (:114-136), solves with up to two
 * binary64 corrections and per-column acceptance (:223-255, _apply :257-263).
 * Arrays are row-major like numpy: matrices [batch][n][n], right-hand sides
 * and solutions [batch][n][nrhs].  Per-matrix status bits go to the caller's
 * device array d_status[batch] (cleared by each call); the calls enqueue work
 * on `stream` and do not synchronise unless stated. */
enum {
  MHM_DENSE_NONFINITE_MATRIX = 1, /* -> ValueError "matrices must be finite ..."    */
  MHM_DENSE_ZERO_ROW = 2,         /* -> LinearSolveError "zero row"                 */
  MHM_DENSE_ZERO_COLUMN = 4,      /* -> LinearSolveError "zero column"              */
  MHM_DENSE_SINGULAR = 8,         /* exactly zero pivot: "singular pivot"            */
  MHM_DENSE_RANK = 16,            /* min|U_jj| <= eps n max|LU|: "rank deficient"    */
  MHM_DENSE_NONFINITE_RHS = 32,   /* -> ValueError "rhs must be finite ..."          */
  MHM_DENSE_NONFINITE_RESULT = 64,/* "nonfinite coefficients"                        */
  MHM_DENSE_RESIDUAL = 128        /* "residual criterion failed"                     */
};

typedef struct mhm_dense mhm_dense;

/* Copy and factor `batch` matrices on `device`; factor status to d_status.
 * Replaces BatchedFactorization.__init__ (cuda.py:178-215). */
int mhm_dense_create(int32_t device, int32_t batch, int32_t n, const double *d_matrices, int32_t *d_status,
                     void *stream, mhm_dense **out);

/* Solve d_rhs[batch][n][nrhs] into d_x against the resident factors with the
 * acceptance max(atol / s, rtol ||b / s||) per column.  A matrix whose factor
 * failed is skipped (its d_x is left as it was).  Grows the solve workspace
 * when nrhs exceeds every earlier call's (that call synchronises the stream).
 * Replaces BatchedFactorization.solve (cuda.py:223-255). */
int mhm_dense_solve(mhm_dense *h, const double *d_rhs, int32_t nrhs, double rtol, double atol, double *d_x,
                    int32_t *d_status, void *stream);

/* Synchronise and release.  Replaces BatchedFactorization.close (cuda.py:265-273). */
int mhm_dense_destroy(mhm_dense *h);

/* One-shot factor and solve of the same batch (the per-cell LocalLinearSolver
 * seam, src/pymhm/core/assembly.py:32-47,150-156, and condense_batched's
 * factor-then-solve, cuda.py:304-309).  Allocates its workspace, synchronises
 * the stream and frees it before returning. */
int mhm_dense_factor_solve(int32_t device, int32_t batch, int32_t n, int32_t nrhs, const double *d_matrices,
                           const double *d_rhs, double *d_x, double rtol, double atol, int32_t *d_status,
                           void *stream);

/* mhm_dense_factor_solve without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a fixed per-device array of 64 MiB held until the process ends;
 * a call whose workspace (mhm_dense_scratch_bytes) exceeds it is rejected with
 * MHM_EINVAL.  Asynchronous.  Every call on one device shares that array:
 * calls that overlap on one GPU (different streams or threads) race. */
int mhm_dense_factor_solve_noalloc(int32_t device, int32_t batch, int32_t n, int32_t nrhs, const double *d_matrices,
                                   const double *d_rhs, double *d_x, double rtol, double atol, int32_t *d_status,
                                   void *stream);

/* Workspace bytes of one mhm_dense_factor_solve(_noalloc) call and the
 * no-allocation limit. */
int mhm_dense_scratch_bytes(int32_t batch, int32_t n, int32_t nrhs, int64_t *bytes, int64_t *limit);

/* ---------------------------------------------------------------- skeleton and boundary forms */

/* Attach the macroface skeleton to a layout (enables mhm_batch_skeleton_coupling).
 * Replaces the mesh-side part of trace_coupling / face_integration
 * (src/pymhm/fem/scalar/triangle.py:160-205, fem/scalar/operators.py:200-244). */
int mhm_layout_set_skeleton(mhm_layout *layout, const mhm_skeleton_desc *desc);

/* Per-macro couplings computed on the device from the layout's skeleton, the
 * side lengths d_side_len[macro][nsides] and orientation signs
 * d_side_sign[macro][nsides] (+1 first owner of the face, -1 second: the face
 * parameter runs against the side and odd modes flip).  Invalidates factors.
 * Replaces trace_coupling (fem/scalar/triangle.py:160-205) per macro. */
int mhm_batch_skeleton_coupling(mhm_batch *batch, const double *d_side_len, const double *d_side_sign, void *stream);

/* Per-macro constraints C = M 1 / (1^T M 1) from the batch's element measures
 * (per-macro geometry when set).  Invalidates factors.  Replaces
 * constraints = physical_mean / (kernel^T physical_mean) of the primal Darcy
 * provider (src/pymhm/_legacy/models/darcy/primal.py:70-72). */
int mhm_batch_mass_constraints(mhm_batch *batch, void *stream);

/* Copy the effective per-macro coupling d_coupling[macro][m][n] and constraint
 * d_constraints[macro][n] (ORIGINAL rows; either may be NULL).  Mirror of the
 * coupling / constraints fields of LocalProblem (core/contracts.py:92-211). */
int mhm_batch_coupling(mhm_batch *batch, double *d_coupling, double *d_constraints, void *stream);

/* Weak Dirichlet trace load over nfaces boundary faces: for face f and mode a,
 * load = d_length[f] * sum_q (d_basis[q][a] d_weights[q]) d_g[f][q]
 * (compensated), written to d_load[d_dof0[f] + a] when d_load is not NULL, and
 * applied to the global system as d_rhs[row] -= load, d_scale[row] += |load|
 * (either may be NULL).  d_basis[nq][fs] holds the face modes at the face
 * quadrature points, d_weights[nq] their weights (sum 1 over the face).
 * Replaces boundary_data (src/pymhm/fem/scalar/operators.py:247-321, scalar
 * Dirichlet data) and the boundary-load subtraction of the global assembly
 * (core/contributions.py:185-195). */
int mhm_dirichlet_load(int32_t nfaces, int32_t nq, int32_t fs, const double *d_basis, const double *d_weights,
                       const double *d_g, const double *d_length, const int64_t *d_dof0, double *d_load,
                       double *d_rhs, double *d_scale, void *stream);

/* ---------------------------------------------------------------- global hybrid solve
 * The assembled global system (CSC pattern of mhm_scatter / HybridScatter,
 * traces then coarse amplitudes) solved on the device by restarted GMRES,
 * right-preconditioned by weighted additive Schwarz over the contribution
 * cells (each cell's assembled w x w block inverted, w <= 64).  Replaces the
 * global solve of solve_hybrid_system (src/pymhm/core/system.py:355-528) and
 * SaddleBlockSolver (src/pymhm/linalg/block.py:99-237): prepare once per
 * matrix, then solve any number of right-hand sides against the resident
 * preconditioner, the reuse contract of a prepared factorization
 * (system.py:442-451, core/offline.py:234-307). */
typedef struct mhm_global mhm_global;

/* Pattern (host arrays): csc_indptr[size + 1], csc_indices[nnz] (rows
 * ascending within a column), and the cells' global indices
 * cell_idx[cell_ptr[c] .. cell_ptr[c + 1]) in cell order.  Every row must
 * belong to a cell.  Replaces SaddleBlockSolver.__init__'s structural part
 * (block.py:115-160). */
int mhm_global_create(int32_t device, int32_t size, int64_t nnz, const int32_t *csc_indptr,
                      const int32_t *csc_indices, int32_t ncells, const int64_t *cell_ptr, const int64_t *cell_idx,
                      mhm_global **out);

/* Device values in CSC order; builds the preconditioner and synchronises the
 * stream.  MHM_ESINGULAR (with *bad_cell) when a cell's assembled block is
 * singular or not finite.  Replaces the factorisation of the preconditioner
 * blocks (block.py:160-175) and factorize(matrix) (linalg/linear.py:467-537). */
int mhm_global_set_values(mhm_global *g, const double *d_csc_values, int32_t *bad_cell, void *stream);

/* x = S^-1 b for one device right-hand side: ||b - S x|| <= rtol ||b|| with the
 * residual in double-double (linalg/linear.py:132-152), at most maxiter
 * Arnoldi steps, restarted every `restart`.  Synchronous (one host round trip
 * per step).  MHM_ERESIDUAL when the criterion is not met.  Replaces
 * SaddleBlockSolver.solve (block.py:176-200). */
int mhm_global_solve(mhm_global *g, const double *d_rhs, double *d_x, double rtol, int32_t maxiter,
                     int32_t restart, int32_t *iterations, double *residual, void *stream);

/* Release.  Replaces SaddleBlockSolver.close (block.py:218-221). */
int mhm_global_destroy(mhm_global *g);

/* Kernel timing with HIP events recorded on the launch stream (no
 * counterpart in the reference, whose benchmarks time phases with
 * perf_counter wall clocks, benchmarks/execution_modes.py:227-258).
 * Kernel ids: MHM_TIMER_ASSEMBLE, _FACTOR, _SOLVE, _RESIDUAL, _CONDENSE.
 * mhm_timing_read synchronises on the recorded events. */
enum {
  MHM_TIMER_ASSEMBLE = 0,
  MHM_TIMER_FACTOR = 1,
  MHM_TIMER_SOLVE = 2,
  MHM_TIMER_RESIDUAL = 3,
  MHM_TIMER_CONDENSE = 4,
  MHM_TIMER_COUNT = 5
};
int mhm_timing_enable(int32_t enable);
int mhm_timing_reset(void);
int mhm_timing_read(int32_t kernel, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* MHM_HIP_H */
