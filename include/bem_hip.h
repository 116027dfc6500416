/*
 * bem_hip.h -- C ABI of the MI355X Galerkin BEM hot path (libbem_hip.so).
 *
 * Drop-in for the boundary-integral-operator assembly and dense solve of
 * GeCao/taichi-BEM.  The reference has no FFI: its boundary is a Python class
 * API over Taichi fields.  Each entry point below names the reference
 * function (file:line) whose work it replaces; the Python layer in
 * taichi-bem_amd/ keeps the reference's class and method names and calls these
 * through ctypes.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (HBM), except where a comment
 *     says host.  The library never allocates or frees caller buffers; its own
 *     scratch lives in the context and is released by bem_destroy.
 *   - Matrices are row-major with leading dimension `ld` (elements).  Complex
 *     values are interleaved (re, im): the layout of a Taichi
 *     Vector.field(2) and of torch complex64/complex128.
 *   - Every call is ordered on `stream` (a hipStream_t, NULL = default
 *     stream); results are visible after a stream synchronisation.
 *   - Return value: BEM_OK (0) or an error code; bem_last_error() gives the
 *     message of the calling thread's last failure.  The library never exits
 *     the process (compare src/managers/BEM_manager.py:833-835).
 */
#ifndef BEM_HIP_H
#define BEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the Python layer maps them to exception types) ---- */
enum {
    BEM_OK = 0,
    BEM_ERR_INVALID = 1,      /* RuntimeError   (BEM_manager.py:41,50,193)   */
    BEM_ERR_UNSUPPORTED = 2,  /* NotImplementedError (utils/functions.py:80-83) */
    BEM_ERR_HIP = 3,          /* HIP runtime failure                           */
    BEM_ERR_SINGULAR = 4      /* exact zero pivot in the LU                     */
};

/* ---- operators (build's numbering) ---- */
enum { BEM_OP_V = 0, BEM_OP_K = 1, BEM_OP_KADJ = 2, BEM_OP_W = 3 };

/* ---- KernelType, src/BEM_functions/utils/types.py:32-35 ---- */
enum { BEM_LAPLACE = 0, BEM_HELMHOLTZ = 1, BEM_HELMHOLTZ_TRANSMISSION = 2 };

/* ---- PanelsRelation, src/BEM_functions/utils/types.py:41-45 ---- */
enum { BEM_SEPARATE = 0, BEM_COINCIDE = 1, BEM_COMMON_VERTEX = 2, BEM_COMMON_EDGE = 3 };

/* ---- element types of caller buffers ---- */
enum { BEM_F64 = 0, BEM_C128 = 1, BEM_F32 = 2, BEM_C64 = 3 };

/* ---- arithmetic of the regular (well-separated) pairs ---- */
enum {
    BEM_PREC_FP64 = 0,  /* everything in FP64 (the parity preset)            */
    BEM_PREC_FP32 = 1   /* regular pairs in FP32 with FP32 accumulation and
                           FP64 panel-centroid differences; near field FP64  */
};

/* ---- quadrature variants; 0 reproduces the reference rules exactly ---- */
enum {
    BEM_FLAG_ROTATE_NEAR = 1,  /* put the shared vertex/edge at the slots the
                                  Sauter-Schwab rules expect (the reference
                                  uses stored order, single_layer_3d.py:133-141) */
    BEM_FLAG_SYM_COINCIDE = 2, /* coincident rule: f(y,x) with swapped bases
                                  instead of f(x,y) twice (single_layer_3d.py:
                                  198-206) */
    BEM_FLAG_DK = 4,           /* bem_apply only: the product with dOp/dk, the
                                  derivative of every entry with respect to the
                                  wave number k (kap = k or k*sqrt_n); FP64.
                                  No reference counterpart (the reference has
                                  no derivatives) */
    BEM_FLAG_TRANSPOSE = 8     /* bem_apply only: y[trial dof] += alpha Op^T
                                  xhat (transpose, not conjugate transpose) of
                                  the same discrete operator, pair rules
                                  unchanged; FP64, not with BEM_FLAG_DK.  No
                                  reference counterpart; serves the adjoint's
                                  gradient with respect to the Dirichlet data,
                                  (1/2 M + K)^T mu, without forming K */
};

typedef struct bem_context bem_context;

/* Operator descriptor.  `kernel`, `k` and `sqrt_n` mean what the reference's
 * BEMManager._kernel_type/_k/sqrt_n mean (BEM_manager.py:31-62): the wave
 * number used in G is k for HELMHOLTZ and k*sqrt_n for
 * HELMHOLTZ_TRANSMISSION; the hypersingular k^2 term always uses
 * k^2 sqrt_n^2 (hypersingular_layer_3d.py:171). */
typedef struct {
    int32_t op;          /* BEM_OP_* */
    int32_t kernel;      /* BEM_LAPLACE / BEM_HELMHOLTZ / ..._TRANSMISSION */
    double k;
    double sqrt_n;
    int32_t deg_test;    /* 0 = P0 panel dofs, 1 = P1 vertex dofs */
    int32_t deg_trial;
    int32_t precision;   /* BEM_PREC_* */
    int32_t dtype;       /* element type of the output / vectors: BEM_* */
    int32_t flags;       /* BEM_FLAG_* */
    int32_t parts;       /* BEM_PART_* mask of the launches to run; 0 = all */
} bem_op_desc;

/* ---- parts of an assembly/apply call (to time them separately) ---- */
enum { BEM_PART_FAR = 1, BEM_PART_NEAR = 2 };

typedef struct {
    int32_t iterations;      /* refinement steps taken */
    int32_t converged;       /* 1 if ||r|| <= tol * ||b|| */
    double residual_norm0;   /* ||b - A x0|| (FP64) */
    double residual_norm;    /* final ||b - A x|| (FP64) */
    double b_norm;
} bem_refine_stats;

/* ---- lifetime -------------------------------------------------------- */

/* Creates a context on `device`.  Replaces the field allocations of
 * BEMManager.initialization, src/managers/BEM_manager.py:82-195. */
int bem_create(bem_context** ctx, int device);

/* Releases everything the context owns.  Replaces BEMManager.kill,
 * src/managers/BEM_manager.py:1075-1108 (and the layers' kill(), e.g.
 * single_layer_3d.py:42-52). */
int bem_destroy(bem_context* ctx);

/* Message of the last failure on the calling thread ("" if none). */
const char* bem_last_error(void);

/* Library ABI version. */
int bem_version(void);

/* ---- mesh and tables ------------------------------------------------- */

/* Uploads a triangle mesh (device pointers: vertices [nv*3] f64, panels
 * [nf*3] i32, 0-based) and builds on the device: panel areas and normals,
 * vertex areas and normals (MeshManager.initialization,
 * src/managers/mesh_manager.py:67-164, geometry at :114-139), the GaussQR
 * tables (BEM_manager.py:166-173, utils/functions.py:5-85) and the
 * near-field pair list of the reference relation classes
 * (mesh_manager.py:538-573).  gauss_qr must be 1..8 or 12, else
 * BEM_ERR_UNSUPPORTED. */
int bem_set_mesh(bem_context* ctx, const double* vertices, const int32_t* panels,
                 int32_t nv, int32_t nf, int32_t gauss_qr, void* stream);

/* Copies the device geometry out (any pointer may be NULL):
 * MeshManager.get_panel_areas/get_panel_normals/get_vert_areas/
 * get_vert_normals, mesh_manager.py:455-469. */
int bem_get_geometry(bem_context* ctx, double* area, double* normal,
                     double* vert_area, double* vert_normal, void* stream);

/* Host array counts[4]: ordered panel pairs per PanelsRelation. */
int bem_nearfield_counts(bem_context* ctx, int64_t counts[4]);

/* ---- operator assembly ---------------------------------------------- */

/* out[test dof, trial dof] += alpha * Op, dense (rows = nf for P0 test or nv
 * for P1 test; ld = row stride of out).  Replaces the forward() kernels:
 *   V  SingleLayer3d.forward         single_layer_3d.py:371-433
 *   K  DoubleLayer3d.forward         double_layer_3d.py:376-438
 *   K' AdjDoubleLayer3d.forward      adj_double_layer_3d.py:378-383 (a no-op
 *      in the reference, which uses K^T at BEM_manager.py:619-622; here K'
 *      can be assembled directly)
 *   W  HypersingularLayer3d.forward  hypersingular_layer_3d.py:437-500
 * together with the block placement of BEMManager.matA_add_M
 * (BEM_manager.py:608-628): pass `out` offset to the block and alpha = the
 * block's multiplier.  Real dtypes (BEM_F64/BEM_F32) take alpha_re only and
 * require the Laplace kernel. */
int bem_assemble(bem_context* ctx, const bem_op_desc* op, void* out, int64_t ld,
                 double alpha_re, double alpha_im, void* stream);

/* Interior/exterior Dirichlet system of the P0 Neumann unknowns with P1
 * Dirichlet data g, in one far-field pass: A += alpha V, and
 * rhs += alpha_k K ghat + alpha_m M ghat (ghat = conj(g) if conj_g; rhs and g
 * of dtype_rhs).  The reference builds the same pieces with
 * SingleLayer3d.forward + matA_add_M (BEM_manager.py:1006-1008) and
 * RHSConstructor3d.forward (rhs_constructor_3d.py:374-386: alpha_m = +-1/2,
 * alpha_k = 1); here V and K ghat share every point-pair evaluation.
 * `op` is the P0 x P0 single-layer descriptor (kernel, k, precision, flags;
 * its dtype is that of A). */
int bem_assemble_dirichlet(bem_context* ctx, const bem_op_desc* op, void* A, int64_t ld,
                           double alpha_re, double alpha_im, const void* g, int32_t conj_g,
                           void* rhs, int32_t dtype_rhs, double alpha_k, double alpha_m,
                           void* stream);

/* The three blocks of the transmission (Calderon) system of one wave number
 * in one pass: V <- V (P0 x P0, nf x nf), K <- K (P0 test x P1 trial,
 * nf x nv), W <- W (P1 x P1, nv x nv); the call zeroes the three outputs
 * first.  Replaces the forward() calls of BEMManager.matrix_layer_forward
 * (BEM_manager.py:829-844) when every panel carries both traces:
 * SingleLayer3d.forward (single_layer_3d.py:371-433),
 * DoubleLayer3d.forward (double_layer_3d.py:376-438) and
 * HypersingularLayer3d.forward (hypersingular_layer_3d.py:437-500), which
 * evaluate the kernel once per operator and ordered pair; here every
 * unordered regular panel pair is evaluated once for all three blocks and
 * both orders.  `op` is the P0 x P0 single-layer descriptor (kernel, k,
 * sqrt_n, precision, dtype of the outputs, flags); ldv >= nf, ldk >= nv and
 * ldw >= nv are the row strides in elements. */
int bem_assemble_vkw(bem_context* ctx, const bem_op_desc* op, void* V, int64_t ldv,
                     void* K, int64_t ldk, void* W, int64_t ldw, void* stream);

/* The k-derivative bilinear forms of the three Calderon blocks over every
 * panel pair, without forming a matrix: out (4 complex128 = 8 doubles, device
 * memory) = (aN^T dV/dk bN, aN^T dK/dk bD, bN^T dK/dk aD, aD^T dW/dk bD),
 * plain products (no conjugation).  aN, bN: complex128 on the panels (nf);
 * aD, bD: complex128 on the vertices (nv).  `op` is the P0 single-layer
 * descriptor of the wave number (Helmholtz kernels, FP64, complex128).  No
 * reference counterpart (the reference has no derivatives); it serves the
 * adjoint wave-number gradient of the transmission system, which needs
 * mu^T dA/dk x of the blocks that BEMManager.matA_add_M places
 * (BEM_manager.py:608-628).  Same result as four bem_apply calls with
 * BEM_FLAG_DK and dots, from one evaluation per unordered point pair. */
int bem_vkw_dk_forms(bem_context* ctx, const bem_op_desc* op, const void* aN, const void* bN,
                     const void* aD, const void* bD, double* out, void* stream);

/* Matrix-free y[test dof] += alpha * Op . xhat, xhat = conj(x) if conj_x.
 * Replaces apply_V_dot_N_boundary (single_layer_3d.py:435-498),
 * apply_K_dot_D_boundary (double_layer_3d.py:440-501),
 * apply_K_dot_N_boundary (adj_double_layer_3d.py:443-506) and
 * apply_W_dot_D_boundary (hypersingular_layer_3d.py:502-565); the reference
 * always conjugates complex data (conj_x = 1).  Sums run in a fixed order
 * (bitwise reproducible) except the far part of a BEM_FLAG_TRANSPOSE product,
 * whose test-panel chunks meet in atomics. */
int bem_apply(bem_context* ctx, const bem_op_desc* op, const void* x, int32_t conj_x,
              void* y, double alpha_re, double alpha_im, void* stream);

/* out[test dof, trial dof] += alpha * M (or out[trial, test] when transpose
 * = 1), M_{s,t} = int phi_s phi_t with the single-panel GaussQR^2 rule.
 * Replaces IdentityLayer3d.matA_add_global_Identity_matrix
 * (identity_layer_3d.py:148-196), which adds M at (N+i, D+v) and its
 * transpose at (D+v, N+i): call twice with offset pointers. */
int bem_mass_assemble(bem_context* ctx, int32_t deg_test, int32_t deg_trial, int32_t dtype,
                      void* out, int64_t ld, double alpha, int32_t transpose, void* stream);

/* y[test dof] += alpha * M . xhat.  Replaces
 * RHSConstructor3d.multiply_identity_to_vector (rhs_constructor_3d.py:256-359). */
int bem_mass_apply(bem_context* ctx, int32_t deg_test, int32_t deg_trial, int32_t dtype,
                   const void* x, int32_t conj_x, void* y, double alpha, void* stream);

/* ---- dense solve ----------------------------------------------------- */

/* In-place LU with partial (row) pivoting of the row-major n x n matrix A
 * (dtype BEM_F64/BEM_C128/BEM_F32/BEM_C64); ipiv (device, n int32) gets the
 * 0-based LAPACK-style row interchanges.  Trailing updates run on MFMA.
 * Together with bem_lu_solve it replaces torch.linalg.solve in
 * BEMManager.solve (BEM_manager.py:1032-1073, the call at :1056). */
int bem_lu_factor(bem_context* ctx, void* A, int64_t ld, int32_t n, int32_t dtype,
                  int32_t* ipiv, void* stream);

/* Solves A x = b in place of b (n entries of dtype) from bem_lu_factor. */
int bem_lu_solve(bem_context* ctx, const void* LU, int64_t ld, int32_t n, int32_t dtype,
                 const int32_t* ipiv, void* b, void* stream);

/* ---- options ---------------------------------------------------------- */
enum {
    BEM_OPT_C64_LU_GEMM = 1  /* arithmetic of the trailing updates of complex64
                                LU factorisations: 0 (default) = FP32 MFMA;
                                2 = the rank-512 updates of single
                                factorisations on the fp16 MFMA, each operand
                                row (column) scaled by a power of two and
                                split into two fp16 parts, three part
                                products per real product, FP32 accumulation
                                (update error about 3e-7 of sum |a||b|; for
                                factors that precondition FP64 refinement).
                                Batched factorisations keep the FP32 MFMA */,
    BEM_OPT_LU_LOOKAHEAD = 2 /* bem_lu_factor outside graph capture: 0 = one
                                stream; N > 0 (default 32) = the panels of the
                                next outer block (256 columns, 512 for FP32
                                and complex64 factors) run on N compute units
                                while the rest of the trailing update runs on
                                the others (two internal streams, default
                                priority, ordered by events against the
                                caller's stream); -1 = the same two streams
                                without compute-unit masks.  Factorisations
                                with block-local pivoting (BEM_OPT_LU_PIVOTING
                                = 2) always use unmasked streams when
                                look-ahead is on */,
    BEM_OPT_LU_PIVOTING = 3  /* 0 (default) = partial pivoting with one launch
                                per column, the pivot rows LAPACK's getrf
                                chooses; 1 = tournament pivoting per 64-column
                                panel (pivot rows chosen by a reduction tree
                                of LDS-resident eliminations, about ten
                                launches per panel): fewer synchronisations
                                per panel, but slower than the persistent
                                partial-pivoting panels of single
                                factorisations, and its element growth can be
                                larger, so use it for batches of
                                well-conditioned systems only; 2 =
                                block-local pivoting: rows interchanged only
                                inside each panel's 64 x 64 diagonal block
                                (two launches per panel, no waits between
                                workgroups), for factors that precondition
                                bem_refine_solve against the original matrix
                                (the mixed-precision solves); the ipiv it
                                returns is valid for every solve entry point */,
    BEM_OPT_LU_GEMM_TILE = 4 /* 1 (default) = trailing updates of at least
                                256 rows use 128-row MFMA tiles: 128 x 64 for
                                complex128, 128 x 64 with three real products
                                per complex product (Gauss) for complex64,
                                128 x 128 for real types; 2 = 128 x 128 with
                                four real products for the complex types too;
                                0 = 64 x 64 tiles for every update */,
    BEM_OPT_LU_PANEL = 5     /* partial-pivoting panels of bem_lu_factor:
                                1 (default) = one persistent launch per
                                64-column panel (one workgroup per compute
                                unit of the panel's stream, grid barriers,
                                8-column sub-panels in LDS) where the rows
                                fit; N > 1 = the same with at most N
                                workgroups; 0 = one launch per column */,
    BEM_OPT_WAIT_LIMIT_US = 6 /* bound, in microseconds, of every wait of a
                                kernel on other workgroups (the persistent
                                panel's grid barriers, the triangular sweeps'
                                ready flags): 0..2,000,000, default 2,000,000.
                                A wait that runs out makes bem_lu_factor
                                return BEM_ERR_HIP and a solve leave NaN, and
                                every kernel still ends.  Lowering it is a
                                diagnostic for those paths */
};

/* Sets a context option.  No reference counterpart. */
int bem_set_option(bem_context* ctx, int32_t option, int32_t value);

/* ---- batches of uneven sizes (wave-number sweeps, several meshes) ---- */

/* One problem of a combined-field batch. */
typedef struct {
    bem_context* ctx;  /* context holding this problem's mesh (bem_set_mesh) */
    double k;          /* wave number */
    double eta;        /* single-layer coupling (Brakhage-Werner: eta = k) */
    void* A;           /* device, nv x nv of the batch dtype, row stride ld */
    int64_t ld;
} bem_cf_problem;

/* Launch shape of the persistent far-field kernel (chosen by an autotuner;
 * zeros select the defaults). */
typedef struct {
    int32_t test_panels_per_block;  /* 1, 2, 4, 8 or 16 (default 2) */
    int32_t blocks_per_cu;          /* grid = blocks_per_cu x compute units (default 4) */
} bem_tuning;

/* A_q += 1/2 M + K - i eta_q V (P1 x P1) for every problem of the batch: the
 * exterior Dirichlet (sound-soft) combined-field operator, built from the
 * reference's K and V blocks (DoubleLayer3d.forward, double_layer_3d.py:
 * 376-438; SingleLayer3d.forward, single_layer_3d.py:371-433; mass,
 * identity_layer_3d.py:117-146); the reference has no combined-field
 * operator of its own.  Meshes and wave numbers may differ per problem; the
 * whole batch runs in three persistent launches (regular pairs, near-field
 * pairs, mass) over one work list.  dtype BEM_C64 or BEM_C128; precision
 * BEM_PREC_FP32 evaluates the regular pairs in FP32 (near field FP64).  The
 * batch context `ctx` owns the work-list table (uploaded, with a stream
 * synchronisation, only when it changes: repeated identical calls may be
 * captured in a HIP graph). */
int bem_assemble_combined_batched(bem_context* ctx, int32_t count, const bem_cf_problem* problems,
                                  int32_t dtype, int32_t precision, int32_t flags,
                                  const bem_tuning* tuning, void* stream);

/* Factors `count` matrices in one launch sequence: A[q] (device, n[q] x
 * n[q], row stride ld[q], all of `dtype`) in place, pivots into ipiv[q]
 * (device) and info[q] (device int32 array of `count`: 0, or 1 + the first
 * exactly-zero pivot column).  Problems of different n share every panel
 * step.  The host arrays are read during the call.  The first call with a
 * given set of arguments uploads a problem table and synchronises the
 * stream; a repeated identical call only launches kernels and may be
 * captured in a HIP graph.  Batched form of bem_lu_factor (the reference
 * solves one system per call with torch.linalg.solve, BEM_manager.py:1056,
 * inside its Python wave-number loop, demo/demo_3d_plot_Sio.py:67-76). */
int bem_lu_factor_batched(bem_context* ctx, int32_t count, void* const* A, const int64_t* ld,
                          const int32_t* n, int32_t dtype, int32_t* const* ipiv, int32_t* info,
                          void* stream);

/* Solves A_q x_q = b_q for the factors of bem_lu_factor_batched (LU[q],
 * ipiv[q], dtype), then runs refine_iters steps of x += LU^-1 (b - A x)
 * with residuals accumulated in FP64 against A[q] (same dtype as the LU).
 * b[q], x[q]: complex128 device vectors (float64 for real dtypes).  resid
 * (device, count doubles, may be NULL) gets ||b - A x|| / ||b||.  A/lda may
 * be NULL when refine_iters = 0 and resid = NULL.  Capturable like
 * bem_lu_factor_batched. */
int bem_batch_solve(bem_context* ctx, int32_t count, const void* const* LU, const int64_t* ld,
                    const int32_t* n, int32_t dtype, const int32_t* const* ipiv,
                    const void* const* A, const int64_t* lda, const void* const* b, void* const* x,
                    int32_t refine_iters, double* resid, void* stream);

/* 1-norm condition estimates: out[2q] = ||A_q||_1 (exact) and out[2q+1] = a
 * lower-bound estimate of ||A_q^-1||_1 from `iters` Hager-Higham steps on
 * the factors (solves with A and A^H) plus the alternating-sign test
 * vector.  out: device, 2 * count doubles.  The reference's norm analysis
 * (BEMManager.get_mat_A1_inv_norm, BEM_manager.py:874-936) takes a dense
 * SVD per wave number; this is the factor-based estimate for the sweep. */
int bem_batch_condest(bem_context* ctx, int32_t count, const void* const* LU, const int64_t* ld,
                      const int32_t* n, int32_t dtype, const int32_t* const* ipiv,
                      const void* const* A, const int64_t* lda, int32_t iters, double* out,
                      void* stream);

/* Solves A^T x = b (transpose, not conjugate transpose) in place of b from
 * the same bem_lu_factor output: the adjoint solve of the gradient below.
 * No reference counterpart (the reference has no derivatives). */
int bem_lu_solve_transpose(bem_context* ctx, const void* LU, int64_t ld, int32_t n, int32_t dtype,
                           const int32_t* ipiv, void* b, void* stream);

/* bem_lu_solve without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory or synchronise inside the call
 * (HIP graph capture, tight serving loops).  Same arguments and result as
 * bem_lu_solve (bitwise); the pivot permutation is built on the device, and
 * the kernels' scratch is a fixed per-device array of the library (about
 * 2.3 MiB) set up by bem_create, so n <= 65536 (larger n: BEM_ERR_INVALID).
 * Every call on one device shares that array: calls that overlap on one
 * device (several streams or host threads) race; consecutive calls are
 * independent.  Outside stream capture a sweep's timed-out wait is reported
 * as BEM_ERR_HIP (one stream synchronisation); under capture the NaN the
 * sweep leaves is the only sign. */
int bem_lu_solve_noalloc(bem_context* ctx, const void* LU, int64_t ld, int32_t n, int32_t dtype,
                         const int32_t* ipiv, void* b, void* stream);

/* r = b - A x in FP64/complex128 (A of dtype BEM_F64/BEM_C128; x, b, r of
 * the matching FP64 type). */
int bem_residual(bem_context* ctx, const void* A, int64_t ld, int32_t n, int32_t dtype,
                 const void* x, const void* b, void* r, void* stream);

/* Mixed-precision solve: x (FP64 type) from the LU of A (dtype_lu, usually
 * BEM_F32/BEM_C64) with iterative refinement whose residuals b - A x are
 * accumulated in FP64 against the matrix A (dtype_a: FP64 or FP32 storage,
 * same real/complex kind as the LU).  Stops when ||b - A x|| <= tol ||b|| or
 * after max_iter steps.  stats (host) may be NULL.  Synchronises the stream. */
int bem_refine_solve(bem_context* ctx, const void* A, int64_t lda, int32_t dtype_a,
                     const void* LU, int64_t ldlu, int32_t n, int32_t dtype_lu,
                     const int32_t* ipiv, const void* b, void* x, int32_t max_iter,
                     double tol, bem_refine_stats* stats, void* stream);

/* ---- derivatives (adjoint) ----------------------------------------- */

/* Adjoint contraction of an operator's derivative, without forming it:
 *   grad_vertices[3 v + d] += Re( alpha sum_{s,t} lam_s u_t dOp_st/dX_{v,d} )
 *   *grad_k                += Re( alpha sum_{s,t} lam_s u_t dOp_st/dk )
 * over every regular and near-field pair, re-evaluating the quadrature of
 * the same discrete operator bem_assemble builds (flags included).  op: V
 * (P0 x P0) or K (P0 test x P1 trial), FP64, dtype BEM_C128 (lam, u complex128
 * of nf and nf|nv entries); grad_vertices [nv*3] f64; grad_k (one f64) may be
 * NULL and is untouched for Laplace.  Every sum runs in a fixed order
 * (per-slot partials, no atomics): the result is bitwise reproducible.  The
 * partials live in a per-mesh buffer that grows to about 1 GiB at most for
 * the trial side (larger meshes take longer test chunks) plus 9 doubles per
 * (256-panel trial block, panel).  No reference counterpart: it differentiates
 * SingleLayer3d/DoubleLayer3d (single_layer_3d.py:155-177,
 * double_layer_3d.py:158-180) for the shape gradient of the Dirichlet
 * problem of BEMManager.run (BEM_manager.py:990-1011). */
int bem_shape_gradient(bem_context* ctx, const bem_op_desc* op, const void* lam, const void* u,
                       double alpha_re, double alpha_im, double* grad_vertices, double* grad_k,
                       void* stream);

/* bem_shape_gradient for K (P0 test x P1 trial) that also adds
 *   kt[v] += (K^T lam)_v   (complex128 [nv], transpose, no conjugation)
 * from the same pair evaluations: the regular pairs inside the contraction,
 * the near pairs by the transposed near-field product.  kt = NULL is
 * bem_shape_gradient.  No reference counterpart: the adjoint of the Dirichlet
 * problem (BEM_manager.py:990-1011) needs (1/2 M + K)^T mu for the gradient
 * with respect to the boundary data g, next to mu^T dK/dX g. */
int bem_shape_gradient_kt(bem_context* ctx, const bem_op_desc* op, const void* lam, const void* u,
                          double alpha_re, double alpha_im, double* grad_vertices, double* grad_k,
                          void* kt, void* stream);

/* grad_vertices[3 v + d] += sum_i coef_i dA_i/dX_{v,d} (panel areas A_i of
 * mesh_manager.py:114-124; coef f64 [nf]).  The derivative of the P0 mass
 * and of area-weighted misfits. */
int bem_area_gradient(bem_context* ctx, const double* coef, double* grad_vertices, void* stream);

/* dst = (dtype_dst) src for a rows x cols block (precision conversion of a
 * matrix or vector; complex <-> complex and real <-> real only). */
int bem_convert(const void* src, int32_t dtype_src, int64_t ld_src, void* dst,
                int32_t dtype_dst, int64_t ld_dst, int32_t rows, int32_t cols, void* stream);

/* dst[r, c] += alpha * src[r, c] (or src[c, r] when transpose = 1) for a
 * rows x cols block of dst; src and dst of the same dtype.  Replaces the
 * block placement loops of BEMManager.matA_add_M (BEM_manager.py:608-628),
 * including the -K^T block at :619-622. */
int bem_matrix_axpy(const void* src, int64_t ld_src, void* dst, int64_t ld_dst, int32_t dtype,
                    int32_t rows, int32_t cols, double alpha_re, double alpha_im, int32_t transpose,
                    void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BEM_HIP_H */
