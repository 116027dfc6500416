/*
 * bem_oracle.h -- CPU FP64 restatement of the reference Galerkin BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the reported CPU baseline ("port").  Only tests/, the smoke() entry
 * and bench.py's cpu_baseline leg may load it.  The product path never links
 * or calls it.
 *
 * Every function restates a reference function (GeCao/taichi-BEM); the
 * citation is given next to each declaration.  Conventions kept from the
 * reference:
 *   - element map x(r1,r2) = (1-r1) x1 + (r1-r2) x2 + r2 x3 on
 *     {0 <= r2 <= r1 <= 1}            (src/managers/BEM_manager.py:427-450)
 *   - P1 shape functions (1-r1, r1-r2, r2); P0 is the constant 1
 *                                      (src/managers/BEM_manager.py:404-425)
 *   - complex values are stored as interleaved (re, im) doubles, the layout of
 *     a Taichi Vector.field(2) and of torch complex128.
 */
#ifndef BEM_ORACLE_H
#define BEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Operator ids (the build's own numbering, shared with include/bem_hip.h). */
enum { BO_OP_V = 0, BO_OP_K = 1, BO_OP_KADJ = 2, BO_OP_W = 3 };
/* KernelType values, src/BEM_functions/utils/types.py:32-35 */
enum { BO_LAPLACE = 0, BO_HELMHOLTZ = 1, BO_HELMHOLTZ_TRANSMISSION = 2 };
/* PanelsRelation values, src/BEM_functions/utils/types.py:41-45 */
enum { BO_SEPARATE = 0, BO_COINCIDE = 1, BO_COMMON_VERTEX = 2, BO_COMMON_EDGE = 3 };
/* Quadrature variants.  0 = the reference's own rules, quirks included. */
enum {
    BO_FLAG_ROTATE_NEAR = 1,   /* move the shared vertex/edge to the slots the
                                  Sauter-Schwab rules expect */
    BO_FLAG_SYM_COINCIDE = 2   /* coincident rule evaluates f(y,x) with swapped
                                  bases instead of f(x,y) twice */
};

typedef struct {
    int32_t op;        /* BO_OP_* */
    int32_t kernel;    /* BO_LAPLACE / BO_HELMHOLTZ / BO_HELMHOLTZ_TRANSMISSION */
    double k;          /* wavenumber */
    double sqrt_n;     /* sqrt of the refractive index (transmission only) */
    int32_t q;         /* GaussQR */
    int32_t deg_test;  /* 0 = P0 (panel dofs), 1 = P1 (vertex dofs) */
    int32_t deg_trial;
    int32_t flags;     /* BO_FLAG_* */
} bo_op;

/* Gauss-Legendre nodes/weights mapped to [0,1] with halved weights, computed
 * by Newton iteration on P_q.  Supported q: 1..8 and 12, as in
 * src/BEM_functions/utils/functions.py:5-85.  Returns 0, or -1 for an
 * unsupported q (the reference raises NotImplementedError). */
int bo_gauss01(int q, double *pts, double *wts);

/* Panel areas and unit normals from (x2-x1) x (x3-x2); vertex areas as the
 * sum of A/3 over incident panels and vertex normals as the normalised sum of
 * panel normals.  src/managers/mesh_manager.py:114-134 */
void bo_geometry(int nv, int nf, const double *verts, const int32_t *tris,
                 double *area, double *normal, double *vert_area, double *vert_normal);

/* Topological relation of panels i and j.  src/managers/mesh_manager.py:538-573 */
int bo_relation(const int32_t *tris, int i, int j);

/* Relation counts over all ordered pairs (O(F^2), small meshes only). */
void bo_relation_counts(int nf, const int32_t *tris, int64_t counts[4]);

/* Galerkin integrals of one operator on the panel pair (i, j) for every basis
 * pair: out[(s * nt + t) * 2 + {0,1}], s over test bases (1 or 3) and t over
 * trial bases.  Sums integrate_on_two_panels over the GaussQR^4 tensor points:
 *   V  src/BEM_functions/single_layer/single_layer_3d.py:55-369
 *   K  src/BEM_functions/double_layer/double_layer_3d.py:56-374
 *   K' src/BEM_functions/adj_double_layer/adj_double_layer_3d.py:57-376
 *   W  src/BEM_functions/hypersingular_layer/hypersingular_layer_3d.py:56-435 */
int bo_pair(const bo_op *op, const double *verts, const int32_t *tris,
            const double *area, const double *normal, int i, int j, int rel,
            double *out);

/* Dense operator, all panels of one dof type (Dirichlet, Neumann and Full
 * boundary-value problems: identity local<->global maps).  out has
 * rows x cols complex entries (row-major); rows = nf (P0 test) or nv (P1 test).
 * out += integral, like the reference's `+=` into _Vmat/_Kmat/_Wmat.
 * Parallel over test panels with nthreads OpenMP threads (<=0: all).
 *   forward(): single_layer_3d.py:371-433, double_layer_3d.py:376-438,
 *   hypersingular_layer_3d.py:437-500 */
int bo_assemble(const bo_op *op, int nv, int nf, const double *verts,
                const int32_t *tris, const double *area, const double *normal,
                double *out, int nthreads);

/* Test panels [row_begin, row_end) only (CPU baseline samples).  For P0 test
 * out holds rows row_begin..row_end-1 at offset 0; for P1 test all nv rows. */
int bo_assemble_rows(const bo_op *op, int nv, int nf, const double *verts,
                     const int32_t *tris, const double *area, const double *normal,
                     int row_begin, int row_end, double *out, int nthreads);

/* Matrix-free y[test dof] += alpha * sum_j Op_pair(i,j) . xhat[trial dof],
 * xhat = conj(x) if conj_x else x.  apply_K_dot_D_boundary
 * (double_layer_3d.py:440-501), apply_V_dot_N_boundary (single_layer_3d.py:
 * 435-498), apply_K_dot_N_boundary (adj_double_layer_3d.py:443-506),
 * apply_W_dot_D_boundary (hypersingular_layer_3d.py:502-565). */
int bo_apply(const bo_op *op, int nv, int nf, const double *verts,
             const int32_t *tris, const double *area, const double *normal,
             const double *x, int conj_x, double alpha, double *y, int nthreads);

/* bo_apply restricted to test panels [row_begin, row_end) (CPU baseline samples). */
int bo_apply_rows(const bo_op *op, int nv, int nf, const double *verts,
                  const int32_t *tris, const double *area, const double *normal,
                  const double *x, int conj_x, double alpha, int row_begin, int row_end,
                  double *y, int nthreads);

/* Mass matrix with the single-panel Q^2 Duffy rule.  out[rows x cols]
 * complex, += alpha * M.  identity_layer_3d.py:46-146 and
 * rhs_constructor_3d.py:203-254. */
int bo_mass(int q, int deg_test, int deg_trial, int nv, int nf,
            const int32_t *tris, const double *area, double alpha, double *out);

/* rhs = s/2 M conj?(g) + K conj?(g) for the Dirichlet BVP (P0 Neumann
 * unknowns, P1 Dirichlet data): RHSConstructor3d.forward,
 * rhs_constructor_3d.py:374-386 with multiply_identity_to_vector (:256-311)
 * and multiply_M_to_vector (:361-372). s = +1 interior, -1 exterior. */
int bo_rhs_dirichlet(const bo_op *kop, int nv, int nf, const double *verts,
                     const int32_t *tris, const double *area, const double *normal,
                     const double *g, int conj_g, double s, double *rhs, int nthreads);

/* Number of OpenMP threads the library would use for nthreads <= 0. */
int bo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
