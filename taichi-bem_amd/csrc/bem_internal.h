// bem_internal.h -- context layout and internal launchers of libbem_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/bem_hip.h"

namespace bem {

// One near-field pair: test panel i, trial panel j, relation, and the vertex
// slot order used for each panel (2 bits per slot: px = s0 | s1<<2 | s2<<4,
// py the same at bit 8).  Identity order reproduces the reference.
struct NearPair {
    int32_t i, j, rel, perm;
};

// Point of a near-field rule: parametric coordinates on both panels and the
// product of the four Gauss weights with the region Jacobian.
struct NearPt {
    double r1x, r2x, r1y, r2y, w;
};

constexpr int PANEL_FRAME = 20;  // doubles per panel of Mesh::frame (18 used)

struct Mesh {
    int nv = 0, nf = 0, q = 0, q2 = 0;
    double* verts = nullptr;      // [nv*3]
    int32_t* tris = nullptr;      // [nf*3]
    double* area = nullptr;       // [nf]
    double* normal = nullptr;     // [nf*3]
    double* vert_area = nullptr;  // [nv]
    double* vert_normal = nullptr;// [nv*3]
    double* curl = nullptr;       // [nf][3 slots][3]  n x grad(phi_s)
    // Regular-rule points of every panel (r1, r2) = (xi, xi*eta), weight
    // w_xi w_eta xi (2A).  FP64 absolute coordinates:
    double* pu = nullptr;         // AoS [nf][q2][4] (x, y, z, w): uniform (scalar) loads
    double* pl = nullptr;         // SoA [q2][4][nf]: per-lane coalesced loads
    // FP32 tables: offsets from the FP64 panel centroid.
    double* cen = nullptr;        // [nf*3]
    float* pu32 = nullptr;        // [nf][q2][4]
    float* pl32 = nullptr;        // [q2][4][nf]
    // FP64 local frame of every panel (the far field's moment form):
    // C (centroid), U = x1 - C, E1 = x2 - x1, E2 = x3 - x2, n, D = -2 U,
    // stride PANEL_FRAME; the points are x(r1, r2) - C = U + r1 E1 + r2 E2.
    double* frame = nullptr;      // [nf][PANEL_FRAME]
    // per regular point: S = |x - C|^2, -2 r1, -2 r2, w  (uniform loads)
    double* pq = nullptr;         // [nf][q2][4]
    float* pq32 = nullptr;        // [nf][q2][4] the same in FP32 (the combined-field sweep's far field)
    double* phi = nullptr;        // [q2][3] P1 shape values at the regular points
    float* phi32 = nullptr;       // [q2][3]
    double* mass_w = nullptr;     // [q2] w_xi w_eta xi (single-panel rule, no area)
    double* r12w = nullptr;       // [q2][3] (r1, r2, w_xi w_eta xi) of the regular rule
    // Near field
    NearPair* near = nullptr;     // sorted by i then j
    int32_t* near_row = nullptr;  // CSR row pointer over test panels [nf+1]
    int64_t n_near = 0;
    int64_t near_counts[4] = {0, 0, 0, 0};
    int near_flags = -1;          // flags the near list was built for
    NearPt* rule = nullptr;       // concatenated near rules
    int32_t rule_off[4] = {0, 0, 0, 0};  // offset per relation (index 0 unused)
    int32_t rule_len[4] = {0, 0, 0, 0};
    std::vector<int32_t> h_tris;  // host copy for near-list construction
    // vertex -> panel corners, CSR in panel order: vp_el[e] = 3 * panel + slot
    // for e in [vp_row[v], vp_row[v + 1]); the fixed order of every per-vertex
    // sum of panel contributions
    int32_t* vp_row = nullptr;    // [nv+1]
    int32_t* vp_el = nullptr;     // [3 nf]
    int32_t* near_tr = nullptr;   // [n_near] index of the mirrored pair (j, i) of near pair (i, j)
    double* near_buf = nullptr;   // scratch for matrix-free near contributions
    double* panel_buf = nullptr;  // [nf][3][2] per-panel partials of matrix-free products into vertex rows
    int64_t near_buf_len = 0;
    double* kpart = nullptr;      // per (test panel, trial block) partial sums of K conj(g), then
    size_t kpart_bytes = 0;       // per (row slot, trial panel) partials of the mirrored terms
    double* adj = nullptr;        // adjoint contractions: per-slot partials of the vertex gradient
    size_t adj_bytes = 0;
};

// Slots of Context::counters.
constexpr int CTR_TRI_FAIL = 0;   // a triangular sweep gave up a wait (sticky within one solve call)

struct Context {
    int device = 0;
    Mesh mesh;
    // LU / solver scratch
    void* work = nullptr;
    size_t work_bytes = 0;
    int* counters = nullptr;      // small per-context flags (CTR_* below)
    size_t counters_len = 0;
    void* tri_sync = nullptr;     // ticket and ready words of single-problem triangular sweeps
    size_t tri_sync_words = 0;
    void* lu_scratch = nullptr;   // LU counters, pivot partials, S_L (per problem of a batch)
    size_t lu_scratch_bytes = 0;
    std::vector<char> lu_table;   // host copy of the batch table last uploaded
    void* lu_table_dev = nullptr; // its device address
    void* batch_ws[2] = {nullptr, nullptr};  // batched solves / condition estimates:
    size_t batch_bytes[2] = {0, 0};           // per-problem vectors + tables
    std::vector<char> batch_tables[2];        // host copies of the tables last uploaded
    void* cf_ws = nullptr;        // combined-field batch table
    size_t cf_bytes = 0;
    std::vector<char> cf_table;
    int num_cu = 0;
    int c64_gemm_h16 = 0;         // c64 LU rank-256 updates: 0 FP32 MFMA, 1 split-fp16 planes on the fp16 MFMA
    void* lu_h16 = nullptr;       // split-fp16 planes of the c64 rank-256 updates
    size_t lu_h16_bytes = 0;
    int lu_lookahead_cus = 32;    // LU look-ahead: 0 off, > 0 panel-stream CUs, -1 unmasked streams
    int lu_pivoting = 0;          // LU panels: 0 partial pivoting (LAPACK rows), 1 tournament
    int lu_panel = 1;             // LU partial-pivoting panels: 1 persistent launch where it fits, 0 per column
    // bound of the kernels' waits on other workgroups (grid barriers, ready flags),
    // wall-clock ticks of 100 MHz: 2 s by default
    long long wait_ticks = 200000000;
    int lu_gemm_tile = 1;         // LU trailing updates: 1 128-row tiles where they fit, 0 64 x 64 only
    int lu_streams_cus = 0;       // the setting the streams were created for
    hipStream_t lu_stream[2] = {nullptr, nullptr};
    hipStream_t lu_stream_free[2] = {nullptr, nullptr};  // look-ahead without CU masks (block-local pivoting)
    hipEvent_t lu_event[3] = {nullptr, nullptr, nullptr};
    hipEvent_t lu_pevent[8] = {};  // per-panel events of an outer block (look-ahead)
};

// error helpers (bem_capi.hip)
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// geometry / tables (bem_geometry.hip)
int mesh_build(Context* ctx, const double* d_verts, const int32_t* d_tris, int nv, int nf, int q,
               hipStream_t s);
int mesh_build_near(Context* ctx, int flags, hipStream_t s);
void mesh_free(Mesh* m);

// assembly (bem_assemble.hip)
int assemble_op(Context* ctx, const bem_op_desc* d, void* out, int64_t ld, double ar, double ai,
                hipStream_t s);
int apply_op(Context* ctx, const bem_op_desc* d, const void* x, int conj_x, void* y, double ar,
             double ai, hipStream_t s);
int mass_assemble(Context* ctx, int deg_test, int deg_trial, int dtype, void* out, int64_t ld,
                  double alpha, int transpose, hipStream_t s);
int assemble_dirichlet(Context* ctx, const bem_op_desc* d, void* A, int64_t ld, double ar, double ai,
                       const void* g, int conj_g, void* rhs, int dtype_rhs, double alpha_k, double alpha_m,
                       hipStream_t s);
int assemble_vkw(Context* ctx, const bem_op_desc* d, void* V, int64_t ldv, void* K, int64_t ldk, void* W,
                 int64_t ldw, hipStream_t s);
int vkw_dk_forms(Context* ctx, const bem_op_desc* d, const void* aN, const void* bN, const void* aD,
                 const void* bD, double* out, hipStream_t s);
// one problem of a combined-field batch (bem_cf_problem with the context resolved)
struct CfInput {
    Context* ctx;
    double k, eta;
    void* A;
    int64_t ld;
};
int assemble_combined_batched(Context* ctx, int count, const CfInput* probs, int dtype, int precision,
                              int flags, const bem_tuning* tune, hipStream_t s);
int mass_apply(Context* ctx, int deg_test, int deg_trial, int dtype, const void* x, int conj_x,
               void* y, double alpha, hipStream_t s);

// adjoint contractions (bem_adjoint.hip)
int shape_gradient(Context* ctx, const bem_op_desc* d, const void* lam, const void* u, double ar, double ai,
                   double* grad, double* grad_k, void* kt, hipStream_t s);
int area_gradient(Context* ctx, const double* coef, double* grad, hipStream_t s);

// dense linear algebra (bem_linalg.hip)
int lu_factor(Context* ctx, void* A, int64_t ld, int n, int dtype, int32_t* ipiv, hipStream_t s);
int lu_factor_batched(Context* ctx, int count, void* const* A, const int64_t* ld, const int32_t* n, int dtype,
                      int32_t* const* ipiv, int32_t* info, hipStream_t s);
int batch_solve(Context* ctx, int count, const void* const* LU, const int64_t* ld, const int32_t* n, int dtype,
                const int32_t* const* ipiv, const void* const* A, const int64_t* lda, const void* const* b,
                void* const* x, int refine_iters, double* resid, hipStream_t s);
int batch_condest(Context* ctx, int count, const void* const* LU, const int64_t* ld, const int32_t* n, int dtype,
                  const int32_t* const* ipiv, const void* const* A, const int64_t* lda, int iters, double* out,
                  hipStream_t s);
int lu_solve(Context* ctx, const void* LU, int64_t ld, int n, int dtype, const int32_t* ipiv,
             void* b, hipStream_t s);
int lu_solve_transpose(Context* ctx, const void* LU, int64_t ld, int n, int dtype, const int32_t* ipiv,
                       void* b, hipStream_t s);
int noalloc_setup(int device);
int lu_solve_noalloc(Context* ctx, const void* LU, int64_t ld, int n, int dtype, const int32_t* ipiv, void* b,
                     hipStream_t s);
int gemv_residual(Context* ctx, const void* A, int64_t ld, int n, int dtype_a, const void* x,
                  const void* b, void* r, hipStream_t s);
int refine_solve(Context* ctx, const void* A, int64_t lda, int dtype_a, const void* LU, int64_t ldlu,
                 int n, int dtype_lu, const int32_t* ipiv, const void* b, void* x, int max_iter,
                 double tol, bem_refine_stats* st, hipStream_t s);
int convert_matrix(const void* src, int dtype_src, int64_t lds, void* dst, int dtype_dst,
                   int64_t ldd, int rows, int cols, hipStream_t s);

int matrix_axpy(const void* src, int64_t lds, void* dst, int64_t ldd, int dtype, int rows, int cols,
                double ar, double ai, int transpose, hipStream_t s);

size_t dtype_size(int dtype);
bool dtype_complex(int dtype);

}  // namespace bem
