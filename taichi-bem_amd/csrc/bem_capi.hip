// bem_capi.hip -- the extern "C" entry points declared in include/bem_hip.h.
#include <string>
#include <vector>

#include "bem_internal.h"

struct bem_context {
    bem::Context impl;
};

namespace bem {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return BEM_OK;
    return set_error(BEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace bem

using bem::check_hip;
using bem::set_error;

namespace {

// Runs `fn` with the context's device current; restores the caller's device.
template <typename F>
int with_device(bem_context* ctx, F fn) {
    if (!ctx) return set_error(BEM_ERR_INVALID, "null context");
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) prev = ctx->impl.device;
    if (prev != ctx->impl.device) {
        int rc = check_hip(hipSetDevice(ctx->impl.device), "hipSetDevice");
        if (rc) return rc;
    }
    int rc = fn();
    if (prev != ctx->impl.device) hipSetDevice(prev);
    return rc;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int bem_version(void) { return 1; }

const char* bem_last_error(void) { return bem::g_last_error.c_str(); }

int bem_create(bem_context** out, int device) {
    if (!out) return set_error(BEM_ERR_INVALID, "null output handle");
    *out = nullptr;
    int count = 0;
    int rc = check_hip(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (rc) return rc;
    if (device < 0 || device >= count) return set_error(BEM_ERR_INVALID, "device index out of range");
    bem_context* ctx = new bem_context();
    ctx->impl.device = device;
    rc = with_device(ctx, [&]() -> int {
        int r = check_hip(hipMalloc((void**)&ctx->impl.counters, 64 * sizeof(int)), "counters");
        if (r) return r;
        return bem::noalloc_setup(device);  // this device's address of the allocation-free scratch
    });
    if (rc) {
        if (ctx->impl.counters) with_device(ctx, [&]() -> int { return check_hip(hipFree(ctx->impl.counters), "free"); });
        delete ctx;
        return rc;
    }
    ctx->impl.counters_len = 64;
    *out = ctx;
    return BEM_OK;
}

int bem_destroy(bem_context* ctx) {
    if (!ctx) return BEM_OK;
    with_device(ctx, [&]() -> int {
        hipDeviceSynchronize();
        bem::mesh_free(&ctx->impl.mesh);
        if (ctx->impl.work) hipFree(ctx->impl.work);
        if (ctx->impl.counters) hipFree(ctx->impl.counters);
        if (ctx->impl.lu_scratch) hipFree(ctx->impl.lu_scratch);
        if (ctx->impl.lu_h16) hipFree(ctx->impl.lu_h16);
        if (ctx->impl.tri_sync) hipFree(ctx->impl.tri_sync);
        for (void* w : ctx->impl.batch_ws)
            if (w) hipFree(w);
        for (hipStream_t st : ctx->impl.lu_stream)
            if (st) hipStreamDestroy(st);
        for (hipStream_t st : ctx->impl.lu_stream_free)
            if (st) hipStreamDestroy(st);
        for (hipEvent_t e : ctx->impl.lu_event)
            if (e) hipEventDestroy(e);
        for (hipEvent_t e : ctx->impl.lu_pevent)
            if (e) hipEventDestroy(e);
        if (ctx->impl.cf_ws) hipFree(ctx->impl.cf_ws);
        return BEM_OK;
    });
    delete ctx;
    return BEM_OK;
}

int bem_set_mesh(bem_context* ctx, const double* vertices, const int32_t* panels, int32_t nv, int32_t nf,
                 int32_t gauss_qr, void* stream) {
    if (!vertices || !panels) return set_error(BEM_ERR_INVALID, "null mesh arrays");
    return with_device(ctx, [&]() -> int {
        return bem::mesh_build(&ctx->impl, vertices, panels, nv, nf, gauss_qr, as_stream(stream));
    });
}

int bem_get_geometry(bem_context* ctx, double* area, double* normal, double* vert_area, double* vert_normal,
                     void* stream) {
    return with_device(ctx, [&]() -> int {
        const bem::Mesh& m = ctx->impl.mesh;
        if (!m.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
        hipStream_t s = as_stream(stream);
        int rc = 0;
        if (area) rc |= check_hip(hipMemcpyAsync(area, m.area, sizeof(double) * m.nf, hipMemcpyDeviceToDevice, s), "area");
        if (normal) rc |= check_hip(hipMemcpyAsync(normal, m.normal, sizeof(double) * m.nf * 3, hipMemcpyDeviceToDevice, s), "normal");
        if (vert_area) rc |= check_hip(hipMemcpyAsync(vert_area, m.vert_area, sizeof(double) * m.nv, hipMemcpyDeviceToDevice, s), "vert_area");
        if (vert_normal) rc |= check_hip(hipMemcpyAsync(vert_normal, m.vert_normal, sizeof(double) * m.nv * 3, hipMemcpyDeviceToDevice, s), "vert_normal");
        return rc ? BEM_ERR_HIP : BEM_OK;
    });
}

int bem_nearfield_counts(bem_context* ctx, int64_t counts[4]) {
    if (!counts) return set_error(BEM_ERR_INVALID, "null counts");
    return with_device(ctx, [&]() -> int {
        const bem::Mesh& m = ctx->impl.mesh;
        if (!m.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
        for (int i = 0; i < 4; ++i) counts[i] = m.near_counts[i];
        return BEM_OK;
    });
}

int bem_assemble(bem_context* ctx, const bem_op_desc* op, void* out, int64_t ld, double alpha_re,
                 double alpha_im, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::assemble_op(&ctx->impl, op, out, ld, alpha_re, alpha_im, as_stream(stream));
    });
}

int bem_apply(bem_context* ctx, const bem_op_desc* op, const void* x, int32_t conj_x, void* y,
              double alpha_re, double alpha_im, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::apply_op(&ctx->impl, op, x, conj_x, y, alpha_re, alpha_im, as_stream(stream));
    });
}

int bem_assemble_dirichlet(bem_context* ctx, const bem_op_desc* op, void* A, int64_t ld, double alpha_re,
                           double alpha_im, const void* g, int32_t conj_g, void* rhs, int32_t dtype_rhs,
                           double alpha_k, double alpha_m, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::assemble_dirichlet(&ctx->impl, op, A, ld, alpha_re, alpha_im, g, conj_g, rhs, dtype_rhs,
                                       alpha_k, alpha_m, as_stream(stream));
    });
}

int bem_assemble_vkw(bem_context* ctx, const bem_op_desc* op, void* V, int64_t ldv, void* K, int64_t ldk, void* W,
                     int64_t ldw, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::assemble_vkw(&ctx->impl, op, V, ldv, K, ldk, W, ldw, as_stream(stream));
    });
}

int bem_vkw_dk_forms(bem_context* ctx, const bem_op_desc* op, const void* aN, const void* bN, const void* aD,
                     const void* bD, double* out, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::vkw_dk_forms(&ctx->impl, op, aN, bN, aD, bD, out, as_stream(stream));
    });
}

int bem_mass_assemble(bem_context* ctx, int32_t deg_test, int32_t deg_trial, int32_t dtype, void* out,
                      int64_t ld, double alpha, int32_t transpose, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::mass_assemble(&ctx->impl, deg_test, deg_trial, dtype, out, ld, alpha, transpose,
                                  as_stream(stream));
    });
}

int bem_mass_apply(bem_context* ctx, int32_t deg_test, int32_t deg_trial, int32_t dtype, const void* x,
                   int32_t conj_x, void* y, double alpha, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::mass_apply(&ctx->impl, deg_test, deg_trial, dtype, x, conj_x, y, alpha, as_stream(stream));
    });
}

int bem_lu_factor(bem_context* ctx, void* A, int64_t ld, int32_t n, int32_t dtype, int32_t* ipiv,
                  void* stream) {
    return with_device(ctx, [&]() -> int { return bem::lu_factor(&ctx->impl, A, ld, n, dtype, ipiv, as_stream(stream)); });
}

int bem_lu_solve(bem_context* ctx, const void* LU, int64_t ld, int32_t n, int32_t dtype, const int32_t* ipiv,
                 void* b, void* stream) {
    return with_device(ctx, [&]() -> int { return bem::lu_solve(&ctx->impl, LU, ld, n, dtype, ipiv, b, as_stream(stream)); });
}

int bem_assemble_combined_batched(bem_context* ctx, int32_t count, const bem_cf_problem* problems, int32_t dtype,
                                  int32_t precision, int32_t flags, const bem_tuning* tuning, void* stream) {
    return with_device(ctx, [&]() -> int {
        if (count > 0 && !problems) return set_error(BEM_ERR_INVALID, "null problem array");
        std::vector<bem::CfInput> in(count > 0 ? count : 0);
        for (int q = 0; q < count; ++q) {
            if (!problems[q].ctx) return set_error(BEM_ERR_INVALID, "null context in the batch");
            if (problems[q].ctx->impl.device != ctx->impl.device)
                return set_error(BEM_ERR_INVALID, "every problem of a batch must be on the batch context's device");
            in[q] = {&problems[q].ctx->impl, problems[q].k, problems[q].eta, problems[q].A, problems[q].ld};
        }
        return bem::assemble_combined_batched(&ctx->impl, count, in.data(), dtype, precision, flags, tuning,
                                              as_stream(stream));
    });
}

int bem_set_option(bem_context* ctx, int32_t option, int32_t value) {
    if (!ctx) return set_error(BEM_ERR_INVALID, "null context");
    switch (option) {
        case BEM_OPT_LU_PIVOTING:
            if (value < 0 || value > 2) return set_error(BEM_ERR_INVALID, "BEM_OPT_LU_PIVOTING takes 0, 1 or 2");
            ctx->impl.lu_pivoting = value;
            return BEM_OK;
        case BEM_OPT_LU_LOOKAHEAD:
            if (value < -1 || value > 1024) return set_error(BEM_ERR_INVALID, "BEM_OPT_LU_LOOKAHEAD takes -1..1024");
            ctx->impl.lu_lookahead_cus = value;
            return BEM_OK;
        case BEM_OPT_LU_PANEL:
            if (value < 0 || value > 1024) return set_error(BEM_ERR_INVALID, "BEM_OPT_LU_PANEL takes 0..1024");
            ctx->impl.lu_panel = value;
            return BEM_OK;
        case BEM_OPT_LU_GEMM_TILE:
            if (value < 0 || value > 2) return set_error(BEM_ERR_INVALID, "BEM_OPT_LU_GEMM_TILE takes 0, 1 or 2");
            ctx->impl.lu_gemm_tile = value;
            return BEM_OK;
        case BEM_OPT_C64_LU_GEMM:
            if (value != 0 && value != 2) return set_error(BEM_ERR_INVALID, "BEM_OPT_C64_LU_GEMM takes 0 or 2");
            ctx->impl.c64_gemm_h16 = value == 2;
            return BEM_OK;
        case BEM_OPT_WAIT_LIMIT_US:
            if (value < 0 || value > 2000000)
                return set_error(BEM_ERR_INVALID, "BEM_OPT_WAIT_LIMIT_US takes 0..2000000");
            ctx->impl.wait_ticks = (long long)value * 100;  // wall clock of 100 MHz
            return BEM_OK;
        default:
            return set_error(BEM_ERR_INVALID, "unknown option");
    }
}

int bem_lu_factor_batched(bem_context* ctx, int32_t count, void* const* A, const int64_t* ld, const int32_t* n,
                          int32_t dtype, int32_t* const* ipiv, int32_t* info, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::lu_factor_batched(&ctx->impl, count, A, ld, n, dtype, ipiv, info, as_stream(stream));
    });
}

int bem_batch_solve(bem_context* ctx, int32_t count, const void* const* LU, const int64_t* ld, const int32_t* n,
                    int32_t dtype, const int32_t* const* ipiv, const void* const* A, const int64_t* lda,
                    const void* const* b, void* const* x, int32_t refine_iters, double* resid, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::batch_solve(&ctx->impl, count, LU, ld, n, dtype, ipiv, A, lda, b, x, refine_iters, resid,
                                as_stream(stream));
    });
}

int bem_batch_condest(bem_context* ctx, int32_t count, const void* const* LU, const int64_t* ld, const int32_t* n,
                      int32_t dtype, const int32_t* const* ipiv, const void* const* A, const int64_t* lda,
                      int32_t iters, double* out, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::batch_condest(&ctx->impl, count, LU, ld, n, dtype, ipiv, A, lda, iters, out, as_stream(stream));
    });
}

int bem_lu_solve_transpose(bem_context* ctx, const void* LU, int64_t ld, int32_t n, int32_t dtype,
                           const int32_t* ipiv, void* b, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::lu_solve_transpose(&ctx->impl, LU, ld, n, dtype, ipiv, b, as_stream(stream));
    });
}

int bem_lu_solve_noalloc(bem_context* ctx, const void* LU, int64_t ld, int32_t n, int32_t dtype,
                         const int32_t* ipiv, void* b, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::lu_solve_noalloc(&ctx->impl, LU, ld, n, dtype, ipiv, b, as_stream(stream));
    });
}

int bem_shape_gradient(bem_context* ctx, const bem_op_desc* op, const void* lam, const void* u, double alpha_re,
                       double alpha_im, double* grad_vertices, double* grad_k, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::shape_gradient(&ctx->impl, op, lam, u, alpha_re, alpha_im, grad_vertices, grad_k, nullptr,
                                   as_stream(stream));
    });
}

int bem_shape_gradient_kt(bem_context* ctx, const bem_op_desc* op, const void* lam, const void* u,
                          double alpha_re, double alpha_im, double* grad_vertices, double* grad_k, void* kt,
                          void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::shape_gradient(&ctx->impl, op, lam, u, alpha_re, alpha_im, grad_vertices, grad_k, kt,
                                   as_stream(stream));
    });
}

int bem_area_gradient(bem_context* ctx, const double* coef, double* grad_vertices, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::area_gradient(&ctx->impl, coef, grad_vertices, as_stream(stream));
    });
}

int bem_residual(bem_context* ctx, const void* A, int64_t ld, int32_t n, int32_t dtype, const void* x,
                 const void* b, void* r, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::gemv_residual(&ctx->impl, A, ld, n, dtype, x, b, r, as_stream(stream));
    });
}

int bem_refine_solve(bem_context* ctx, const void* A, int64_t lda, int32_t dtype_a, const void* LU, int64_t ldlu,
                     int32_t n, int32_t dtype_lu, const int32_t* ipiv, const void* b, void* x, int32_t max_iter,
                     double tol, bem_refine_stats* stats, void* stream) {
    return with_device(ctx, [&]() -> int {
        return bem::refine_solve(&ctx->impl, A, lda, dtype_a, LU, ldlu, n, dtype_lu, ipiv, b, x, max_iter, tol,
                                 stats, as_stream(stream));
    });
}

int bem_convert(const void* src, int32_t dtype_src, int64_t ld_src, void* dst, int32_t dtype_dst, int64_t ld_dst,
                int32_t rows, int32_t cols, void* stream) {
    return bem::convert_matrix(src, dtype_src, ld_src, dst, dtype_dst, ld_dst, rows, cols, as_stream(stream));
}

int bem_matrix_axpy(const void* src, int64_t ld_src, void* dst, int64_t ld_dst, int32_t dtype, int32_t rows,
                    int32_t cols, double alpha_re, double alpha_im, int32_t transpose, void* stream) {
    return bem::matrix_axpy(src, ld_src, dst, ld_dst, dtype, rows, cols, alpha_re, alpha_im, transpose,
                            as_stream(stream));
}

}  // extern "C"
