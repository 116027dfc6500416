"""ctypes binding of libbem_hip.so (the C ABI in include/bem_hip.h).

Loads the in-tree library built by ``build()`` (csrc/Makefile).  There is no
fallback: if the library is missing or cannot be loaded, every call raises.
Status codes map to the exception types the reference raises for the same
conditions (RuntimeError for invalid configuration, BEM_manager.py:41,193;
NotImplementedError for an unsupported GaussQR, utils/functions.py:80-83).
"""
from __future__ import annotations

import atexit
import ctypes
import os
import subprocess
import sys
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbem_hip.so")
CSRC = os.path.join(_HERE, "csrc")

BEM_OK, BEM_ERR_INVALID, BEM_ERR_UNSUPPORTED, BEM_ERR_HIP, BEM_ERR_SINGULAR = 0, 1, 2, 3, 4
OP_V, OP_K, OP_KADJ, OP_W = 0, 1, 2, 3
F64, C128, F32, C64 = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
FLAG_ROTATE_NEAR, FLAG_SYM_COINCIDE, FLAG_DK, FLAG_TRANSPOSE = 1, 2, 4, 8
PART_FAR, PART_NEAR = 1, 2
KernelType_LAPLACE, KernelType_HELMHOLTZ, KernelType_TRANSMISSION = 0, 1, 2
OPT_C64_LU_GEMM, OPT_LU_LOOKAHEAD, OPT_LU_PIVOTING, OPT_LU_GEMM_TILE, OPT_LU_PANEL, OPT_WAIT_LIMIT_US = 1, 2, 3, 4, 5, 6

# Every symbol include/bem_hip.h declares.
EXPORTS = (
    "bem_create", "bem_destroy", "bem_last_error", "bem_version", "bem_set_mesh",
    "bem_get_geometry", "bem_nearfield_counts", "bem_assemble", "bem_assemble_dirichlet", "bem_assemble_vkw",
    "bem_vkw_dk_forms", "bem_apply",
    "bem_mass_assemble", "bem_mass_apply", "bem_lu_factor", "bem_lu_solve",
    "bem_residual", "bem_refine_solve", "bem_convert", "bem_matrix_axpy", "bem_lu_solve_transpose",
    "bem_shape_gradient", "bem_shape_gradient_kt", "bem_area_gradient", "bem_assemble_combined_batched",
    "bem_lu_factor_batched",
    "bem_batch_solve", "bem_batch_condest", "bem_set_option", "bem_lu_solve_noalloc",
)


class BemOpDesc(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("k", ctypes.c_double), ("sqrt_n", ctypes.c_double),
        ("deg_test", ctypes.c_int32), ("deg_trial", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("parts", ctypes.c_int32),
    ]


class BemRefineStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
        ("residual_norm0", ctypes.c_double), ("residual_norm", ctypes.c_double),
        ("b_norm", ctypes.c_double),
    ]


class BemCfProblem(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("k", ctypes.c_double), ("eta", ctypes.c_double),
                ("A", ctypes.c_void_p), ("ld", ctypes.c_int64)]


class BemTuning(ctypes.Structure):
    _fields_ = [("test_panels_per_block", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32)]


class BEMError(RuntimeError):
    pass


class SingularMatrixError(BEMError):
    pass


def build(jobs: int = 4) -> str:
    """Compiles csrc/ into libbem_hip.so for gfx950 (always recompiles
    stale objects; make decides)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BEMError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        sig = {
            "bem_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int]),
            "bem_destroy": (ctypes.c_int, [vp]),
            "bem_last_error": (ctypes.c_char_p, []),
            "bem_version": (ctypes.c_int, []),
            "bem_set_mesh": (ctypes.c_int, [vp, vp, vp, i32, i32, i32, vp]),
            "bem_get_geometry": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
            "bem_nearfield_counts": (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
            "bem_assemble": (ctypes.c_int, [vp, ctypes.POINTER(BemOpDesc), vp, i64, dbl, dbl, vp]),
            "bem_apply": (ctypes.c_int, [vp, ctypes.POINTER(BemOpDesc), vp, i32, vp, dbl, dbl, vp]),
            "bem_assemble_dirichlet": (ctypes.c_int, [vp, ctypes.POINTER(BemOpDesc), vp, i64, dbl, dbl, vp, i32, vp,
                                                      i32, dbl, dbl, vp]),
            "bem_assemble_vkw": (ctypes.c_int, [vp, ctypes.POINTER(BemOpDesc), vp, i64, vp, i64, vp, i64, vp]),
            "bem_vkw_dk_forms": (ctypes.c_int, [vp, ctypes.POINTER(BemOpDesc), vp, vp, vp, vp, vp, vp]),
            "bem_mass_assemble": (ctypes.c_int, [vp, i32, i32, i32, vp, i64, dbl, i32, vp]),
            "bem_mass_apply": (ctypes.c_int, [vp, i32, i32, i32, vp, i32, vp, dbl, vp]),
            "bem_lu_factor": (ctypes.c_int, [vp, vp, i64, i32, i32, vp, vp]),
            "bem_lu_solve": (ctypes.c_int, [vp, vp, i64, i32, i32, vp, vp, vp]),
            "bem_lu_solve_transpose": (ctypes.c_int, [vp, vp, i64, i32, i32, vp, vp, vp]),
            "bem_lu_solve_noalloc": (ctypes.c_int, [vp, vp, i64, i32, i32, vp, vp, vp]),
            "bem_shape_gradient": (ctypes.c_int, [vp, ctypes.POINTER(BemOpDesc), vp, vp, dbl, dbl, vp, vp, vp]),
            "bem_shape_gradient_kt": (ctypes.c_int,
                                      [vp, ctypes.POINTER(BemOpDesc), vp, vp, dbl, dbl, vp, vp, vp, vp]),
            "bem_area_gradient": (ctypes.c_int, [vp, vp, vp, vp]),
            "bem_set_option": (ctypes.c_int, [vp, i32, i32]),
            "bem_assemble_combined_batched": (ctypes.c_int, [vp, i32, ctypes.POINTER(BemCfProblem), i32, i32, i32,
                                                             ctypes.POINTER(BemTuning), vp]),
            "bem_lu_factor_batched": (ctypes.c_int, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(i64),
                                                     ctypes.POINTER(i32), i32, ctypes.POINTER(vp), vp, vp]),
            "bem_batch_solve": (ctypes.c_int, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(i64), ctypes.POINTER(i32),
                                               i32, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i64),
                                               ctypes.POINTER(vp), ctypes.POINTER(vp), i32, vp, vp]),
            "bem_batch_condest": (ctypes.c_int, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(i64),
                                                 ctypes.POINTER(i32), i32, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                                 ctypes.POINTER(i64), i32, vp, vp]),
            "bem_residual": (ctypes.c_int, [vp, vp, i64, i32, i32, vp, vp, vp, vp]),
            "bem_refine_solve": (ctypes.c_int, [vp, vp, i64, i32, vp, i64, i32, i32, vp, vp, vp, i32, dbl,
                                                ctypes.POINTER(BemRefineStats), vp]),
            "bem_convert": (ctypes.c_int, [vp, i32, i64, vp, i32, i64, i32, i32, vp]),
            "bem_matrix_axpy": (ctypes.c_int, [vp, i64, vp, i64, i32, i32, i32, dbl, dbl, i32, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc == BEM_OK:
        return
    msg = lib().bem_last_error().decode(errors="replace")
    if rc == BEM_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == BEM_ERR_SINGULAR:
        raise SingularMatrixError(msg)
    raise BEMError(msg)


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def current_stream() -> int:
    import torch

    return torch.cuda.current_stream().cuda_stream


def dtype_code(t) -> int:
    import torch

    m = {torch.float64: F64, torch.complex128: C128, torch.float32: F32, torch.complex64: C64}
    if t.dtype not in m:
        raise BEMError(f"unsupported dtype {t.dtype}")
    return m[t.dtype]


def _require_cuda(*ts) -> None:
    """Every tensor crossing the C ABI is a device tensor whose storage holds
    its values: a lazy conjugate or negative view (torch's conj/neg bits)
    would hand the library the unconjugated, unnegated memory."""
    for t in ts:
        if t is None:
            continue
        if t.is_conj() or t.is_neg():
            raise BEMError("lazy conj/neg view: call resolve_conj()/resolve_neg() before passing it")
        if not t.is_cuda:
            raise BEMError("the BEM library takes device (HIP) tensors only")


# Contexts still open when the interpreter exits are closed by an atexit
# hook, which runs before module teardown and before the C runtime's own exit
# handlers (HIP's and a profiler's): the library's streams and device memory
# are then released while the runtime is whole, never from a finaliser at
# interpreter shutdown.
_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def close_all_contexts() -> None:
    for c in list(_live_contexts):
        c.close()


class Context:
    """Owns one bem_context (device scratch, mesh tables, near-field list)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().bem_create(ctypes.byref(h), int(device)))
        self._h = h
        self.device = int(device)
        self.nv = self.nf = 0
        self.gauss_qr = 0
        _live_contexts.add(self)

    def close(self) -> None:
        """Releases the context's device memory and streams (bem_destroy).
        Idempotent; later calls on the context raise BEMError (null context)."""
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().bem_destroy(self._h)
        self._h = None
        _live_contexts.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the atexit hook has already closed every
        # context; a finaliser then makes no HIP call (the function is bound
        # here because module globals may be gone by then)
        if getattr(self, "_h", None) is None or _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_mesh(self, vertices, panels, gauss_qr: int) -> None:
        """vertices [nv,3] float64 and panels [nf,3] int32, device tensors."""
        import torch

        v = vertices.contiguous().to(torch.float64)
        p = panels.contiguous().to(torch.int32)
        _require_cuda(v, p)
        check(lib().bem_set_mesh(self._h, v.data_ptr(), p.data_ptr(), int(v.shape[0]), int(p.shape[0]),
                                 int(gauss_qr), current_stream()))
        self.nv, self.nf, self.gauss_qr = int(v.shape[0]), int(p.shape[0]), int(gauss_qr)
        self._verts, self._panels = v, p

    def geometry(self):
        import torch

        dev = torch.device("cuda", self.device)
        area = torch.empty(self.nf, dtype=torch.float64, device=dev)
        normal = torch.empty(self.nf, 3, dtype=torch.float64, device=dev)
        va = torch.empty(self.nv, dtype=torch.float64, device=dev)
        vn = torch.empty(self.nv, 3, dtype=torch.float64, device=dev)
        check(lib().bem_get_geometry(self._h, area.data_ptr(), normal.data_ptr(), va.data_ptr(),
                                     vn.data_ptr(), current_stream()))
        return area, normal, va, vn

    def nearfield_counts(self) -> list[int]:
        c = (ctypes.c_int64 * 4)()
        check(lib().bem_nearfield_counts(self._h, c))
        return [int(x) for x in c]

    @staticmethod
    def op_desc(op: int, kernel: int, k: float, sqrt_n: float, deg_test: int, deg_trial: int,
                precision: int, dtype: int, flags: int = 0, parts: int = 0) -> BemOpDesc:
        return BemOpDesc(int(op), int(kernel), float(k), float(sqrt_n), int(deg_test), int(deg_trial),
                         int(precision), int(dtype), int(flags), int(parts))

    def assemble(self, desc: BemOpDesc, out, alpha: complex = 1.0) -> None:
        """out (a 2-D device tensor or a view of a block with unit column
        stride) += alpha * Op."""
        _require_cuda(out)
        if out.stride(-1) != 1:
            raise BEMError("output block needs unit column stride")
        desc.dtype = dtype_code(out)
        a = complex(alpha)
        check(lib().bem_assemble(self._h, ctypes.byref(desc), out.data_ptr(), int(out.stride(0)),
                                 a.real, a.imag, current_stream()))

    def assemble_dirichlet(self, desc: BemOpDesc, A, g, rhs, alpha: complex = 1.0, alpha_k: float = 1.0,
                           alpha_m: float = 0.5, conj_g: bool = True) -> None:
        """A += alpha V and rhs += alpha_k K ghat + alpha_m M ghat in one far-field
        pass (P0 Neumann unknowns, P1 Dirichlet data g)."""
        _require_cuda(A, g, rhs)
        if A.stride(-1) != 1:
            raise BEMError("matrix needs unit column stride")
        if g.dtype != rhs.dtype:
            raise BEMError("g and rhs must share a dtype")
        desc.dtype = dtype_code(A)
        a = complex(alpha)
        check(lib().bem_assemble_dirichlet(self._h, ctypes.byref(desc), A.data_ptr(), int(A.stride(0)), a.real,
                                           a.imag, g.data_ptr(), int(bool(conj_g)), rhs.data_ptr(),
                                           dtype_code(rhs), float(alpha_k), float(alpha_m), current_stream()))

    def assemble_vkw(self, desc: BemOpDesc, V, K, W) -> None:
        """V <- V (P0 x P0), K <- K (P0 x P1), W <- W (P1 x P1) of one wave number
        in one far-field pass; `desc` is the P0 single-layer descriptor."""
        _require_cuda(V, K, W)
        if V.dtype != K.dtype or V.dtype != W.dtype:
            raise BEMError("V, K and W must share a dtype")
        for M, shape in ((V, (self.nf, self.nf)), (K, (self.nf, self.nv)), (W, (self.nv, self.nv))):
            if M.dim() != 2 or M.stride(-1) != 1:
                raise BEMError("outputs must be matrices with unit column stride")
            if tuple(M.shape) != shape:
                raise BEMError(f"output of shape {tuple(M.shape)}, expected {shape}")
        desc.dtype = dtype_code(V)
        check(lib().bem_assemble_vkw(self._h, ctypes.byref(desc), V.data_ptr(), int(V.stride(0)), K.data_ptr(),
                                     int(K.stride(0)), W.data_ptr(), int(W.stride(0)), current_stream()))

    def vkw_dk_forms(self, desc: BemOpDesc, aN, bN, aD, bD):
        """The four k-derivative bilinear forms of the Calderon blocks,
        (aN^T dV bN, aN^T dK bD, bN^T dK aD, aD^T dW bD), plain products over
        every panel pair; complex128 vectors, `desc` the P0 single-layer
        descriptor of the wave number.  Returns a complex128 tensor of 4."""
        import torch

        _require_cuda(aN, bN, aD, bD)
        for x, n in ((aN, self.nf), (bN, self.nf), (aD, self.nv), (bD, self.nv)):
            if x.dtype != torch.complex128 or x.dim() != 1 or x.shape[0] != n or not x.is_contiguous():
                raise BEMError(f"vectors must be contiguous complex128 of length nf={self.nf} or nv={self.nv}")
        out = torch.empty(4, dtype=torch.complex128, device=aN.device)
        desc.dtype = C128
        check(lib().bem_vkw_dk_forms(self._h, ctypes.byref(desc), aN.data_ptr(), bN.data_ptr(), aD.data_ptr(),
                                     bD.data_ptr(), out.data_ptr(), current_stream()))
        return out

    def apply(self, desc: BemOpDesc, x, y, alpha: complex = 1.0, conj_x: bool = True) -> None:
        """y[test dof] += alpha Op xhat, or y[trial dof] += alpha Op^T xhat when
        desc.flags has FLAG_TRANSPOSE (xhat = conj(x) if conj_x)."""
        _require_cuda(x, y)
        if x.dtype != y.dtype:
            raise BEMError("x and y must share a dtype")
        desc.dtype = dtype_code(y)
        a = complex(alpha)
        check(lib().bem_apply(self._h, ctypes.byref(desc), x.data_ptr(), int(bool(conj_x)), y.data_ptr(),
                              a.real, a.imag, current_stream()))

    def mass_assemble(self, deg_test: int, deg_trial: int, out, alpha: float = 1.0,
                      transpose: bool = False) -> None:
        _require_cuda(out)
        check(lib().bem_mass_assemble(self._h, int(deg_test), int(deg_trial), dtype_code(out),
                                      out.data_ptr(), int(out.stride(0)), float(alpha), int(bool(transpose)),
                                      current_stream()))

    def mass_apply(self, deg_test: int, deg_trial: int, x, y, alpha: float = 1.0, conj_x: bool = True) -> None:
        _require_cuda(x, y)
        check(lib().bem_mass_apply(self._h, int(deg_test), int(deg_trial), dtype_code(y), x.data_ptr(),
                                   int(bool(conj_x)), y.data_ptr(), float(alpha), current_stream()))

    def set_option(self, option: int, value: int) -> None:
        """bem_set_option (OPT_C64_LU_GEMM: 0 FP32 MFMA (default), 2 split-fp16 planes on the fp16 MFMA)."""
        check(lib().bem_set_option(self._h, int(option), int(value)))

    def lu_factor(self, A, ipiv) -> None:
        _require_cuda(A, ipiv)
        n = int(A.shape[0])
        check(lib().bem_lu_factor(self._h, A.data_ptr(), int(A.stride(0)), n, dtype_code(A), ipiv.data_ptr(),
                                  current_stream()))

    def lu_solve(self, LU, ipiv, b) -> None:
        _require_cuda(LU, ipiv, b)
        check(lib().bem_lu_solve(self._h, LU.data_ptr(), int(LU.stride(0)), int(LU.shape[0]), dtype_code(LU),
                                 ipiv.data_ptr(), b.data_ptr(), current_stream()))

    def lu_solve_noalloc(self, LU, ipiv, b) -> None:
        """lu_solve without any allocation in the library (fixed per-device
        scratch, n <= 65536): for graph capture and serving loops."""
        _require_cuda(LU, ipiv, b)
        check(lib().bem_lu_solve_noalloc(self._h, LU.data_ptr(), int(LU.stride(0)), int(LU.shape[0]),
                                         dtype_code(LU), ipiv.data_ptr(), b.data_ptr(), current_stream()))

    def lu_solve_transpose(self, LU, ipiv, b) -> None:
        """b <- A^-T b (plain transpose) from the factors of A."""
        _require_cuda(LU, ipiv, b)
        check(lib().bem_lu_solve_transpose(self._h, LU.data_ptr(), int(LU.stride(0)), int(LU.shape[0]),
                                           dtype_code(LU), ipiv.data_ptr(), b.data_ptr(), current_stream()))

    def shape_gradient(self, desc: BemOpDesc, lam, u, grad, grad_k=None, alpha: complex = 1.0, kt=None) -> None:
        """grad [nv,3] += Re(alpha sum lam_s u_t dOp_st/dX); grad_k [1] +=
        Re(alpha sum lam_s u_t dOp_st/dk).  lam, u complex128 device vectors.
        kt (K only, complex128 [nv]) += K^T lam from the same evaluations."""
        import torch

        _require_cuda(lam, u, grad, grad_k, kt)
        if kt is not None and (kt.dtype != torch.complex128 or not kt.is_contiguous()):
            raise BEMError("kt must be a contiguous complex128 tensor")
        if lam.dtype != torch.complex128 or u.dtype != torch.complex128:
            raise BEMError("shape gradient coefficients must be complex128")
        if grad.dtype != torch.float64 or not grad.is_contiguous():
            raise BEMError("grad must be a contiguous float64 tensor")
        desc.dtype = C128
        a = complex(alpha)
        if kt is None:
            check(lib().bem_shape_gradient(self._h, ctypes.byref(desc), lam.data_ptr(), u.data_ptr(), a.real,
                                           a.imag, grad.data_ptr(), _ptr(grad_k), current_stream()))
        else:
            check(lib().bem_shape_gradient_kt(self._h, ctypes.byref(desc), lam.data_ptr(), u.data_ptr(), a.real,
                                              a.imag, grad.data_ptr(), _ptr(grad_k), kt.data_ptr(),
                                              current_stream()))

    def area_gradient(self, coef, grad) -> None:
        """grad [nv,3] += sum_i coef_i dA_i/dX (coef float64 [nf])."""
        _require_cuda(coef, grad)
        check(lib().bem_area_gradient(self._h, coef.contiguous().data_ptr(), grad.data_ptr(), current_stream()))

    def residual(self, A, x, b, r) -> None:
        _require_cuda(A, x, b, r)
        check(lib().bem_residual(self._h, A.data_ptr(), int(A.stride(0)), int(A.shape[0]), dtype_code(A),
                                 x.data_ptr(), b.data_ptr(), r.data_ptr(), current_stream()))

    def refine_solve(self, A64, LU, ipiv, b, x, max_iter: int = 30, tol: float = 1e-14) -> BemRefineStats:
        """x from the LU (any precision) refined with FP64-accumulated
        residuals against A64 (FP64 or FP32 storage)."""
        _require_cuda(A64, LU, ipiv, b, x)
        st = BemRefineStats()
        check(lib().bem_refine_solve(self._h, A64.data_ptr(), int(A64.stride(0)), dtype_code(A64), LU.data_ptr(),
                                     int(LU.stride(0)), int(A64.shape[0]), dtype_code(LU), ipiv.data_ptr(),
                                     b.data_ptr(), x.data_ptr(), int(max_iter), float(tol), ctypes.byref(st),
                                     current_stream()))
        return st


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[None if t is None else t.data_ptr() for t in ts])


def _i64s(xs):
    return (ctypes.c_int64 * len(xs))(*[int(x) for x in xs])


def _i32s(xs):
    return (ctypes.c_int32 * len(xs))(*[int(x) for x in xs])


class Batch:
    """Batched entry points over a list of square device matrices (uneven
    sizes allowed).  Tables are cached in the owning context, so a repeated
    call with the same tensors can run inside a HIP graph capture."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def assemble_combined(self, problems, dtype_code_: int, precision: int = PREC_FP32, flags: int = 0,
                          tuning: tuple[int, int] | None = None) -> None:
        """problems: list of (Context, k, eta, A) with A an (nv, nv) complex
        device tensor: A += 1/2 M + K - i eta V."""
        arr = (BemCfProblem * len(problems))()
        for q, (c, k, eta, A) in enumerate(problems):
            _require_cuda(A)
            arr[q] = BemCfProblem(c.handle.value, float(k), float(eta), A.data_ptr(), int(A.stride(0)))
        tn = BemTuning(*(tuning or (0, 0)))
        check(lib().bem_assemble_combined_batched(self.ctx.handle, len(problems), arr, int(dtype_code_),
                                                  int(precision), int(flags), ctypes.byref(tn), current_stream()))

    def lu_factor(self, mats, ipivs, info) -> None:
        _require_cuda(*mats, *ipivs, info)
        self._keep = (_ptrs(mats), _i64s([m.stride(0) for m in mats]), _i32s([m.shape[0] for m in mats]),
                      _ptrs(ipivs))
        a, ld, n, ip = self._keep
        check(lib().bem_lu_factor_batched(self.ctx.handle, len(mats), a, ld, n, dtype_code(mats[0]), ip,
                                          info.data_ptr(), current_stream()))

    def solve(self, lus, ipivs, mats, bs, xs, refine_iters: int = 0, resid=None) -> None:
        """x_q = A_q^-1 b_q from the factors, then refine_iters FP64-residual
        refinement steps against mats (same dtype as the factors; may be None
        without refinement or residuals).  b, x complex128 (float64 for real)."""
        _require_cuda(*lus, *ipivs, *bs, *xs)
        A = _ptrs(mats) if mats is not None else None
        lda = _i64s([m.stride(0) for m in mats]) if mats is not None else None
        check(lib().bem_batch_solve(self.ctx.handle, len(lus), _ptrs(lus), _i64s([m.stride(0) for m in lus]),
                                    _i32s([m.shape[0] for m in lus]), dtype_code(lus[0]), _ptrs(ipivs), A, lda,
                                    _ptrs(bs), _ptrs(xs), int(refine_iters), _ptr(resid), current_stream()))

    def condest(self, lus, ipivs, mats, out, iters: int = 3) -> None:
        """out [count, 2] float64: ||A||_1 and an estimate of ||A^-1||_1."""
        _require_cuda(*lus, *ipivs, *mats, out)
        check(lib().bem_batch_condest(self.ctx.handle, len(lus), _ptrs(lus), _i64s([m.stride(0) for m in lus]),
                                      _i32s([m.shape[0] for m in lus]), dtype_code(lus[0]), _ptrs(ipivs),
                                      _ptrs(mats), _i64s([m.stride(0) for m in mats]), int(iters), out.data_ptr(),
                                      current_stream()))


def convert(src, dst) -> None:
    """dst[...] = src with precision conversion on the device (2-D or 1-D)."""
    _require_cuda(src, dst)
    s2 = src if src.dim() == 2 else src.view(1, -1)
    d2 = dst if dst.dim() == 2 else dst.view(1, -1)
    check(lib().bem_convert(s2.data_ptr(), dtype_code(src), int(s2.stride(0)), d2.data_ptr(), dtype_code(dst),
                            int(d2.stride(0)), int(s2.shape[0]), int(s2.shape[1]), current_stream()))


def matrix_axpy(src, dst, alpha: complex = 1.0, transpose: bool = False) -> None:
    """dst += alpha * src (or src^T): 2-D device blocks of one dtype."""
    _require_cuda(src, dst)
    if src.dtype != dst.dtype:
        raise BEMError("axpy needs one dtype")
    a = complex(alpha)
    rows, cols = int(dst.shape[0]), int(dst.shape[1])
    check(lib().bem_matrix_axpy(src.data_ptr(), int(src.stride(0)), dst.data_ptr(), int(dst.stride(0)),
                                dtype_code(dst), rows, cols, a.real, a.imag, int(bool(transpose)),
                                current_stream()))
