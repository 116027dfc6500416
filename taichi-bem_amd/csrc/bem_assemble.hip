// bem_assemble.hip -- Galerkin operator assembly and matrix-free products.
//
// Work split (the reference runs one thread per test panel over all trial
// panels, basis pairs and Gauss points serially, single_layer_3d.py:385-433):
//
//   far field  : regular (well-separated) pairs with the tensor rule.  A
//                block owns one test panel i (wave-uniform: its quadrature
//                points come through scalar loads) and 256 trial panels j
//                (one per lane, coalesced SoA loads).  Every lane evaluates
//                the kernel once per point pair and accumulates all basis
//                pairs from that single evaluation.  Pairs that share a
//                vertex are skipped here.
//   near field : coincident / edge / vertex pairs from the near list, one
//                wave per pair, lanes striding over the Sauter-Schwab points
//                of the pair's relation, FP64 always, wave-reduced sums.
//
// Output: out[row * ld + col] += alpha * value with row/col the test/trial
// dof (panel for P0, vertex for P1).  Assembled P1 rows or columns collect
// from several panels and use atomics; P0 x P0 entries have a single writer.
// Matrix-free products into vertex rows store per-panel and per-pair partials
// and add them per vertex in panel order (bitwise reproducible).
#include <algorithm>
#include <cstring>
#include <vector>

#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

namespace {

struct OpParams {
    int nf, nv, q2, flags;
    const int32_t* __restrict__ tris;
    const double* __restrict__ verts;
    const double* __restrict__ area;
    const double* __restrict__ normal;
    const double* __restrict__ curl;
    const double* __restrict__ cen;
    const double* __restrict__ pu;
    const double* __restrict__ pl;
    const float* __restrict__ pu32;
    const float* __restrict__ pl32;
    const double* __restrict__ phi;
    const float* __restrict__ phi32;
    const double* __restrict__ frame;  // [nf][PANEL_FRAME] local frames
    const double* __restrict__ pq;     // [nf][q2][4] (|x - C|^2, -2 r1, -2 r2, w)
    const float* __restrict__ pq32;    // the same in FP32
    const NearPair* __restrict__ near;
    const int32_t* __restrict__ near_row;
    const NearPt* __restrict__ rule;
    int32_t rule_off[4], rule_len[4];
    int64_t n_near;
    double kap;   // wave number inside G
    double k2n;   // k^2 sqrt_n^2 of the hypersingular term
    double dkap;  // d kap / d k        (BEM_FLAG_DK products)
    double dk2n;  // d k2n / d k = 2 k sqrt_n^2
    // output / vectors
    void* out;
    int64_t ld;
    int dtype;
    double ar, ai;
    const void* x;
    int conj_x;
    int dtype_x;
    void* y;
    double* near_buf;
    double* panel_buf;                    // [nf][3][2] far-field partials of P1 test rows
    const int32_t* __restrict__ near_tr;  // mirrored pair of each near pair
    const int32_t* __restrict__ vp_row;   // vertex -> panel corners (CSR, panel order)
    const int32_t* __restrict__ vp_el;
};

template <typename T> struct Ptr;
template <> struct Ptr<double> {
    static __device__ __forceinline__ const double* pq(const OpParams& p) { return p.pq; }
    static __device__ __forceinline__ const double* pu(const OpParams& p) { return p.pu; }
    static __device__ __forceinline__ const double* pl(const OpParams& p) { return p.pl; }
    static __device__ __forceinline__ const double* phi(const OpParams& p) { return p.phi; }
};
template <> struct Ptr<float> {
    static __device__ __forceinline__ const float* pq(const OpParams& p) { return p.pq32; }
    static __device__ __forceinline__ const float* pu(const OpParams& p) { return p.pu32; }
    static __device__ __forceinline__ const float* pl(const OpParams& p) { return p.pl32; }
    static __device__ __forceinline__ const float* phi(const OpParams& p) { return p.phi32; }
};

// dtype-generic accumulate into an output element.
__device__ __forceinline__ void add_elem(void* out, int dtype, int64_t idx, double vr, double vi,
                                         bool atomic) {
    switch (dtype) {
        case BEM_F64: {
            double* o = (double*)out + idx;
            if (atomic) atomicAdd(o, vr); else *o += vr;
        } break;
        case BEM_C128: {
            double* o = (double*)out + 2 * idx;
            if (atomic) { atomicAdd(o, vr); atomicAdd(o + 1, vi); }
            else { o[0] += vr; o[1] += vi; }
        } break;
        case BEM_F32: {
            float* o = (float*)out + idx;
            if (atomic) atomicAdd(o, (float)vr); else *o += (float)vr;
        } break;
        default: {
            float* o = (float*)out + 2 * idx;
            if (atomic) { atomicAdd(o, (float)vr); atomicAdd(o + 1, (float)vi); }
            else { o[0] += (float)vr; o[1] += (float)vi; }
        } break;
    }
}

__device__ __forceinline__ void load_elem(const void* x, int dtype, int64_t idx, int conj,
                                          double& xr, double& xi) {
    switch (dtype) {
        case BEM_F64: xr = ((const double*)x)[idx]; xi = 0.0; break;
        case BEM_C128: xr = ((const double*)x)[2 * idx]; xi = ((const double*)x)[2 * idx + 1]; break;
        case BEM_F32: xr = ((const float*)x)[idx]; xi = 0.0; break;
        default: xr = ((const float*)x)[2 * idx]; xi = ((const float*)x)[2 * idx + 1]; break;
    }
    if (conj) xi = -xi;
}

__device__ __forceinline__ bool shares_vertex(const int32_t* t, int i, int j) {
    int a0 = t[3 * i], a1 = t[3 * i + 1], a2 = t[3 * i + 2];
    int b0 = t[3 * j], b1 = t[3 * j + 1], b2 = t[3 * j + 2];
    return a0 == b0 || a0 == b1 || a0 == b2 || a1 == b0 || a1 == b1 || a1 == b2 || a2 == b0 ||
           a2 == b1 || a2 == b2;
}

// ---------------------------------------------------------------------------
// Regular pair (i test, j trial): full basis block in FP64 (values already
// multiplied by 1/(4 pi) and, for W, combined with the curl term).
//   V  G                          (single_layer_3d.py:155-177)
//   K  dG/dn_y(x, y)              (double_layer_3d.py:158-180)
//   K' dG/dn_y with x<->y, n_x    (adj_double_layer_3d.py:160-182)
//   W  G [curl.curl - k^2 n phi phi n_x.n_y]   (hypersingular_layer_3d.py:186-211)
// DK = true gives the derivative of the entries with respect to k instead:
//   dG/dk = dkap i exp(i kap r) / (4 pi),
//   d(dG/dn_y)/dk = dkap kap (x-y).n_y exp(i kap r) / (4 pi r),
//   W: dG/dk curl.curl - (dk2n G + k2n dG/dk) phi phi n_x.n_y.
// ---------------------------------------------------------------------------
template <typename R, int OP, bool HELM, int DT, int DR, bool DK = false>
__device__ __forceinline__ void far_pair(const OpParams& p, int i, int j, double (&vr)[3][3],
                                         double (&vi)[3][3]) {
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    const int q2 = p.q2;
    const int nf = p.nf;
    // test panel i is wave-uniform in every caller: its points and the basis
    // table are read with scalar loads
    const ConstPtr<R> PU = as_const(Ptr<R>::pu(p) + (int64_t)i * q2 * 4);
    const R* __restrict__ PL = Ptr<R>::pl(p);
    const ConstPtr<R> PHI = as_const(Ptr<R>::phi(p));
    // FP32: d = (c_i - c_j) + (o_a - o_b); FP64: absolute coordinates.
    R dcx = 0, dcy = 0, dcz = 0;
    if constexpr (sizeof(R) == 4) {
        dcx = (R)(p.cen[3 * i] - p.cen[3 * j]);
        dcy = (R)(p.cen[3 * i + 1] - p.cen[3 * j + 1]);
        dcz = (R)(p.cen[3 * i + 2] - p.cen[3 * j + 2]);
    }
    R nnx = 0, nny = 0, nnz = 0;
    if constexpr (OP == BEM_OP_K) {
        nnx = (R)p.normal[3 * j]; nny = (R)p.normal[3 * j + 1]; nnz = (R)p.normal[3 * j + 2];
    } else if constexpr (OP == BEM_OP_KADJ) {
        // (y - x).n_x = -(x - y).n_x
        nnx = -(R)p.normal[3 * i]; nny = -(R)p.normal[3 * i + 1]; nnz = -(R)p.normal[3 * i + 2];
    }
    const R kap = (R)p.kap;
    const R dkap = (R)p.dkap, k2n = (R)p.k2n, dk2n = (R)p.dk2n;
    R ar[NS][NT], ai[NS][NT];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t) { ar[s][t] = 0; ai[s][t] = 0; }
    R s0r = 0, s0i = 0;  // W: sum of w G (curl term)

    for (int b = 0; b < q2; ++b) {
        const R ex = dcx - PL[((int64_t)b * 4 + 0) * nf + j];
        const R ey = dcy - PL[((int64_t)b * 4 + 1) * nf + j];
        const R ez = dcz - PL[((int64_t)b * 4 + 2) * nf + j];
        const R wb = PL[((int64_t)b * 4 + 3) * nf + j];
        R pr[NS], pi[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) { pr[s] = 0; pi[s] = 0; }
        R p1r = 0, p1i = 0;  // W with DK: sum of w dG/dk
        for (int a = 0; a < q2; ++a) {
            const R dx = PU[4 * a + 0] + ex;
            const R dy = PU[4 * a + 1] + ey;
            const R dz = PU[4 * a + 2] + ez;
            const R wa = PU[4 * a + 3];
            const R r2 = dx * dx + dy * dy + dz * dz;
            const R rinv = Math<R>::rsqrt(r2);
            R fr, fi;
            if constexpr (DK && (OP == BEM_OP_V || OP == BEM_OP_W)) {
                R sn, cs;
                Math<R>::sincos(kap * (r2 * rinv), &sn, &cs);
                const R f1r = -wa * sn * dkap, f1i = wa * cs * dkap;
                if constexpr (OP == BEM_OP_W) {
                    const R g = wa * rinv;
                    fr = dk2n * g * cs + k2n * f1r;
                    fi = dk2n * g * sn + k2n * f1i;
                    p1r += f1r; p1i += f1i;
                } else { fr = f1r; fi = f1i; }
            } else if constexpr (DK) {
                const R dn = dx * nnx + dy * nny + dz * nnz;
                R sn, cs;
                Math<R>::sincos(kap * (r2 * rinv), &sn, &cs);
                const R base = wa * dn * rinv * kap * dkap;
                fr = base * cs; fi = base * sn;
            } else if constexpr (OP == BEM_OP_V || OP == BEM_OP_W) {
                const R g = wa * rinv;
                if constexpr (HELM) {
                    R sn, cs;
                    Math<R>::sincos(kap * (r2 * rinv), &sn, &cs);
                    fr = g * cs; fi = g * sn;
                } else { fr = g; fi = 0; }
            } else {
                const R dn = dx * nnx + dy * nny + dz * nnz;
                const R base = wa * dn * (rinv * rinv * rinv);
                if constexpr (HELM) {
                    const R kr = kap * (r2 * rinv);
                    R sn, cs;
                    Math<R>::sincos(kr, &sn, &cs);
                    fr = base * (cs + kr * sn); fi = base * (sn - kr * cs);
                } else { fr = base; fi = 0; }
            }
            if constexpr (DT) {
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    const R ph = PHI[3 * a + s];
                    pr[s] += fr * ph; pi[s] += fi * ph;
                }
            } else { pr[0] += fr; pi[0] += fi; }
        }
        if constexpr (OP == BEM_OP_W && DK) {
            s0r += wb * p1r;
            s0i += wb * p1i;
        } else if constexpr (OP == BEM_OP_W) {
            // sum_a w_a G: the P1 bases sum to one, so it is the sum of the
            // basis-weighted partials
            R sr = 0, si = 0;
#pragma unroll
            for (int s = 0; s < NS; ++s) { sr += pr[s]; si += pi[s]; }
            s0r += wb * sr;
            s0i += wb * si;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const R wt = DR ? wb * PHI[3 * b + t] : wb;
#pragma unroll
            for (int s = 0; s < NS; ++s) { ar[s][t] += pr[s] * wt; ai[s][t] += pi[s] * wt; }
        }
    }
    if constexpr (OP == BEM_OP_W) {
        const double nxy = p.normal[3 * i] * p.normal[3 * j] + p.normal[3 * i + 1] * p.normal[3 * j + 1] +
                           p.normal[3 * i + 2] * p.normal[3 * j + 2];
        const double kf = DK ? 1.0 : p.k2n;  // DK: the k^2 factor is inside the sums
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                double cc = 0.0;
                if (DT && DR)
                    for (int d = 0; d < 3; ++d) cc += p.curl[9 * i + 3 * s + d] * p.curl[9 * j + 3 * t + d];
                vr[s][t] = kInv4Pi * (cc * (double)s0r - kf * nxy * (double)ar[s][t]);
                vi[s][t] = kInv4Pi * (cc * (double)s0i - kf * nxy * (double)ai[s][t]);
            }
    } else {
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                vr[s][t] = kInv4Pi * (double)ar[s][t];
                vi[s][t] = kInv4Pi * (double)ai[s][t];
            }
    }
}

template <typename R, int OP, bool HELM, int DT, int DR>
__global__ __launch_bounds__(256) void far_assemble_kernel(OpParams p, int njb) {
    const int i = blockIdx.x / njb;
    const int j = (blockIdx.x % njb) * 256 + threadIdx.x;
    if (j >= p.nf) return;
    if (shares_vertex(p.tris, i, j)) return;  // near field
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    double vr[3][3], vi[3][3];
    far_pair<R, OP, HELM, DT, DR>(p, i, j, vr, vi);
    const bool atomic = DT || DR;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int64_t row = DT ? p.tris[3 * i + s] : i;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int64_t col = DR ? p.tris[3 * j + t] : j;
            add_elem(p.out, p.dtype, row * p.ld + col, p.ar * vr[s][t] - p.ai * vi[s][t],
                     p.ar * vi[s][t] + p.ai * vr[s][t], atomic);
        }
    }
}

// ---------------------------------------------------------------------------
// Near pair: Sauter-Schwab rule of the pair's relation in FP64, lanes
// striding over the rule points, then a wave reduction.  The coincident rule
// evaluates its second term f(y, x) with the same bases (reference,
// single_layer_3d.py:198-206) unless BEM_FLAG_SYM_COINCIDE.
// ---------------------------------------------------------------------------
// near_pair's pseudo-operator for the combined field K - i eta V (P1 x P1,
// Helmholtz): one kernel evaluation per rule point serves both operators.
constexpr int kOpCf = 16;

template <int OP, bool HELM, int DT, int DR, bool DK = false>
__device__ __forceinline__ void near_pair(const OpParams& p, const NearPair& np, int lane,
                                          double (&vr)[3][3], double (&vi)[3][3], double eta = 0.0) {
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    const int i = np.i, j = np.j, rel = np.rel;
    int px[3], py[3], invx[3], invy[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        px[m] = (np.perm >> (2 * m)) & 3;
        py[m] = (np.perm >> (8 + 2 * m)) & 3;
    }
    // position of stored slot s in the (possibly rotated) parametrisation
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        invx[sl] = px[0] == sl ? 0 : (px[1] == sl ? 1 : 2);
        invy[sl] = py[0] == sl ? 0 : (py[1] == sl ? 1 : 2);
    }
    double X[3][3], Y[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            X[m][d] = p.verts[3 * p.tris[3 * i + px[m]] + d];
            Y[m][d] = p.verts[3 * p.tris[3 * j + py[m]] + d];
        }
    const double nx0 = p.normal[3 * i], nx1 = p.normal[3 * i + 1], nx2 = p.normal[3 * i + 2];
    const double ny0 = p.normal[3 * j], ny1 = p.normal[3 * j + 1], ny2 = p.normal[3 * j + 2];
    const double w4a = (p.area[i] * 2.0) * (p.area[j] * 2.0);
    const double kap = p.kap;
    const bool twice = rel == BEM_COINCIDE;
    const bool sym = (p.flags & BEM_FLAG_SYM_COINCIDE) != 0;
    const NearPt* __restrict__ rule = p.rule + p.rule_off[rel];
    const int len = p.rule_len[rel];

    double ar[NS][NT], ai[NS][NT];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t) { ar[s][t] = 0.0; ai[s][t] = 0.0; }
    double s0r = 0.0, s0i = 0.0;

    for (int e = lane; e < len; e += 64) {
        const NearPt q = rule[e];
        double x[3], y[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[d] = (1.0 - q.r1x) * X[0][d] + (q.r1x - q.r2x) * X[1][d] + q.r2x * X[2][d];
            y[d] = (1.0 - q.r1y) * Y[0][d] + (q.r1y - q.r2y) * Y[1][d] + q.r2y * Y[2][d];
        }
        const double w = q.w * w4a;
        for (int pass = 0; pass <= (int)twice; ++pass) {
            // kernel argument order: pass 1 evaluates f(y, x)
            const double sg = pass ? -1.0 : 1.0;
            const double dx = sg * (x[0] - y[0]), dy = sg * (x[1] - y[1]), dz = sg * (x[2] - y[2]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double rinv = rsqrt_f64(r2);
            double fr, fi, f1r = 0.0, f1i = 0.0;
            if constexpr (DK && (OP == BEM_OP_V || OP == BEM_OP_W)) {
                double sn, cs;
                Math<double>::sincos(kap * (r2 * rinv), &sn, &cs);
                f1r = -sn * p.dkap; f1i = cs * p.dkap;
                if constexpr (OP == BEM_OP_W) {
                    fr = p.dk2n * rinv * cs + p.k2n * f1r;
                    fi = p.dk2n * rinv * sn + p.k2n * f1i;
                } else { fr = f1r; fi = f1i; }
            } else if constexpr (DK) {
                const double dn = (OP == BEM_OP_K) ? (dx * ny0 + dy * ny1 + dz * ny2)
                                                   : -(dx * nx0 + dy * nx1 + dz * nx2);
                double sn, cs;
                Math<double>::sincos(kap * (r2 * rinv), &sn, &cs);
                const double base = dn * rinv * kap * p.dkap;
                fr = base * cs; fi = base * sn;
            } else if constexpr (OP == kOpCf) {
                // K - i eta V from one evaluation: K (dn rinv^3)(cs + kr sn, sn - kr cs), V rinv (cs, sn)
                const double dn = dx * ny0 + dy * ny1 + dz * ny2;
                const double kr = kap * (r2 * rinv);
                double sn, cs;
                Math<double>::sincos(kr, &sn, &cs);
                const double base = dn * (rinv * rinv * rinv);
                fr = base * (cs + kr * sn) + eta * (rinv * sn);
                fi = base * (sn - kr * cs) - eta * (rinv * cs);
            } else if constexpr (OP == BEM_OP_V || OP == BEM_OP_W) {
                if constexpr (HELM) {
                    double sn, cs;
                    Math<double>::sincos(kap * (r2 * rinv), &sn, &cs);
                    fr = rinv * cs; fi = rinv * sn;
                } else { fr = rinv; fi = 0.0; }
            } else {
                const double dn = (OP == BEM_OP_K) ? (dx * ny0 + dy * ny1 + dz * ny2)
                                                   : -(dx * nx0 + dy * nx1 + dz * nx2);
                const double base = dn * (rinv * rinv * rinv);
                if constexpr (HELM) {
                    const double kr = kap * (r2 * rinv);
                    double sn, cs;
                    Math<double>::sincos(kr, &sn, &cs);
                    fr = base * (cs + kr * sn); fi = base * (sn - kr * cs);
                } else { fr = base; fi = 0.0; }
            }
            fr *= w; fi *= w;
            const bool swap_b = pass && sym;
            const double bx1 = swap_b ? q.r1y : q.r1x, bx2 = swap_b ? q.r2y : q.r2x;
            const double by1 = swap_b ? q.r1x : q.r1y, by2 = swap_b ? q.r2x : q.r2y;
            if constexpr (OP == BEM_OP_W && DK) { s0r += f1r * w; s0i += f1i * w; }
            else if constexpr (OP == BEM_OP_W) { s0r += fr; s0i += fi; }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double phx = DT ? shape(bx1, bx2, invx[s]) : 1.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const double ph = phx * (DR ? shape(by1, by2, invy[t]) : 1.0);
                    ar[s][t] += fr * ph; ai[s][t] += fi * ph;
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t) { ar[s][t] = wave_sum(ar[s][t]); ai[s][t] = wave_sum(ai[s][t]); }
    if constexpr (OP == BEM_OP_W) {
        s0r = wave_sum(s0r); s0i = wave_sum(s0i);
        const double nxy = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
        const double kf = DK ? 1.0 : p.k2n;
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                double cc = 0.0;
                if (DT && DR)
                    for (int d = 0; d < 3; ++d) cc += p.curl[9 * i + 3 * s + d] * p.curl[9 * j + 3 * t + d];
                vr[s][t] = kInv4Pi * (cc * s0r - kf * nxy * ar[s][t]);
                vi[s][t] = kInv4Pi * (cc * s0i - kf * nxy * ai[s][t]);
            }
    } else {
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) { vr[s][t] = kInv4Pi * ar[s][t]; vi[s][t] = kInv4Pi * ai[s][t]; }
    }
}

template <int OP, bool HELM, int DT, int DR>
__global__ __launch_bounds__(256) void near_assemble_kernel(OpParams p) {
    const int lane = threadIdx.x & 63;
    const int64_t pid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pid >= p.n_near) return;
    const NearPair np = p.near[pid];
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    double vr[3][3], vi[3][3];
    near_pair<OP, HELM, DT, DR>(p, np, lane, vr, vi);
    if (lane != 0) return;
    const bool atomic = DT || DR;
    for (int s = 0; s < NS; ++s) {
        const int64_t row = DT ? p.tris[3 * np.i + s] : np.i;
        for (int t = 0; t < NT; ++t) {
            const int64_t col = DR ? p.tris[3 * np.j + t] : np.j;
            add_elem(p.out, p.dtype, row * p.ld + col, p.ar * vr[s][t] - p.ai * vi[s][t],
                     p.ar * vi[s][t] + p.ai * vr[s][t], atomic);
        }
    }
}

// ---------------------------------------------------------------------------
// Matrix-free products y[test] += alpha Op xhat.
// ---------------------------------------------------------------------------
template <typename R, int OP, bool HELM, int DT, int DR, bool DK = false>
__global__ __launch_bounds__(256) void far_apply_kernel(OpParams p) {
    const int i = blockIdx.x;
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    double yr[NS], yi[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) { yr[s] = 0.0; yi[s] = 0.0; }
    for (int j = threadIdx.x; j < p.nf; j += 256) {
        if (shares_vertex(p.tris, i, j)) continue;
        double vr[3][3], vi[3][3];
        far_pair<R, OP, HELM, DT, DR, DK>(p, i, j, vr, vi);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            double xr, xi;
            load_elem(p.x, p.dtype_x, DR ? p.tris[3 * j + t] : j, p.conj_x, xr, xi);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                yr[s] += vr[s][t] * xr - vi[s][t] * xi;
                yi[s] += vr[s][t] * xi + vi[s][t] * xr;
            }
        }
    }
    __shared__ double red[4][6];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = wave_sum(yr[s]), b = wave_sum(yi[s]);
        if (lane == 0) { red[wid][2 * s] = a; red[wid][2 * s + 1] = b; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int s = 0; s < NS; ++s) {
            double a = red[0][2 * s] + red[1][2 * s] + red[2][2 * s] + red[3][2 * s];
            double b = red[0][2 * s + 1] + red[1][2 * s + 1] + red[2][2 * s + 1] + red[3][2 * s + 1];
            if (DT) {
                // vertex rows collect from several panels: vertex_rows_kernel adds them in panel order
                p.panel_buf[6 * i + 2 * s] = a;
                p.panel_buf[6 * i + 2 * s + 1] = b;
            } else {
                add_elem(p.y, p.dtype_x, i, p.ar * a - p.ai * b, p.ar * b + p.ai * a, false);
            }
        }
    }
}

template <int OP, bool HELM, int DT, int DR, bool DK = false>
__global__ __launch_bounds__(256) void near_apply_kernel(OpParams p) {
    const int lane = threadIdx.x & 63;
    const int64_t pid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pid >= p.n_near) return;
    const NearPair np = p.near[pid];
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    double vr[3][3], vi[3][3];
    near_pair<OP, HELM, DT, DR, DK>(p, np, lane, vr, vi);
    if (lane != 0) return;
    for (int s = 0; s < NS; ++s) {
        double yr = 0.0, yi = 0.0;
        for (int t = 0; t < NT; ++t) {
            double xr, xi;
            load_elem(p.x, p.dtype_x, DR ? p.tris[3 * np.j + t] : np.j, p.conj_x, xr, xi);
            yr += vr[s][t] * xr - vi[s][t] * xi;
            yi += vr[s][t] * xi + vi[s][t] * xr;
        }
        p.near_buf[pid * 6 + 2 * s] = yr;
        p.near_buf[pid * 6 + 2 * s + 1] = yi;
    }
}

// ---------------------------------------------------------------------------
// Transposed products y[trial] += alpha Op^T xhat (BEM_FLAG_TRANSPOSE), FP64,
// from the same pair evaluations as the forward products.  Far: a block
// holds 256 trial panels (one per lane) and a chunk of TCH test panels
// (wave-uniform, as in far_apply_kernel); a lane sums Op[s][t] x[s] over its
// chunk and adds its trial dofs once, with atomics (the blocks of the test
// chunks meet there, in no fixed order).  Near: one wave per pair writes its
// per-pair sums, and trial_gather_kernel adds them in a fixed order.
// ---------------------------------------------------------------------------
constexpr int TCH = 64;

template <int OP, bool HELM, int DT, int DR>
__global__ __launch_bounds__(256) void far_apply_t_kernel(OpParams p, int njb) {
    const int ib = blockIdx.x / njb, jb = blockIdx.x % njb;
    const int j = jb * 256 + threadIdx.x;
    if (j >= p.nf) return;
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    double yr[NT], yi[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { yr[t] = 0.0; yi[t] = 0.0; }
    const int i_end = min(p.nf, (ib + 1) * TCH);
    for (int i = ib * TCH; i < i_end; ++i) {
        if (shares_vertex(p.tris, i, j)) continue;
        double vr[3][3], vi[3][3];
        far_pair<double, OP, HELM, DT, DR>(p, i, j, vr, vi);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double xr, xi;
            load_elem(p.x, p.dtype_x, DT ? p.tris[3 * i + s] : i, p.conj_x, xr, xi);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                yr[t] += vr[s][t] * xr - vi[s][t] * xi;
                yi[t] += vr[s][t] * xi + vi[s][t] * xr;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int64_t col = DR ? p.tris[3 * j + t] : j;
        add_elem(p.y, p.dtype_x, col, p.ar * yr[t] - p.ai * yi[t], p.ar * yi[t] + p.ai * yr[t], true);
    }
}

template <int OP, bool HELM, int DT, int DR>
__global__ __launch_bounds__(256) void near_apply_t_kernel(OpParams p) {
    const int lane = threadIdx.x & 63;
    const int64_t pid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pid >= p.n_near) return;
    const NearPair np = p.near[pid];
    constexpr int NS = DT ? 3 : 1, NT = DR ? 3 : 1;
    double vr[3][3], vi[3][3];
    near_pair<OP, HELM, DT, DR>(p, np, lane, vr, vi);
    if (lane != 0) return;
    for (int t = 0; t < NT; ++t) {
        double yr = 0.0, yi = 0.0;
        for (int s = 0; s < NS; ++s) {
            double xr, xi;
            load_elem(p.x, p.dtype_x, DT ? p.tris[3 * np.i + s] : np.i, p.conj_x, xr, xi);
            yr += vr[s][t] * xr - vi[s][t] * xi;
            yi += vr[s][t] * xi + vi[s][t] * xr;
        }
        p.near_buf[pid * 6 + 2 * t] = yr;
        p.near_buf[pid * 6 + 2 * t + 1] = yi;
    }
}

// Sums of the near pairs of test panel i for its slot s, in list order.
__device__ __forceinline__ void near_row_sum(const OpParams& p, int i, int s, double& a, double& b) {
    for (int e = p.near_row[i]; e < p.near_row[i + 1]; ++e) {
        a += p.near_buf[(int64_t)e * 6 + 2 * s];
        b += p.near_buf[(int64_t)e * 6 + 2 * s + 1];
    }
}

// Deterministic per-row sum of the near contributions, P0 test rows.
__global__ void near_reduce_kernel(OpParams p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nf) return;
    double a = 0.0, b = 0.0;
    near_row_sum(p, i, 0, a, b);
    add_elem(p.y, p.dtype_x, i, p.ar * a - p.ai * b, p.ar * b + p.ai * a, false);
}

// P1 test rows: one thread per vertex adds, over the vertex's panel corners
// (i, s) in panel order, the far-field partial of the panel (use_far) and its
// near pairs in list order (use_near).
__global__ void vertex_rows_kernel(OpParams p, int use_far, int use_near) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= p.nv) return;
    double a = 0.0, b = 0.0;
    for (int c = p.vp_row[v]; c < p.vp_row[v + 1]; ++c) {
        const int i = p.vp_el[c] / 3, s = p.vp_el[c] % 3;
        if (use_far) { a += p.panel_buf[6 * i + 2 * s]; b += p.panel_buf[6 * i + 2 * s + 1]; }
        if (use_near) near_row_sum(p, i, s, a, b);
    }
    add_elem(p.y, p.dtype_x, v, p.ar * a - p.ai * b, p.ar * b + p.ai * a, false);
}

// Near part of a transposed product: trial dof j (P0) or v (P1) collects the
// pairs whose trial panel is j, found as the mirrors of row j's pairs, in
// row order (and for P1 over the vertex's panel corners in panel order).
__device__ __forceinline__ void near_col_sum(const OpParams& p, int j, int t, double& a, double& b) {
    for (int e = p.near_row[j]; e < p.near_row[j + 1]; ++e) {
        const int64_t f = p.near_tr[e];
        a += p.near_buf[f * 6 + 2 * t];
        b += p.near_buf[f * 6 + 2 * t + 1];
    }
}

__global__ void trial_gather_kernel(OpParams p, int deg_trial) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (!deg_trial) {
        if (v >= p.nf) return;
        near_col_sum(p, v, 0, a, b);
    } else {
        if (v >= p.nv) return;
        for (int c = p.vp_row[v]; c < p.vp_row[v + 1]; ++c) near_col_sum(p, p.vp_el[c] / 3, p.vp_el[c] % 3, a, b);
    }
    add_elem(p.y, p.dtype_x, v, p.ar * a - p.ai * b, p.ar * b + p.ai * a, false);
}

// ---------------------------------------------------------------------------
// Mass matrix, single-panel rule (identity_layer_3d.py:117-146).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mass_block(const OpParams& p, const double* mw, int i, int dt, int dr,
                                           double (&m)[3][3]) {
    for (int s = 0; s < 3; ++s)
        for (int t = 0; t < 3; ++t) m[s][t] = 0.0;
    const double a2 = p.area[i] * 2.0;
    for (int e = 0; e < p.q2; ++e) {
        const double w = mw[e] * a2;
        for (int s = 0; s < (dt ? 3 : 1); ++s) {
            const double ps = dt ? p.phi[3 * e + s] : 1.0;
            for (int t = 0; t < (dr ? 3 : 1); ++t) m[s][t] += ps * (dr ? p.phi[3 * e + t] : 1.0) * w;
        }
    }
}

__global__ void mass_assemble_kernel(OpParams p, const double* mw, int dt, int dr, int transpose) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nf) return;
    double m[3][3];
    mass_block(p, mw, i, dt, dr, m);
    const bool atomic = dt && dr;
    for (int s = 0; s < (dt ? 3 : 1); ++s)
        for (int t = 0; t < (dr ? 3 : 1); ++t) {
            int64_t row = dt ? p.tris[3 * i + s] : i, col = dr ? p.tris[3 * i + t] : i;
            if (transpose) { int64_t tmp = row; row = col; col = tmp; }
            add_elem(p.out, p.dtype, row * p.ld + col, p.ar * m[s][t], 0.0, atomic);
        }
}

// P0 test rows: one thread per panel.
__global__ void mass_apply_kernel(OpParams p, const double* mw, int dr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nf) return;
    double m[3][3];
    mass_block(p, mw, i, 0, dr, m);
    double yr = 0.0, yi = 0.0;
    for (int t = 0; t < (dr ? 3 : 1); ++t) {
        double xr, xi;
        load_elem(p.x, p.dtype_x, dr ? p.tris[3 * i + t] : i, p.conj_x, xr, xi);
        yr += m[0][t] * xr; yi += m[0][t] * xi;
    }
    add_elem(p.y, p.dtype_x, i, p.ar * yr, p.ar * yi, false);
}

// P1 test rows: one thread per vertex over its panel corners in panel order.
__global__ void mass_apply_vertex_kernel(OpParams p, const double* mw, int dr) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= p.nv) return;
    double yr = 0.0, yi = 0.0;
    for (int c = p.vp_row[v]; c < p.vp_row[v + 1]; ++c) {
        const int i = p.vp_el[c] / 3, s = p.vp_el[c] % 3;
        double m[3][3];
        mass_block(p, mw, i, 1, dr, m);
        for (int t = 0; t < (dr ? 3 : 1); ++t) {
            double xr, xi;
            load_elem(p.x, p.dtype_x, dr ? p.tris[3 * i + t] : i, p.conj_x, xr, xi);
            yr += m[s][t] * xr; yi += m[s][t] * xi;
        }
    }
    add_elem(p.y, p.dtype_x, v, p.ar * yr, p.ar * yi, false);
}

OpParams make_params(Context* ctx, const bem_op_desc* d) {
    Mesh& m = ctx->mesh;
    OpParams p;
    p.nf = m.nf; p.nv = m.nv; p.q2 = m.q2; p.flags = d ? d->flags : 0;
    p.tris = m.tris; p.verts = m.verts; p.area = m.area; p.normal = m.normal; p.curl = m.curl;
    p.cen = m.cen; p.pu = m.pu; p.pl = m.pl; p.pu32 = m.pu32; p.pl32 = m.pl32;
    p.phi = m.phi; p.phi32 = m.phi32; p.frame = m.frame; p.pq = m.pq; p.pq32 = m.pq32; p.near = m.near; p.near_row = m.near_row; p.rule = m.rule;
    for (int r = 0; r < 4; ++r) { p.rule_off[r] = m.rule_off[r]; p.rule_len[r] = m.rule_len[r]; }
    p.n_near = m.n_near;
    p.kap = 0.0; p.k2n = 0.0; p.dkap = 0.0; p.dk2n = 0.0;
    if (d) {
        const bool tr = d->kernel == BEM_HELMHOLTZ_TRANSMISSION;
        p.kap = tr ? d->k * d->sqrt_n : d->k;
        p.k2n = d->k * d->k * d->sqrt_n * d->sqrt_n;
        p.dkap = tr ? d->sqrt_n : 1.0;
        p.dk2n = 2.0 * d->k * d->sqrt_n * d->sqrt_n;
    }
    p.out = nullptr; p.ld = 0; p.dtype = BEM_C128; p.ar = 1.0; p.ai = 0.0;
    p.x = nullptr; p.conj_x = 0; p.dtype_x = BEM_C128; p.y = nullptr; p.near_buf = m.near_buf;
    p.panel_buf = m.panel_buf; p.near_tr = m.near_tr; p.vp_row = m.vp_row; p.vp_el = m.vp_el;
    return p;
}

// ---------------------------------------------------------------------------
// Dirichlet system in one far-field pass (P0 test panels): V (P0 trial) into
// the matrix and the matrix-free K conj(g) (P1 trial) for the right-hand side.
// The tensor rule makes V symmetric (V_ij = V_ji: the same point pairs, G
// symmetric), so only pairs i < j are evaluated, and one evaluation of r, 1/r
// and exp(i kap r) per point pair yields
//   V_ij, written to A[i][j] and A[j][i],
//   the K conj(g) term of row i (normal n_j, data on panel j) and
//   the K conj(g) term of row j (normal n_i, data on panel i).
// A block takes a tile of DTI test panels (wave-uniform, scalar loads) and
// 256 trial panels j (one per lane), over the upper triangle only.  Row-i
// terms are reduced over the block into one partial per (test panel, trial
// block); row-j terms stay in the lane over the tile and are added once per
// lane (atomics).  A second kernel sums the partials in a fixed order.
// ---------------------------------------------------------------------------
constexpr int DTI = 8;             // test panels per tile
constexpr int DTG = 256 / DTI;     // tiles per 256-row group

// block b of the upper-triangle tiling -> (tile, trial block): the tiles of
// group g (rows [256 g, 256 g + 256)) pair with trial blocks g .. njb-1.
__device__ __forceinline__ void sym_block(int64_t b, int njb, int ngroups, int& tile, int& jb) {
    auto prefix = [&](int g) -> int64_t { return (int64_t)DTG * ((int64_t)g * njb - (int64_t)g * (g - 1) / 2); };
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix(mid) <= b) lo = mid; else hi = mid - 1;
    }
    const int64_t local = b - prefix(lo);
    const int cnt = njb - lo;
    tile = lo * DTG + (int)(local / cnt);
    jb = lo + (int)(local % cnt);
}

template <typename R, bool HELM>
__global__ __launch_bounds__(256) void far_dirichlet_kernel(OpParams p, int njb, int ngroups, double* kpart,
                                                            double* krow) {
    int tile, jb;
    sym_block(blockIdx.x, njb, ngroups, tile, jb);
    const int nf = p.nf, q2 = p.q2;
    const int i0 = tile * DTI;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int j = jb * 256 + threadIdx.x;
    const bool jvalid = j < nf;
    const int jj = jvalid ? j : 0;
    __shared__ double gts[DTI][144][2];  // data interpolated at the test points (Q^2 <= 144) of each tile panel
    __shared__ double red[4][2];
    __shared__ double2 sct[256];         // sin/cos(j pi/128) for the FP64 phase
    if (i0 >= nf) return;  // the last group's empty tiles (uniform)
    if constexpr (HELM && sizeof(R) == 8) sincos_table_fill(sct);
    for (int e = threadIdx.x; e < DTI * q2; e += 256) {
        const int ti = e / q2, a = e % q2, i = i0 + ti;
        double gr = 0.0, gi = 0.0;
        if (i < nf)
            for (int s = 0; s < 3; ++s) {
                double xr, xi;
                load_elem(p.x, p.dtype_x, p.tris[3 * i + s], p.conj_x, xr, xi);
                const double ph = p.phi[3 * a + s];
                gr += ph * xr; gi += ph * xi;
            }
        gts[ti][a][0] = gr;
        gts[ti][a][1] = gi;
    }
    __syncthreads();
    const R* __restrict__ PL = Ptr<R>::pl(p);
    const ConstPtr<R> PHI = as_const(Ptr<R>::phi(p));  // scalar loads
    const R kap = (R)p.kap;
    // trial-side data of this lane's panel j
    const R nnx = (R)p.normal[3 * jj], nny = (R)p.normal[3 * jj + 1], nnz = (R)p.normal[3 * jj + 2];
    R gjr[3], gji[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        double xr, xi;
        load_elem(p.x, p.dtype_x, p.tris[3 * jj + t], p.conj_x, xr, xi);
        gjr[t] = (R)xr; gji[t] = (R)xi;
    }
    double kjr = 0.0, kji = 0.0;  // row-j term, summed over the tile
    double vtr[DTI], vti[DTI];    // V_ij of the tile, for A[j][i0 + ti]
    unsigned done = 0u;           // bit ti: pair (i0 + ti, j) evaluated here
#pragma unroll
    for (int ti = 0; ti < DTI; ++ti) { vtr[ti] = 0.0; vti[ti] = 0.0; }
    const int i_end = min(nf, i0 + DTI);
#pragma unroll
    for (int ti = 0; ti < DTI; ++ti) {
        const int i = i0 + ti;
        double kir = 0.0, kii = 0.0;
        if (i < i_end && jvalid && j > i && !shares_vertex(p.tris, i, j)) {
            // wave-uniform test points, read with scalar loads
            const ConstPtr<R> PU = as_const(Ptr<R>::pu(p) + (int64_t)i * q2 * 4);
            R dcx = 0, dcy = 0, dcz = 0;
            if constexpr (sizeof(R) == 4) {
                dcx = (R)(p.cen[3 * i] - p.cen[3 * j]);
                dcy = (R)(p.cen[3 * i + 1] - p.cen[3 * j + 1]);
                dcz = (R)(p.cen[3 * i + 2] - p.cen[3 * j + 2]);
            }
            const R mix = (R)p.normal[3 * i], miy = (R)p.normal[3 * i + 1], miz = (R)p.normal[3 * i + 2];
            // (x - y).n_i = x_a.n_i + e_b.n_i, and x_a.n_i is the same for every
            // test point of the flat panel i: one dot per trial point, not per pair
            const R pun = PU[0] * mix + PU[1] * miy + PU[2] * miz;
            R vr = 0, vi = 0, ar = 0, ai = 0, br = 0, bi = 0;
            for (int b = 0; b < q2; ++b) {
                const R ex = dcx - PL[((int64_t)b * 4 + 0) * nf + j];
                const R ey = dcy - PL[((int64_t)b * 4 + 1) * nf + j];
                const R ez = dcz - PL[((int64_t)b * 4 + 2) * nf + j];
                const R wb = PL[((int64_t)b * 4 + 3) * nf + j];
                const R dni = pun + ex * mix + ey * miy + ez * miz;
                R pvr = 0, pvi = 0, pkr = 0, pki = 0, qkr = 0, qki = 0;
                for (int a = 0; a < q2; ++a) {
                    const R dx = PU[4 * a + 0] + ex;
                    const R dy = PU[4 * a + 1] + ey;
                    const R dz = PU[4 * a + 2] + ez;
                    const R wa = PU[4 * a + 3];
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    const R rinv = Math<R>::rsqrt(r2);
                    const R g = wa * rinv;
                    const R g3 = g * (rinv * rinv);
                    const R bj = g3 * (dx * nnx + dy * nny + dz * nnz);    // (x - y).n_j: row i
                    const R bi_ = -g3 * dni;  // (y - x).n_i: row j
                    const R gar = (R)gts[ti][a][0], gai = (R)gts[ti][a][1];
                    if constexpr (HELM) {
                        const R kr = kap * (r2 * rinv);
                        R sn, cs;
                        if constexpr (sizeof(R) == 8) sincos_f64_tab(kr, sct, &sn, &cs);
                        else Math<R>::sincos(kr, &sn, &cs);
                        pvr += g * cs; pvi += g * sn;
                        const R c1 = cs + kr * sn, c2 = sn - kr * cs;
                        pkr += bj * c1; pki += bj * c2;
                        // row j: bi_ (c1 + i c2) times the test-point data
                        const R tr = bi_ * c1, tim = bi_ * c2;
                        qkr += tr * gar - tim * gai;
                        qki += tr * gai + tim * gar;
                    } else {
                        pvr += g;
                        pkr += bj;
                        qkr += bi_ * gar;
                        qki += bi_ * gai;
                    }
                }
                vr += pvr * wb; vi += pvi * wb;
                const R ph0 = PHI[3 * b], ph1 = PHI[3 * b + 1], ph2 = PHI[3 * b + 2];
                const R xr = ph0 * gjr[0] + ph1 * gjr[1] + ph2 * gjr[2];
                const R xi = ph0 * gji[0] + ph1 * gji[1] + ph2 * gji[2];
                const R wr = pkr * wb, wi = pki * wb;
                ar += wr * xr - wi * xi;
                ai += wr * xi + wi * xr;
                br += qkr * wb; bi += qki * wb;
            }
            const double Vr = kInv4Pi * (double)vr, Vi = kInv4Pi * (double)vi;
            add_elem(p.out, p.dtype, (int64_t)i * p.ld + j, p.ar * Vr - p.ai * Vi, p.ar * Vi + p.ai * Vr, false);
            vtr[ti] = Vr; vti[ti] = Vi;
            done |= 1u << ti;
            kir = kInv4Pi * (double)ar; kii = kInv4Pi * (double)ai;
            kjr += kInv4Pi * (double)br; kji += kInv4Pi * (double)bi;
        }
        // row i: reduce over the block into the (i, jb) partial
        kir = wave_sum(kir);
        kii = wave_sum(kii);
        if (lane == 0) { red[wid][0] = kir; red[wid][1] = kii; }
        __syncthreads();
        if (threadIdx.x == 0 && i < nf) {
            kpart[((int64_t)i * njb + jb) * 2] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
            kpart[((int64_t)i * njb + jb) * 2 + 1] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
        }
        __syncthreads();
    }
    if (!jvalid) return;
    // the mirrored entries A[j][i0 .. i0 + DTI) of this lane (contiguous)
#pragma unroll
    for (int ti = 0; ti < DTI; ++ti) {
        const int i = i0 + ti;
        if (done & (1u << ti))
            add_elem(p.out, p.dtype, (int64_t)j * p.ld + i, p.ar * vtr[ti] - p.ai * vti[ti],
                     p.ar * vti[ti] + p.ai * vtr[ti], false);
    }
    if (kjr != 0.0 || kji != 0.0) {
        atomicAdd(krow + 2 * j, kjr);
        atomicAdd(krow + 2 * j + 1, kji);
    }
}

__global__ void zero_kernel(double* x, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = 0.0;
}

// rhs[i] += alpha (krow[i] + sum_{jb >= i / 256} kpart[i][jb]) (fixed order
// over the partials; krow holds the row's terms from the mirrored pairs).
__global__ void kpart_reduce_kernel(const double* kpart, const double* krow, int nf, int njb, void* y, int dtype,
                                    double ar, double ai) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    double a = krow[2 * i], b = krow[2 * i + 1];
    for (int q = i / 256; q < njb; ++q) {
        a += kpart[((int64_t)i * njb + q) * 2];
        b += kpart[((int64_t)i * njb + q) * 2 + 1];
    }
    add_elem(y, dtype, i, ar * a - ai * b, ar * b + ai * a, false);
}

// ---------------------------------------------------------------------------
// The FP64 Dirichlet pass in moment form.  Same pairs and outputs as
// far_dirichlet_kernel (V into A[i][j] and A[j][i], K conj(g) of both rows),
// with the test panel i in its local frame (Mesh::frame): a test point is
// x_a - C = U + r1_a E1 + r2_a E2, and with Y = y_b - C for the trial point
//   |x_a - y_b|^2   = S_a + Y.(Y + D) + p1_a (E1.Y) + p2_a (E2.Y)
//   (x_a - y_b).n_j = (U - Y).n_j - (p1_a (E1.n_j) + p2_a (E2.n_j)) / 2
//   (y_b - x_a).n_i = Y.n_i                      (x_a - C lies in panel i)
// (p1 = -2 r1, p2 = -2 r2, S_a = |x_a - C|^2, D = -2U: Mesh::pq).  So per
// point pair only r^2 (three operations), 1/r, exp(i kap r) and three
// complex moments over the test points,
//   T0 = sum_a t_a,  T1 = sum_a p1_a t_a,  T2 = sum_a p2_a t_a,
//   t_a = w_a (1 - i kap r) e^{i kap r} / r^3,
// are formed.  For Helmholtz the pair works in units of 1/|kap|: r^2 is
// formed as (kap r)^2 (S_a, Q, P1, P2 scaled by kap^2), so kap r is
// (kap r)^2 / (kap r) with no multiply by kap, and V and the moments come
// out divided by |kap| and |kap|^3, restored once per panel pair; kap < 0
// conjugates the point sums (exp(i kap r) = conj(exp(i |kap| r))), and
// kap = 0 takes the Laplace instance.  Both K rows follow from the moments
// per trial point: row i through
// the affine (x_a - y_b).n_j, row j through the P1 data on panel i, which is
// affine in (r1, r2) too (phi = 1 - r1, r1 - r2, r2).
// Blocks take DROWS test panels (DTPB tiles of DTI) of one 256-row group
// against 256 trial panels; the row-j terms are summed over the block's rows
// in the lane and stored once per (row slot, trial panel), so that every sum
// runs in a fixed order and the pass is bitwise reproducible.
// ---------------------------------------------------------------------------
constexpr int DTPB = 4;                // tiles per block
constexpr int DROWS = DTI * DTPB;      // test panels per block
constexpr int DSUB = 256 / DROWS;      // row slots per 256-row group

// block b -> (group g, row slot sb, trial block jb): the blocks of group g
// pair each of its DSUB row slots with trial blocks g .. njb-1.
__device__ __forceinline__ void sym_block_rows(int64_t b, int njb, int ngroups, int& g, int& sb, int& jb) {
    auto prefix = [&](int gg) -> int64_t { return (int64_t)DSUB * ((int64_t)gg * njb - (int64_t)gg * (gg - 1) / 2); };
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix(mid) <= b) lo = mid; else hi = mid - 1;
    }
    const int64_t local = b - prefix(lo);
    const int cnt = njb - lo;
    g = lo;
    sb = (int)(local / cnt);
    jb = lo + (int)(local % cnt);
}

template <bool HELM, int NQ2>
__global__ __launch_bounds__(256) void far_dirichlet_f64_kernel(OpParams p, int njb, int ngroups, double* kpart,
                                                                double* jpart) {
    int g, sb, jb;
    sym_block_rows(blockIdx.x, njb, ngroups, g, sb, jb);
    const int nf = p.nf;
    const int q2 = NQ2 > 0 ? NQ2 : p.q2;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int j = jb * 256 + threadIdx.x;
    const bool jvalid = j < nf;
    const int jj = jvalid ? j : 0;
    const int r0 = g * 256 + sb * DROWS;
    __shared__ double red[4][2];
    __shared__ double2 sct[HELM ? kSinCosTab2048 : 1];  // {sin, cos}(2 pi j / 2048)
    if (r0 >= nf) {               // an empty row slot of the last group (uniform):
        if (jvalid) {             // its partials are read all the same
            jpart[((int64_t)(g * DSUB + sb) * nf + j) * 2] = 0.0;
            jpart[((int64_t)(g * DSUB + sb) * nf + j) * 2 + 1] = 0.0;
        }
        return;
    }
    if constexpr (HELM) sincos_table2048_fill(sct);
    __syncthreads();
    const double* __restrict__ PL = p.pl;
    const ConstPtr<double> PHI = as_const(p.phi);
    const double kap = p.kap;
    const double k2 = HELM ? kap * kap : 1.0;            // scale of r^2
    const double ka = HELM ? __builtin_fabs(kap) : 1.0;  // V carries 1/ka, the moments 1/ka^3
    // trial side of this lane's panel j
    const double nx = p.normal[3 * jj], ny = p.normal[3 * jj + 1], nz = p.normal[3 * jj + 2];
    double gjr[3], gji[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) load_elem(p.x, p.dtype_x, p.tris[3 * jj + t], p.conj_x, gjr[t], gji[t]);
    double kjr = 0.0, kji = 0.0;  // row-j terms of this block
    for (int tile = 0; tile < DTPB; ++tile) {
        const int i0 = r0 + tile * DTI;
        if (i0 >= nf) break;  // uniform
        double vtr[DTI], vti[DTI];  // V_ij of the tile, for A[j][i0 + ti]
        unsigned done = 0u;
#pragma unroll
        for (int ti = 0; ti < DTI; ++ti) { vtr[ti] = 0.0; vti[ti] = 0.0; }
#pragma unroll 1
        for (int ti = 0; ti < DTI; ++ti) {
            const int i = i0 + ti;
            double kir = 0.0, kii = 0.0;
            if (i < nf && jvalid && j > i && !shares_vertex(p.tris, i, j)) {
                // test panel frame and points: wave-uniform, scalar loads
                const ConstPtr<double> FR = as_const(p.frame + (int64_t)i * PANEL_FRAME);
                const ConstPtr<double> PQ = as_const(p.pq + (int64_t)i * q2 * 4);
                const double cx = FR[0], cy = FR[1], cz = FR[2];
                const double dx = FR[15], dy = FR[16], dz = FR[17];
                const double e1x = FR[6], e1y = FR[7], e1z = FR[8];
                const double e2x = FR[9], e2y = FR[10], e2z = FR[11];
                const double mix = FR[12], miy = FR[13], miz = FR[14];
                const double un = FR[3] * nx + FR[4] * ny + FR[5] * nz;  // U.n_j
                const double h1 = -0.5 * (e1x * nx + e1y * ny + e1z * nz);
                const double h2 = -0.5 * (e2x * nx + e2y * ny + e2z * nz);
                double vr = 0, vi = 0, ar = 0, ai = 0;
                double b0r = 0, b0i = 0, b1r = 0, b1i = 0, b2r = 0, b2i = 0;
#pragma unroll 1
                for (int b = 0; b < q2; ++b) {
                    const double yx = PL[((int64_t)b * 4 + 0) * nf + j] - cx;
                    const double yy = PL[((int64_t)b * 4 + 1) * nf + j] - cy;
                    const double yz = PL[((int64_t)b * 4 + 2) * nf + j] - cz;
                    const double wb = PL[((int64_t)b * 4 + 3) * nf + j];
                    const double P1 = k2 * (e1x * yx + e1y * yy + e1z * yz);
                    const double P2 = k2 * (e2x * yx + e2y * yy + e2z * yz);
                    const double Q = k2 * (yx * (yx + dx) + yy * (yy + dy) + yz * (yz + dz));
                    double pvr = 0, pvi = 0, t0r = 0, t0i = 0, t1r = 0, t1i = 0, t2r = 0, t2i = 0;
#pragma unroll 5
                    for (int a = 0; a < (NQ2 > 0 ? NQ2 : q2); ++a) {
                        const double p1 = PQ[4 * a + 1], p2 = PQ[4 * a + 2], wa = PQ[4 * a + 3];
                        const double r2 = HELM ? __builtin_fma(p2, P2, __builtin_fma(p1, P1, __builtin_fma(PQ[4 * a], k2, Q)))
                                               : __builtin_fma(p2, P2, __builtin_fma(p1, P1, PQ[4 * a] + Q));
                        const double rinv = rsqrt_f64(r2);
                        const double gv = wa * rinv;
                        if constexpr (HELM) {
                            const double th = r2 * rinv;  // |kap| r
                            double sn, cs;
                            sincos_f64_t2048(th, sct, &sn, &cs);
                            pvr = __builtin_fma(gv, cs, pvr);
                            pvi = __builtin_fma(gv, sn, pvi);
                            // (1 - i th) e^{i th} / th^3 with th rinv = 1:
                            // (sin th + cos th / th - i (cos th - sin th / th)) / th^2
                            const double g2 = gv * rinv;
                            const double tr = g2 * __builtin_fma(cs, rinv, sn);
                            const double tim = g2 * __builtin_fma(sn, rinv, -cs);
                            t0r += tr; t0i += tim;
                            t1r = __builtin_fma(p1, tr, t1r); t1i = __builtin_fma(p1, tim, t1i);
                            t2r = __builtin_fma(p2, tr, t2r); t2i = __builtin_fma(p2, tim, t2i);
                        } else {
                            const double g3 = gv * (rinv * rinv);
                            pvr += gv;
                            t0r += g3;
                            t1r = __builtin_fma(p1, g3, t1r);
                            t2r = __builtin_fma(p2, g3, t2r);
                        }
                    }
                    if (HELM && kap < 0.0) {  // uniform
                        pvi = -pvi; t0i = -t0i; t1i = -t1i; t2i = -t2i;
                    }
                    vr = __builtin_fma(wb, pvr, vr);
                    vi = __builtin_fma(wb, pvi, vi);
                    // row i: sum_a t_a (x_a - y_b).n_j times the data at y_b
                    const double c0 = un - (yx * nx + yy * ny + yz * nz);
                    const double kr = __builtin_fma(h2, t2r, __builtin_fma(h1, t1r, c0 * t0r));
                    const double ki = __builtin_fma(h2, t2i, __builtin_fma(h1, t1i, c0 * t0i));
                    const double ph0 = PHI[3 * b], ph1 = PHI[3 * b + 1], ph2 = PHI[3 * b + 2];
                    const double xr = ph0 * gjr[0] + ph1 * gjr[1] + ph2 * gjr[2];
                    const double xi = ph0 * gji[0] + ph1 * gji[1] + ph2 * gji[2];
                    const double wr = wb * kr, wi = wb * ki;
                    ar = __builtin_fma(wr, xr, ar); ar = __builtin_fma(-wi, xi, ar);
                    ai = __builtin_fma(wr, xi, ai); ai = __builtin_fma(wi, xr, ai);
                    // row j: w_b (y_b - x_a).n_i, the same for every test point
                    const double sw = wb * (yx * mix + yy * miy + yz * miz);
                    b0r = __builtin_fma(sw, t0r, b0r); b0i = __builtin_fma(sw, t0i, b0i);
                    b1r = __builtin_fma(sw, t1r, b1r); b1i = __builtin_fma(sw, t1i, b1i);
                    b2r = __builtin_fma(sw, t2r, b2r); b2i = __builtin_fma(sw, t2i, b2i);
                }
                const double sv = kInv4Pi * ka, sk = kInv4Pi * (ka * ka * ka);
                const double Vr = sv * vr, Vi = sv * vi;
                add_elem(p.out, p.dtype, (int64_t)i * p.ld + j, p.ar * Vr - p.ai * Vi, p.ar * Vi + p.ai * Vr, false);
                vtr[ti] = Vr; vti[ti] = Vi;
                done |= 1u << ti;
                kir = sk * ar; kii = sk * ai;
                // row j: the data on panel i at its test points is
                // g0 (1 + p1/2) + g1 (p2 - p1)/2 - g2 p2/2
                double g0r, g0i, g1r, g1i, g2r, g2i;
                load_elem(p.x, p.dtype_x, p.tris[3 * i], p.conj_x, g0r, g0i);
                load_elem(p.x, p.dtype_x, p.tris[3 * i + 1], p.conj_x, g1r, g1i);
                load_elem(p.x, p.dtype_x, p.tris[3 * i + 2], p.conj_x, g2r, g2i);
                const double c0r = b0r + 0.5 * b1r, c0i = b0i + 0.5 * b1i;  // coefficient of g0
                const double c1r = 0.5 * (b2r - b1r), c1i = 0.5 * (b2i - b1i);
                const double c2r = -0.5 * b2r, c2i = -0.5 * b2i;
                const double sr = c0r * g0r - c0i * g0i + c1r * g1r - c1i * g1i + c2r * g2r - c2i * g2i;
                const double si = c0r * g0i + c0i * g0r + c1r * g1i + c1i * g1r + c2r * g2i + c2i * g2r;
                kjr += sk * sr;
                kji += sk * si;
            }
            // row i: reduce over the block into the (i, jb) partial
            kir = wave_sum(kir);
            kii = wave_sum(kii);
            if (lane == 0) { red[wid][0] = kir; red[wid][1] = kii; }
            __syncthreads();
            if (threadIdx.x == 0 && i < nf) {
                kpart[((int64_t)i * njb + jb) * 2] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
                kpart[((int64_t)i * njb + jb) * 2 + 1] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
            }
            __syncthreads();
        }
        if (jvalid) {
            // the mirrored entries A[j][i0 .. i0 + DTI) of this lane (contiguous)
#pragma unroll
            for (int ti = 0; ti < DTI; ++ti)
                if (done & (1u << ti))
                    add_elem(p.out, p.dtype, (int64_t)j * p.ld + i0 + ti, p.ar * vtr[ti] - p.ai * vti[ti],
                             p.ar * vti[ti] + p.ai * vtr[ti], false);
        }
    }
    if (jvalid) {
        // one store per (row slot, trial panel): every slot with g <= jb is written
        double* o = jpart + ((int64_t)(g * DSUB + sb) * nf + j) * 2;
        o[0] = kjr;
        o[1] = kji;
    }
}

// rhs[i] += alpha (sum_{jb >= i/256} kpart[i][jb] + sum over the row slots
// of the groups g <= i/256 of jpart[slot][i]), in a fixed order.
__global__ void kpart_jpart_reduce_kernel(const double* kpart, const double* jpart, int nf, int njb, void* y,
                                          int dtype, double ar, double ai) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    double a = 0.0, b = 0.0;
    for (int q = i / 256; q < njb; ++q) {
        a += kpart[((int64_t)i * njb + q) * 2];
        b += kpart[((int64_t)i * njb + q) * 2 + 1];
    }
    const int nslot = (i / 256 + 1) * DSUB;
    for (int q = 0; q < nslot; ++q) {
        a += jpart[((int64_t)q * nf + i) * 2];
        b += jpart[((int64_t)q * nf + i) * 2 + 1];
    }
    add_elem(y, dtype, i, ar * a - ai * b, ar * b + ai * a, false);
}

// ---------------------------------------------------------------------------
// The blocks of the Calderon (transmission) system of one wave number in one
// far-field pass: V (P0 x P0), K (P0 test x P1 trial) and W (P1 x P1).  The
// tensor rule evaluates the same point pairs for (i, j) and (j, i), so only
// pairs i < j are visited, and one evaluation of r, 1/r and exp(i kap r) per
// point pair yields
//   V_ij, stored to V[i][j] and V[j][i];
//   K[i][v] for the vertices v of panel j (normal n_j, trial bases on j) and
//   K[j][u] for the vertices u of panel i (normal n_i, trial bases on i);
//   the W entries of the pair, W_ij[s][t] = W_ji[t][s] (G and the curl and
//   normal products are symmetric): the curl term needs only the V integral,
//   and the k^2 term the P1 x P1 integrals of G.
// W entries are added once, at (min(u, v), max(u, v)), and a second kernel
// copies the upper triangle to the lower one.  Regular pairs share no
// vertex, so they never reach the diagonal; the near field, added after
// the copy, brings both orders of its pairs itself.  Blocks tile the upper
// triangle like the Dirichlet pass (DTI test panels, wave-uniform, times 256
// trial panels, one per lane).
// ---------------------------------------------------------------------------
struct VkwOut {
    void* V; int64_t ldv;
    void* K; int64_t ldk;
    void* W; int64_t ldw;
};

// NQ2 > 0 fixes the rule size so the test-point loop unrolls (batched
// scalar loads of the test points).
template <typename R, bool HELM, int NQ2 = 0>
__global__ __launch_bounds__(256) void far_vkw_kernel(OpParams p, VkwOut o, int njb, int ngroups) {
    int tile, jb;
    sym_block(blockIdx.x, njb, ngroups, tile, jb);
    const int nf = p.nf, q2 = NQ2 > 0 ? NQ2 : p.q2;
    const int i0 = tile * DTI;
    const int j = jb * 256 + threadIdx.x;
    // FP64 Helmholtz: the pair sums in units of 1/|kap| (r^2 formed as
    // (kap r)^2, so kap r needs no multiply by kap; V-type sums rescaled by
    // |kap| and K-type ones by |kap|^3 once per panel pair; kap < 0
    // conjugates the point sums) and sin/cos from the 2048-entry table
    constexpr bool KS = HELM && sizeof(R) == 8;
    __shared__ double2 sct[KS ? kSinCosTab2048 : 1];  // {sin, cos}(2 pi j / 2048)
    if (i0 >= nf) return;  // the last group's empty tiles (uniform)
    if constexpr (KS) {
        sincos_table2048_fill(sct);
        __syncthreads();
    }
    if (j >= nf) return;  // no barrier below
    const R* __restrict__ PL = Ptr<R>::pl(p);
    const ConstPtr<R> PHI = as_const(Ptr<R>::phi(p));  // scalar loads
    const R kap = (R)p.kap;
    const R k2 = KS ? kap * kap : R(1), ka = KS ? (R)fabs((double)kap) : R(1);
    const R nnx = (R)p.normal[3 * j], nny = (R)p.normal[3 * j + 1], nnz = (R)p.normal[3 * j + 2];
    const int64_t vj[3] = {p.tris[3 * j], p.tris[3 * j + 1], p.tris[3 * j + 2]};
    const int i_end = min(nf, i0 + DTI);
    for (int i = i0; i < i_end; ++i) {
        if (j <= i || shares_vertex(p.tris, i, j)) continue;
        // Moment form (the test panel in its local frame, as in the Dirichlet
        // pass): with Y = y_b - C_i and x_a - C_i = U + r1 E1 + r2 E2,
        //   |x_a - y_b|^2   = S_a + Y.(Y + D) + p1 (E1.Y) + p2 (E2.Y),
        //   (x_a - y_b).n_j = (U - Y).n_j - p1 (E1.n_j)/2 - p2 (E2.n_j)/2,
        //   (y_b - x_a).n_i = Y.n_i,
        // (S_a, p1, p2, w_a) from the per-point table (scalar loads).
        const ConstPtr<R> PQ = as_const(Ptr<R>::pq(p) + (int64_t)i * q2 * 4);
        const ConstPtr<R> PU = as_const(Ptr<R>::pu(p) + (int64_t)i * q2 * 4);  // FP32: offsets from C_i
        const ConstPtr<double> FR = as_const(p.frame + (int64_t)i * PANEL_FRAME);
        R U[3], E1[3], E2[3], D[3], ni[3], oY[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            U[d] = (R)FR[3 + d];
            E1[d] = (R)FR[6 + d];
            E2[d] = (R)FR[9 + d];
            ni[d] = (R)FR[12 + d];
            D[d] = (R)FR[15 + d];
            // FP64 trial points are absolute (Y = y - C_i); FP32 ones are
            // offsets from their centroid (Y = (y - C_j) + (C_j - C_i))
            oY[d] = sizeof(R) == 8 ? (R)(-FR[d]) : (R)(p.cen[3 * j + d] - p.cen[3 * i + d]);
        }
        const R e1n = R(-0.5) * (E1[0] * nnx + E1[1] * nny + E1[2] * nnz);
        const R e2n = R(-0.5) * (E2[0] * nnx + E2[1] * nny + E2[2] * nnz);
        const R un = U[0] * nnx + U[1] * nny + U[2] * nnz;
        R kar[3], kai[3], kbr[3], kbi[3], wr[3][3], wi[3][3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            kar[s] = kai[s] = kbr[s] = kbi[s] = 0;
#pragma unroll
            for (int t = 0; t < 3; ++t) wr[s][t] = wi[s][t] = 0;
        }
        for (int b = 0; b < q2; ++b) {
            const R yx = PL[((int64_t)b * 4 + 0) * nf + j] + oY[0];
            const R yy = PL[((int64_t)b * 4 + 1) * nf + j] + oY[1];
            const R yz = PL[((int64_t)b * 4 + 2) * nf + j] + oY[2];
            const R wb = PL[((int64_t)b * 4 + 3) * nf + j];
            const R yd = k2 * (yx * (yx + D[0]) + yy * (yy + D[1]) + yz * (yz + D[2]));
            const R y1 = k2 * (E1[0] * yx + E1[1] * yy + E1[2] * yz);
            const R y2 = k2 * (E2[0] * yx + E2[1] * yy + E2[2] * yz);
            const R c0 = un - (yx * nnx + yy * nny + yz * nnz);
            const R dyi = yx * ni[0] + yy * ni[1] + yz * ni[2];  // (y - x).n_i
            R pwr[3], pwi[3], qkr[3], qki[3], pkr = 0, pki = 0;
#pragma unroll
            for (int s = 0; s < 3; ++s) pwr[s] = pwi[s] = qkr[s] = qki[s] = 0;
#pragma unroll 5
            for (int a = 0; a < q2; ++a) {
                R r2, wa, dnj;
                if constexpr (sizeof(R) == 8) {
                    const R sa = PQ[4 * a + 0], p1 = PQ[4 * a + 1], p2 = PQ[4 * a + 2];
                    wa = PQ[4 * a + 3];
                    r2 = p2 * y2 + (p1 * y1 + (KS ? __builtin_fma(sa, k2, yd) : sa + yd));
                    dnj = p2 * e2n + (p1 * e1n + c0);
                } else {
                    // FP32 keeps the direct form (its K rows lost digits in the
                    // moment form at no gain: the FP32 pass is atomic-bound)
                    const R dx = PU[4 * a + 0] - yx, dy = PU[4 * a + 1] - yy, dz = PU[4 * a + 2] - yz;
                    wa = PU[4 * a + 3];
                    r2 = dx * dx + dy * dy + dz * dz;
                    dnj = dx * nnx + dy * nny + dz * nnz;
                }
                const R rinv = Math<R>::rsqrt(r2);
                const R g = wa * rinv;
                const R g3 = g * (rinv * rinv);
                const R bj = g3 * dnj;   // (x - y).n_j: K row i
                const R bi_ = g3 * dyi;  // (y - x).n_i: K row j
                if constexpr (HELM) {
                    const R kr = KS ? r2 * rinv : kap * (r2 * rinv);  // |kap| r (KS) or kap r
                    R sn, cs;
                    if constexpr (KS) sincos_f64_t2048(kr, sct, &sn, &cs);
                    else Math<R>::sincos(kr, &sn, &cs);
                    const R gr = g * cs, gi = g * sn;
                    const R c1 = cs + kr * sn, c2 = sn - kr * cs;
                    pkr += bj * c1; pki += bj * c2;
                    const R tr = bi_ * c1, tim = bi_ * c2;
#pragma unroll
                    for (int s = 0; s < 3; ++s) {
                        const R ph = PHI[3 * a + s];
                        pwr[s] += gr * ph; pwi[s] += gi * ph;
                        qkr[s] += tr * ph; qki[s] += tim * ph;
                    }
                } else {
                    pkr += bj;
#pragma unroll
                    for (int s = 0; s < 3; ++s) {
                        const R ph = PHI[3 * a + s];
                        pwr[s] += g * ph;
                        qkr[s] += bi_ * ph;
                    }
                }
            }
            if (KS && kap < R(0)) {  // uniform
                pki = -pki;
#pragma unroll
                for (int s = 0; s < 3; ++s) { pwi[s] = -pwi[s]; qki[s] = -qki[s]; }
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const R wt = wb * PHI[3 * b + t];
                kar[t] += pkr * wt; kai[t] += pki * wt;
#pragma unroll
                for (int s = 0; s < 3; ++s) { wr[s][t] += pwr[s] * wt; wi[s][t] += pwi[s] * wt; }
            }
#pragma unroll
            for (int s = 0; s < 3; ++s) { kbr[s] += qkr[s] * wb; kbi[s] += qki[s] * wb; }
        }
        if constexpr (KS) {
            const R ka3 = ka * ka * ka;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                kar[s] *= ka3; kai[s] *= ka3; kbr[s] *= ka3; kbi[s] *= ka3;
#pragma unroll
                for (int t = 0; t < 3; ++t) { wr[s][t] *= ka; wi[s][t] *= ka; }
            }
        }
        // the V integral: the P1 bases sum to one on either panel
        double s0r = 0.0, s0i = 0.0;
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int t = 0; t < 3; ++t) { s0r += (double)wr[s][t]; s0i += (double)wi[s][t]; }
        const double Vr = kInv4Pi * s0r, Vi = kInv4Pi * s0i;
        add_elem(o.V, p.dtype, (int64_t)i * o.ldv + j, Vr, Vi, false);
        add_elem(o.V, p.dtype, (int64_t)j * o.ldv + i, Vr, Vi, false);
        const int64_t vi[3] = {p.tris[3 * i], p.tris[3 * i + 1], p.tris[3 * i + 2]};
#pragma unroll
        for (int t = 0; t < 3; ++t)
            add_elem(o.K, p.dtype, (int64_t)i * o.ldk + vj[t], kInv4Pi * (double)kar[t], kInv4Pi * (double)kai[t],
                     true);
#pragma unroll
        for (int s = 0; s < 3; ++s)
            add_elem(o.K, p.dtype, (int64_t)j * o.ldk + vi[s], kInv4Pi * (double)kbr[s], kInv4Pi * (double)kbi[s],
                     true);
        const double nxy = p.normal[3 * i] * p.normal[3 * j] + p.normal[3 * i + 1] * p.normal[3 * j + 1] +
                           p.normal[3 * i + 2] * p.normal[3 * j + 2];
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const double cc = p.curl[9 * i + 3 * s] * p.curl[9 * j + 3 * t] +
                                  p.curl[9 * i + 3 * s + 1] * p.curl[9 * j + 3 * t + 1] +
                                  p.curl[9 * i + 3 * s + 2] * p.curl[9 * j + 3 * t + 2];
                const double er = kInv4Pi * (cc * s0r - p.k2n * nxy * (double)wr[s][t]);
                const double ei = kInv4Pi * (cc * s0i - p.k2n * nxy * (double)wi[s][t]);
                const int64_t lo = min(vi[s], vj[t]), hi = max(vi[s], vj[t]);
                add_elem(o.W, p.dtype, lo * o.ldw + hi, er, ei, true);
            }
    }
}

// W[hi][lo] = W[lo][hi] for lo < hi, in 32 x 32 tiles through LDS: block
// (tr, tc) with tr >= tc writes the strict lower part of tile row tr,
// column tc from the transposed upper tile.  E is the element type.
template <typename E>
__global__ __launch_bounds__(256) void mirror_upper_kernel(E* W, int64_t ld, int n, int ntiles) {
    // block -> (tr, tc), tc <= tr, row-major over the lower triangle of tiles
    const int64_t b = blockIdx.x;
    int tr = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((int64_t)tr * (tr + 1) / 2 > b) --tr;
    while ((int64_t)(tr + 1) * (tr + 2) / 2 <= b) ++tr;
    const int tc = (int)(b - (int64_t)tr * (tr + 1) / 2);
    if (tr >= ntiles) return;
    __shared__ E t[32][33];
    const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
    // load the upper tile: rows tc*32 + y, cols tr*32 + x
    for (int y = y0; y < 32; y += 8) {
        const int r = tc * 32 + y, c = tr * 32 + x;
        if (r < n && c < n && r < c) t[y][x] = W[(int64_t)r * ld + c];
    }
    __syncthreads();
    for (int y = y0; y < 32; y += 8) {
        const int r = tr * 32 + y, c = tc * 32 + x;
        if (r < n && c < n && r > c) W[(int64_t)r * ld + c] = t[x][y];
    }
}

// ---------------------------------------------------------------------------
// The k-derivative bilinear forms of the Calderon blocks over the regular
// pairs, without forming a matrix or a vector:
//   c0 = aN^T dV bN,  c1 = aN^T dK bD,  c2 = bN^T dK aD,  c3 = aD^T dW bD
// (aN, bN on the P0 panels, aD, bD on the P1 vertices, plain products, no
// conjugation).  Tiling and the single evaluation per unordered point pair
// are those of far_vkw_kernel, with the k-derivative integrands
//   dG/dk       = dkap i exp(i kap r) / (4 pi),
//   d(dG/dn)/dk = dkap kap ((x - y).n) exp(i kap r) / (4 pi r),
//   dW/dk       = dG/dk curl.curl - (dk2n G + k2n dG/dk) phi phi n_x.n_y.
// Each pair is contracted with the vectors in both orders, each lane keeps
// the four sums, and the block writes its reduction to part[block][8].
// ---------------------------------------------------------------------------
struct DkVecs {
    const double* aN;  // complex, [nf]
    const double* bN;
    const double* aD;  // complex, [nv]
    const double* bD;
};

__device__ __forceinline__ void cmac(double& zr, double& zi, double xr, double xi, double yr, double yi) {
    zr += xr * yr - xi * yi;
    zi += xr * yi + xi * yr;
}

template <int NQ2 = 0>
__global__ __launch_bounds__(256) void far_vkw_dk_kernel(OpParams p, DkVecs v, int njb, int ngroups, double* part) {
    int tile, jb;
    sym_block(blockIdx.x, njb, ngroups, tile, jb);
    const int nf = p.nf, q2 = NQ2 > 0 ? NQ2 : p.q2;
    const int i0 = tile * DTI;
    const int j = jb * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __shared__ double2 sct[256];
    __shared__ double red[4][8];
    if (i0 >= nf) {  // the last group's empty tiles (uniform)
        if (threadIdx.x < 8) part[(int64_t)blockIdx.x * 8 + threadIdx.x] = 0.0;
        return;
    }
    sincos_table_fill(sct);
    __syncthreads();
    double c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) c[q] = 0.0;
    if (j < nf) {
        const double* __restrict__ PL = p.pl;
        const ConstPtr<double> PHI = as_const(p.phi);  // scalar loads
        const double kap = p.kap, dkap = p.dkap, k2n = p.k2n, dk2n = p.dk2n, kd = p.kap * p.dkap;
        const double nnx = p.normal[3 * j], nny = p.normal[3 * j + 1], nnz = p.normal[3 * j + 2];
        const int vj[3] = {p.tris[3 * j], p.tris[3 * j + 1], p.tris[3 * j + 2]};
        const double aNj[2] = {v.aN[2 * j], v.aN[2 * j + 1]}, bNj[2] = {v.bN[2 * j], v.bN[2 * j + 1]};
        double aDj[3][2], bDj[3][2];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            aDj[t][0] = v.aD[2 * vj[t]]; aDj[t][1] = v.aD[2 * vj[t] + 1];
            bDj[t][0] = v.bD[2 * vj[t]]; bDj[t][1] = v.bD[2 * vj[t] + 1];
        }
        const int i_end = min(nf, i0 + DTI);
        for (int i = i0; i < i_end; ++i) {
            if (j <= i || shares_vertex(p.tris, i, j)) continue;
            const ConstPtr<double> PU = as_const(p.pu + (int64_t)i * q2 * 4);  // wave-uniform test points
            const double mix = p.normal[3 * i], miy = p.normal[3 * i + 1], miz = p.normal[3 * i + 2];
            const double pun = PU[0] * mix + PU[1] * miy + PU[2] * miz;
            double vr = 0, vi = 0, kar[3], kai[3], kbr[3], kbi[3], wr[3][3], wi[3][3];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                kar[s] = kai[s] = kbr[s] = kbi[s] = 0;
#pragma unroll
                for (int t = 0; t < 3; ++t) wr[s][t] = wi[s][t] = 0;
            }
            for (int b = 0; b < q2; ++b) {
                const double ex = -PL[((int64_t)b * 4 + 0) * nf + j];
                const double ey = -PL[((int64_t)b * 4 + 1) * nf + j];
                const double ez = -PL[((int64_t)b * 4 + 2) * nf + j];
                const double wb = PL[((int64_t)b * 4 + 3) * nf + j];
                const double dni = pun + ex * mix + ey * miy + ez * miz;
                double pvr = 0, pvi = 0, pkr = 0, pki = 0, pwr[3], pwi[3], qkr[3], qki[3];
#pragma unroll
                for (int s = 0; s < 3; ++s) pwr[s] = pwi[s] = qkr[s] = qki[s] = 0;
#pragma unroll 5
                for (int a = 0; a < q2; ++a) {
                    const double dx = PU[4 * a + 0] + ex;
                    const double dy = PU[4 * a + 1] + ey;
                    const double dz = PU[4 * a + 2] + ez;
                    const double wa = PU[4 * a + 3];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double rinv = rsqrt_f64(r2);
                    const double kr = kap * (r2 * rinv);
                    double sn, cs;
                    sincos_f64_tab(kr, sct, &sn, &cs);
                    const double wd = wa * dkap;
                    const double f1r = -wd * sn, f1i = wd * cs;  // dG/dk (times 4 pi / w_b)
                    pvr += f1r; pvi += f1i;
                    const double g = wa * rinv;
                    const double hr = dk2n * g * cs + k2n * f1r, hi = dk2n * g * sn + k2n * f1i;
                    const double gk = g * kd;
                    const double bj = gk * (dx * nnx + dy * nny + dz * nnz);  // (x - y).n_j
                    const double bi_ = -gk * dni;                              // (y - x).n_i
                    pkr += bj * cs; pki += bj * sn;
                    const double tr = bi_ * cs, tim = bi_ * sn;
#pragma unroll
                    for (int s = 0; s < 3; ++s) {
                        const double ph = PHI[3 * a + s];
                        pwr[s] += hr * ph; pwi[s] += hi * ph;
                        qkr[s] += tr * ph; qki[s] += tim * ph;
                    }
                }
                vr += pvr * wb; vi += pvi * wb;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const double wt = wb * PHI[3 * b + t];
                    kar[t] += pkr * wt; kai[t] += pki * wt;
#pragma unroll
                    for (int s = 0; s < 3; ++s) { wr[s][t] += pwr[s] * wt; wi[s][t] += pwi[s] * wt; }
                }
#pragma unroll
                for (int s = 0; s < 3; ++s) { kbr[s] += qkr[s] * wb; kbi[s] += qki[s] * wb; }
            }
            // contraction of the pair in both orders
            const int vi_[3] = {p.tris[3 * i], p.tris[3 * i + 1], p.tris[3 * i + 2]};
            const double aNi[2] = {v.aN[2 * i], v.aN[2 * i + 1]}, bNi[2] = {v.bN[2 * i], v.bN[2 * i + 1]};
            double aDi[3][2], bDi[3][2];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                aDi[s][0] = v.aD[2 * vi_[s]]; aDi[s][1] = v.aD[2 * vi_[s] + 1];
                bDi[s][0] = v.bD[2 * vi_[s]]; bDi[s][1] = v.bD[2 * vi_[s] + 1];
            }
            const double dvr = kInv4Pi * vr, dvi = kInv4Pi * vi;
            // c0: aN_i dV bN_j + aN_j dV bN_i
            double xr = 0, xi = 0;
            cmac(xr, xi, aNi[0], aNi[1], bNj[0], bNj[1]);
            cmac(xr, xi, aNj[0], aNj[1], bNi[0], bNi[1]);
            cmac(c[0], c[1], dvr, dvi, xr, xi);
            // c1, c2: row i against panel j's vertices, row j against panel i's
            double s1r = 0, s1i = 0, s2r = 0, s2i = 0, u1r = 0, u1i = 0, u2r = 0, u2i = 0;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                cmac(s1r, s1i, kar[t], kai[t], bDj[t][0], bDj[t][1]);
                cmac(s2r, s2i, kar[t], kai[t], aDj[t][0], aDj[t][1]);
                cmac(u1r, u1i, kbr[t], kbi[t], bDi[t][0], bDi[t][1]);
                cmac(u2r, u2i, kbr[t], kbi[t], aDi[t][0], aDi[t][1]);
            }
            double yr = 0, yi = 0, zr = 0, zi = 0;
            cmac(yr, yi, aNi[0], aNi[1], s1r, s1i);
            cmac(yr, yi, aNj[0], aNj[1], u1r, u1i);
            cmac(zr, zi, bNi[0], bNi[1], s2r, s2i);
            cmac(zr, zi, bNj[0], bNj[1], u2r, u2i);
            c[2] += kInv4Pi * yr; c[3] += kInv4Pi * yi;
            c[4] += kInv4Pi * zr; c[5] += kInv4Pi * zi;
            // c3: sum_st dW[s][t] (aD_is bD_jt + aD_jt bD_is)
            const double nxy = mix * nnx + miy * nny + miz * nnz;
            double wsr = 0, wsi = 0;
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const double cc = p.curl[9 * i + 3 * s] * p.curl[9 * j + 3 * t] +
                                      p.curl[9 * i + 3 * s + 1] * p.curl[9 * j + 3 * t + 1] +
                                      p.curl[9 * i + 3 * s + 2] * p.curl[9 * j + 3 * t + 2];
                    const double er = cc * vr - nxy * wr[s][t], ei = cc * vi - nxy * wi[s][t];
                    double mr = 0, mi = 0;
                    cmac(mr, mi, aDi[s][0], aDi[s][1], bDj[t][0], bDj[t][1]);
                    cmac(mr, mi, aDj[t][0], aDj[t][1], bDi[s][0], bDi[s][1]);
                    cmac(wsr, wsi, er, ei, mr, mi);
                }
            c[6] += kInv4Pi * wsr; c[7] += kInv4Pi * wsi;
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double t = wave_sum(c[q]);
        if (lane == 0) red[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < 8)
        part[(int64_t)blockIdx.x * 8 + threadIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out[0..8) = the sums of part[nparts][8] (fixed order) plus the near-field
// dots x_q . y_q (complex, plain) of the four forms; one block.
struct DkNear {
    const double* x[4];  // complex vectors
    const double* y[4];
    int64_t n[4];
};
__global__ __launch_bounds__(256) void dk_finish_kernel(const double* part, int64_t nparts, DkNear nr, double* out) {
    __shared__ double red[4][8];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) c[q] = 0.0;
    for (int64_t b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] += part[b * 8 + q];
#pragma unroll
    for (int f = 0; f < 4; ++f)
        for (int64_t e = threadIdx.x; e < nr.n[f]; e += blockDim.x)
            cmac(c[2 * f], c[2 * f + 1], nr.x[f][2 * e], nr.x[f][2 * e + 1], nr.y[f][2 * e], nr.y[f][2 * e + 1]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double t = wave_sum(c[q]);
        if (lane == 0) red[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < 8) out[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Zeroes row blockIdx.x of a matrix: row_bytes bytes at out + row * ld_bytes
// (16-byte stores when the rows are 16-byte aligned, else 4-byte ones).
__global__ __launch_bounds__(256) void zero_rows_kernel(char* out, int64_t row_bytes, int64_t ld_bytes, int wide) {
    char* row = out + (int64_t)blockIdx.x * ld_bytes;
    if (wide) {
        for (int64_t w = threadIdx.x; w < row_bytes / 16; w += blockDim.x) ((uint4*)row)[w] = make_uint4(0, 0, 0, 0);
    } else {
        for (int64_t w = threadIdx.x; w < row_bytes / 4; w += blockDim.x) ((uint32_t*)row)[w] = 0u;
    }
}

// ---------------------------------------------------------------------------
// Combined-field operator of the exterior Dirichlet problem,
//   A = 1/2 M + K - i eta V      (P1 x P1; Brakhage-Werner, eta = k usually),
// for a batch of problems (meshes of different sizes, wave numbers) in
// persistent launches: a fixed grid walks one work list spanning every
// problem.  Far items are (problem, TI test panels, 256 trial panels); V and
// K share each point-pair evaluation.  Near items are the near-field pairs
// of every problem (FP64, as in assembly), mass items the panels.  P1 rows and
// columns collect from several panels through atomics.
// ---------------------------------------------------------------------------
struct CfProb {
    OpParams p;          // the problem's mesh tables, kap = k, out = A
    const double* mass_w;
    double eta;
    int64_t far_begin;   // first far item of the problem in the work list
    int64_t near_begin;  // first near item
    int64_t mass_begin;  // first mass item
    int nib, njb;        // far tiling: test-panel groups x trial blocks
};

__device__ __forceinline__ int cf_find(const CfProb* tbl, int count, int64_t item, int which) {
    // last problem whose begin <= item (begins are non-decreasing)
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t b = which == 0 ? tbl[mid].far_begin : (which == 1 ? tbl[mid].near_begin : tbl[mid].mass_begin);
        if (b <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// K - i eta V of a regular P1 x P1 pair in both orders.  The point-pair
// integrand of V is symmetric, and K with test panel j and trial panel i
// differs from K with test panel i only by the normal it projects on
// ((y - x).n_i instead of (x - y).n_j), so one evaluation of r, 1/r and
// exp(i kap r) per point pair feeds all four kernels; only the two
// combinations are accumulated.  C_ij[s][t] pairs test basis s on panel i
// with trial basis t on panel j; C_ji[t][s] is the transposed pair.
template <typename R>
__device__ __forceinline__ void far_pair_cf_sym(const OpParams& p, double eta, int i, int j, double (&Cr)[3][3],
                                                double (&Ci)[3][3], double (&Dr)[3][3], double (&Di)[3][3]) {
    const int q2 = p.q2, nf = p.nf;
    // wave-uniform test panel (the problem table is read per item, so the
    // compiler is told): scalar loads of its points and of the basis table
    const ConstPtr<R> PU = as_const_uniform(Ptr<R>::pu(p) + (int64_t)i * q2 * 4);
    const R* __restrict__ PL = Ptr<R>::pl(p);
    const ConstPtr<R> PHI = as_const_uniform(Ptr<R>::phi(p));
    R dcx = 0, dcy = 0, dcz = 0;
    if constexpr (sizeof(R) == 4) {
        dcx = (R)(p.cen[3 * i] - p.cen[3 * j]);
        dcy = (R)(p.cen[3 * i + 1] - p.cen[3 * j + 1]);
        dcz = (R)(p.cen[3 * i + 2] - p.cen[3 * j + 2]);
    }
    const R nnx = (R)p.normal[3 * j], nny = (R)p.normal[3 * j + 1], nnz = (R)p.normal[3 * j + 2];
    const R mix = (R)p.normal[3 * i], miy = (R)p.normal[3 * i + 1], miz = (R)p.normal[3 * i + 2];
    // (x - y).n_i = x_a.n_i + e_b.n_i, with x_a.n_i the same for every test point of the flat panel i
    const R pun = PU[0] * mix + PU[1] * miy + PU[2] * miz;
    const R kap = (R)p.kap, et = (R)eta;
    R acr[3][3], aci[3][3], bcr[3][3], bci[3][3];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) { acr[s][t] = aci[s][t] = bcr[s][t] = bci[s][t] = 0; }
    for (int b = 0; b < q2; ++b) {
        const R ex = dcx - PL[((int64_t)b * 4 + 0) * nf + j];
        const R ey = dcy - PL[((int64_t)b * 4 + 1) * nf + j];
        const R ez = dcz - PL[((int64_t)b * 4 + 2) * nf + j];
        const R wb = PL[((int64_t)b * 4 + 3) * nf + j];
        const R dni = pun + ex * mix + ey * miy + ez * miz;
        R pcr[3], pci[3], qcr[3], qci[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) { pcr[s] = pci[s] = qcr[s] = qci[s] = 0; }
        for (int a = 0; a < q2; ++a) {
            const R dx = PU[4 * a + 0] + ex;
            const R dy = PU[4 * a + 1] + ey;
            const R dz = PU[4 * a + 2] + ez;
            const R wa = PU[4 * a + 3];
            const R r2 = dx * dx + dy * dy + dz * dz;
            const R rinv = Math<R>::rsqrt(r2);
            const R g = wa * rinv;
            const R g3 = g * (rinv * rinv);
            const R kr = kap * (r2 * rinv);
            R sn, cs;
            Math<R>::sincos(kr, &sn, &cs);
            // K: base (cs + kr sn, sn - kr cs); -i eta V: eta g (sn, -cs)
            const R c1 = cs + kr * sn, c2 = sn - kr * cs;
            const R egs = et * g * sn, egc = et * g * cs;
            const R bj = g3 * (dx * nnx + dy * nny + dz * nnz), bi_ = -g3 * dni;
            const R fr = bj * c1 + egs, fi = bj * c2 - egc;
            const R hr = bi_ * c1 + egs, hi = bi_ * c2 - egc;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const R ph = PHI[3 * a + s];
                pcr[s] += fr * ph; pci[s] += fi * ph;
                qcr[s] += hr * ph; qci[s] += hi * ph;
            }
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const R wt = wb * PHI[3 * b + t];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                acr[s][t] += pcr[s] * wt; aci[s][t] += pci[s] * wt;
                bcr[s][t] += qcr[s] * wt; bci[s][t] += qci[s] * wt;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            Cr[s][t] = kInv4Pi * (double)acr[s][t];
            Ci[s][t] = kInv4Pi * (double)aci[s][t];
            Dr[t][s] = kInv4Pi * (double)bcr[s][t];
            Di[t][s] = kInv4Pi * (double)bci[s][t];
        }
}

// The same pair in FP32 with the test panel in its local frame (FP64 frame
// table, converted once per pair): with Y = y_b - C_i the trial point
// relative to the test centroid and x_a - C_i = U + r1 E1 + r2 E2,
//   |x_a - y_b|^2      = S_a + Y.(Y + D) + p1 (E1.Y) + p2 (E2.Y),
//   (x_a - y_b).n_j    = (U - Y).n_j - p1 (E1.n_j)/2 - p2 (E2.n_j)/2,
//   (y_b - x_a).n_i    = Y.n_i                    (x_a - C_i lies in panel i),
// with (S_a, p1, p2) = (|x_a - C_i|^2, -2 r1, -2 r2) from the per-point table
// and D = -2U: three operations for r^2 and two for the projection per point
// pair instead of six and three.  The test bases weight the sums directly
// (phi from the uniform basis table), so no basis is formed as a difference
// of moments.  NQ2 > 0 fixes the rule size for unrolling.
template <int NQ2>
__device__ __forceinline__ void far_pair_cf_sym_f32(const OpParams& p, double eta, int i, int j, double (&Cr)[3][3],
                                                    double (&Ci)[3][3], double (&Dr)[3][3], double (&Di)[3][3]) {
    const int q2 = NQ2 > 0 ? NQ2 : p.q2, nf = p.nf;
    const ConstPtr<float> PQ = as_const_uniform(p.pq32 + (int64_t)i * q2 * 4);
    const ConstPtr<float> PHI = as_const_uniform(p.phi32);
    const ConstPtr<double> FR = as_const_uniform(p.frame + (int64_t)i * PANEL_FRAME);
    const float* __restrict__ PL = p.pl32;
    float U[3], E1[3], E2[3], D[3], ni[3], nj[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        U[d] = (float)FR[3 + d];
        E1[d] = (float)FR[6 + d];
        E2[d] = (float)FR[9 + d];
        ni[d] = (float)FR[12 + d];
        D[d] = (float)FR[15 + d];
        nj[d] = (float)p.normal[3 * j + d];
    }
    // Y = (y_b - C_j) + (C_j - C_i)
    const float dcx = (float)(p.cen[3 * j] - p.cen[3 * i]);
    const float dcy = (float)(p.cen[3 * j + 1] - p.cen[3 * i + 1]);
    const float dcz = (float)(p.cen[3 * j + 2] - p.cen[3 * i + 2]);
    const float e1n = -0.5f * (E1[0] * nj[0] + E1[1] * nj[1] + E1[2] * nj[2]);
    const float e2n = -0.5f * (E2[0] * nj[0] + E2[1] * nj[1] + E2[2] * nj[2]);
    const float un = U[0] * nj[0] + U[1] * nj[1] + U[2] * nj[2];
    const float kap = (float)p.kap, et = (float)eta;
    float acr[3][3], aci[3][3], bcr[3][3], bci[3][3];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) { acr[s][t] = aci[s][t] = bcr[s][t] = bci[s][t] = 0.f; }
    for (int b = 0; b < q2; ++b) {
        const float yx = PL[((int64_t)b * 4 + 0) * nf + j] + dcx;
        const float yy = PL[((int64_t)b * 4 + 1) * nf + j] + dcy;
        const float yz = PL[((int64_t)b * 4 + 2) * nf + j] + dcz;
        const float wb = PL[((int64_t)b * 4 + 3) * nf + j];
        const float yd = yx * (yx + D[0]) + yy * (yy + D[1]) + yz * (yz + D[2]);
        const float y1 = E1[0] * yx + E1[1] * yy + E1[2] * yz;
        const float y2 = E2[0] * yx + E2[1] * yy + E2[2] * yz;
        const float c0 = un - (yx * nj[0] + yy * nj[1] + yz * nj[2]);
        const float dyi = yx * ni[0] + yy * ni[1] + yz * ni[2];  // (y - x).n_i
        float pcr[3], pci[3], qcr[3], qci[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) { pcr[s] = pci[s] = qcr[s] = qci[s] = 0.f; }
#pragma unroll 5
        for (int a = 0; a < q2; ++a) {
            const float sa = PQ[4 * a + 0], p1 = PQ[4 * a + 1], p2 = PQ[4 * a + 2], wa = PQ[4 * a + 3];
            const float r2 = __builtin_fmaf(p2, y2, __builtin_fmaf(p1, y1, sa + yd));
            const float rinv = rsqrt_f32(r2);
            const float g = wa * rinv;
            const float g3 = g * (rinv * rinv);
            const float kr = kap * (r2 * rinv);
            float sn, cs;
            sincos_f32(kr, &sn, &cs);
            // K: base (cs + kr sn, sn - kr cs); -i eta V: eta g (sn, -cs)
            const float c1 = __builtin_fmaf(kr, sn, cs), c2 = __builtin_fmaf(-kr, cs, sn);
            const float etg = et * g;
            const float egs = etg * sn, egc = etg * cs;
            const float dnj = __builtin_fmaf(p2, e2n, __builtin_fmaf(p1, e1n, c0));
            const float bj = g3 * dnj, bi_ = g3 * dyi;
            const float fr = __builtin_fmaf(bj, c1, egs), fi = __builtin_fmaf(bj, c2, -egc);
            const float hr = __builtin_fmaf(bi_, c1, egs), hi = __builtin_fmaf(bi_, c2, -egc);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const float ph = PHI[3 * a + s];
                pcr[s] = __builtin_fmaf(fr, ph, pcr[s]);
                pci[s] = __builtin_fmaf(fi, ph, pci[s]);
                qcr[s] = __builtin_fmaf(hr, ph, qcr[s]);
                qci[s] = __builtin_fmaf(hi, ph, qci[s]);
            }
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const float wt = wb * PHI[3 * b + t];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                acr[s][t] = __builtin_fmaf(pcr[s], wt, acr[s][t]);
                aci[s][t] = __builtin_fmaf(pci[s], wt, aci[s][t]);
                bcr[s][t] = __builtin_fmaf(qcr[s], wt, bcr[s][t]);
                bci[s][t] = __builtin_fmaf(qci[s], wt, bci[s][t]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            Cr[s][t] = kInv4Pi * (double)acr[s][t];
            Ci[s][t] = kInv4Pi * (double)aci[s][t];
            Dr[t][s] = kInv4Pi * (double)bcr[s][t];
            Di[t][s] = kInv4Pi * (double)bci[s][t];
        }
}

// Far item -> (test group ib, trial block jb) over the upper triangle: the
// test groups of 256-row group gr (256 / TI of them) pair with trial blocks
// gr .. njb-1, and a lane only evaluates pairs with j > i.
template <int TI>
__device__ __forceinline__ void cf_sym_item(int64_t local, int njb, int nib, int& ib, int& jb) {
    constexpr int GP = 256 / TI;
    auto prefix = [&](int g) -> int64_t { return (int64_t)GP * ((int64_t)g * njb - (int64_t)g * (g - 1) / 2); };
    int lo = 0, hi = (nib + GP - 1) / GP - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix(mid) <= local) lo = mid; else hi = mid - 1;
    }
    const int64_t off = local - prefix(lo);
    const int cnt = njb - lo;
    ib = lo * GP + (int)(off / cnt);
    jb = lo + (int)(off % cnt);
}

// FP32 items hold the sums of up to CF_SLOTS test vertices per lane and add
// them to the matrix once a slot is needed for another vertex (and at the
// end of the item): the TI consecutive test panels of an item share most
// vertices (8 panels of a subdivided icosphere touch about 10, not 24), so
// the lane's atomics drop by that factor.  Which slot holds which vertex is
// the same for every lane (the test panels are wave-uniform).
constexpr int CF_SLOTS = 6;

// Three waves per SIMD: the register budget (168 VGPRs) spills about 22
// registers of the FP32 variant, outside the point loops (measured 3 % faster
// than two waves without spills).
template <typename R, int TI, int NQ2 = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void cf_far_kernel(const CfProb* tbl,
                                                                                            int count, int64_t total) {
    for (int64_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int q = cf_find(tbl, count, item, 0);
        const CfProb& P = tbl[q];
        const OpParams& p = P.p;
        int ib, jb;
        cf_sym_item<TI>(item - P.far_begin, P.njb, P.nib, ib, jb);
        const int j = jb * 256 + threadIdx.x;
        const int i_end = min(p.nf, (ib + 1) * TI);
        if constexpr (sizeof(R) == 8) {
            if (j >= p.nf) continue;
            for (int i = ib * TI; i < i_end; ++i) {
                if (j <= i || shares_vertex(p.tris, i, j)) continue;
                double Cr[3][3], Ci[3][3], Dr[3][3], Di[3][3];
                far_pair_cf_sym<R>(p, P.eta, i, j, Cr, Ci, Dr, Di);
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        const int64_t vi = p.tris[3 * i + s], vj = p.tris[3 * j + t];
                        add_elem(p.out, p.dtype, vi * p.ld + vj, Cr[s][t], Ci[s][t], true);
                        add_elem(p.out, p.dtype, vj * p.ld + vi, Dr[t][s], Di[t][s], true);
                    }
            }
        } else {
            const bool jv = j < p.nf;
            const int jj = jv ? j : 0;
            int64_t vt[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) vt[t] = p.tris[3 * jj + t];
            int sv[CF_SLOTS];  // test vertex of each slot (-1: free), the same in every lane
            float ar[CF_SLOTS][3], ai[CF_SLOTS][3], br[CF_SLOTS][3], bi[CF_SLOTS][3];
#pragma unroll
            for (int k = 0; k < CF_SLOTS; ++k) {
                sv[k] = -1;
#pragma unroll
                for (int t = 0; t < 3; ++t) ar[k][t] = ai[k][t] = br[k][t] = bi[k][t] = 0.f;
            }
            // adds slot k's sums to A[u][v_t] and A[v_t][u] and frees it
            auto flush = [&](int k) {
#pragma unroll
                for (int kk = 0; kk < CF_SLOTS; ++kk) {
                    if (kk != k) continue;
                    const int64_t u = sv[kk];
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        if (ar[kk][t] != 0.f || ai[kk][t] != 0.f)
                            add_elem(p.out, p.dtype, u * p.ld + vt[t], ar[kk][t], ai[kk][t], true);
                        if (br[kk][t] != 0.f || bi[kk][t] != 0.f)
                            add_elem(p.out, p.dtype, vt[t] * p.ld + u, br[kk][t], bi[kk][t], true);
                        ar[kk][t] = ai[kk][t] = br[kk][t] = bi[kk][t] = 0.f;
                    }
                    sv[kk] = -1;
                }
            };
            for (int i = ib * TI; i < i_end; ++i) {
                const int a[3] = {p.tris[3 * i], p.tris[3 * i + 1], p.tris[3 * i + 2]};
                int slot[3];
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    slot[s] = -1;
#pragma unroll
                    for (int k = 0; k < CF_SLOTS; ++k)
                        if (sv[k] == a[s]) slot[s] = k;
                }
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    if (slot[s] >= 0) continue;
                    int k = -1;
#pragma unroll
                    for (int kk = CF_SLOTS - 1; kk >= 0; --kk)
                        if (sv[kk] < 0) k = kk;
                    if (k < 0) {
                        // evict the first slot this panel does not use
#pragma unroll
                        for (int kk = CF_SLOTS - 1; kk >= 0; --kk)
                            if (kk != slot[0] && kk != slot[1] && kk != slot[2]) k = kk;
                        flush(k);
                    }
                    sv[k] = a[s];
                    slot[s] = k;
                }
                if (!jv || j <= i || shares_vertex(p.tris, i, j)) continue;
                double Cr[3][3], Ci[3][3], Dr[3][3], Di[3][3];
                far_pair_cf_sym_f32<NQ2>(p, P.eta, i, j, Cr, Ci, Dr, Di);
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int k = 0; k < CF_SLOTS; ++k) {
                        if (k != slot[s]) continue;
#pragma unroll
                        for (int t = 0; t < 3; ++t) {
                            ar[k][t] += (float)Cr[s][t];
                            ai[k][t] += (float)Ci[s][t];
                            br[k][t] += (float)Dr[t][s];
                            bi[k][t] += (float)Di[t][s];
                        }
                    }
            }
#pragma unroll
            for (int k = 0; k < CF_SLOTS; ++k)
                if (sv[k] >= 0) flush(k);
        }
    }
}

// near items: one wave per near pair (V and K P1 x P1, FP64)
__global__ __launch_bounds__(256) void cf_near_kernel(const CfProb* tbl, int count, int64_t total) {
    const int lane = threadIdx.x & 63;
    for (int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); item < total; item += (int64_t)gridDim.x * 4) {
        const int q = cf_find(tbl, count, item, 1);
        const CfProb& P = tbl[q];
        const NearPair np = P.p.near[item - P.near_begin];
        double Cr[3][3], Ci[3][3];
        near_pair<kOpCf, true, 1, 1>(P.p, np, lane, Cr, Ci, P.eta);
        if (lane == 0)
            for (int s = 0; s < 3; ++s)
                for (int t = 0; t < 3; ++t)
                    add_elem(P.p.out, P.p.dtype, (int64_t)P.p.tris[3 * np.i + s] * P.p.ld + P.p.tris[3 * np.j + t],
                             Cr[s][t], Ci[s][t], true);
    }
}

// 1/2 M (P1 x P1), one thread per panel item
__global__ void cf_mass_kernel(const CfProb* tbl, int count, int64_t total) {
    for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < total;
         item += (int64_t)gridDim.x * blockDim.x) {
        const int q = cf_find(tbl, count, item, 2);
        const CfProb& P = tbl[q];
        const int i = (int)(item - P.mass_begin);
        double m[3][3];
        mass_block(P.p, P.mass_w, i, 1, 1, m);
        for (int s = 0; s < 3; ++s)
            for (int t = 0; t < 3; ++t)
                add_elem(P.p.out, P.p.dtype, (int64_t)P.p.tris[3 * i + s] * P.p.ld + P.p.tris[3 * i + t],
                         0.5 * m[s][t], 0.0, true);
    }
}

// ------------------------- dispatch -----------------------------------------
using FarFn = void (*)(OpParams, int);
using NearFn = void (*)(OpParams);

template <typename R, int OP, bool HELM>
FarFn far_assemble_fn(int dt, int dr) {
    if (dt == 0 && dr == 0) return far_assemble_kernel<R, OP, HELM, 0, 0>;
    if (dt == 0 && dr == 1) return far_assemble_kernel<R, OP, HELM, 0, 1>;
    if (dt == 1 && dr == 0) return far_assemble_kernel<R, OP, HELM, 1, 0>;
    return far_assemble_kernel<R, OP, HELM, 1, 1>;
}
template <typename R, int OP, bool HELM>
NearFn far_apply_fn(int dt, int dr) {
    if (dt == 0 && dr == 0) return far_apply_kernel<R, OP, HELM, 0, 0>;
    if (dt == 0 && dr == 1) return far_apply_kernel<R, OP, HELM, 0, 1>;
    if (dt == 1 && dr == 0) return far_apply_kernel<R, OP, HELM, 1, 0>;
    return far_apply_kernel<R, OP, HELM, 1, 1>;
}
template <int OP, bool HELM>
NearFn near_fn(int dt, int dr, bool apply) {
    if (apply) {
        if (dt == 0 && dr == 0) return near_apply_kernel<OP, HELM, 0, 0>;
        if (dt == 0 && dr == 1) return near_apply_kernel<OP, HELM, 0, 1>;
        if (dt == 1 && dr == 0) return near_apply_kernel<OP, HELM, 1, 0>;
        return near_apply_kernel<OP, HELM, 1, 1>;
    }
    if (dt == 0 && dr == 0) return near_assemble_kernel<OP, HELM, 0, 0>;
    if (dt == 0 && dr == 1) return near_assemble_kernel<OP, HELM, 0, 1>;
    if (dt == 1 && dr == 0) return near_assemble_kernel<OP, HELM, 1, 0>;
    return near_assemble_kernel<OP, HELM, 1, 1>;
}

template <typename R, int OP>
FarFn far_assemble_k(bool helm, int dt, int dr) {
    return helm ? far_assemble_fn<R, OP, true>(dt, dr) : far_assemble_fn<R, OP, false>(dt, dr);
}
template <typename R>
FarFn far_assemble_any(int op, bool helm, int dt, int dr) {
    switch (op) {
        case BEM_OP_V: return far_assemble_k<R, BEM_OP_V>(helm, dt, dr);
        case BEM_OP_K: return far_assemble_k<R, BEM_OP_K>(helm, dt, dr);
        case BEM_OP_KADJ: return far_assemble_k<R, BEM_OP_KADJ>(helm, dt, dr);
        default: return far_assemble_k<R, BEM_OP_W>(helm, dt, dr);
    }
}
template <typename R, int OP>
NearFn far_apply_k(bool helm, int dt, int dr) {
    return helm ? far_apply_fn<R, OP, true>(dt, dr) : far_apply_fn<R, OP, false>(dt, dr);
}
template <typename R>
NearFn far_apply_any(int op, bool helm, int dt, int dr) {
    switch (op) {
        case BEM_OP_V: return far_apply_k<R, BEM_OP_V>(helm, dt, dr);
        case BEM_OP_K: return far_apply_k<R, BEM_OP_K>(helm, dt, dr);
        case BEM_OP_KADJ: return far_apply_k<R, BEM_OP_KADJ>(helm, dt, dr);
        default: return far_apply_k<R, BEM_OP_W>(helm, dt, dr);
    }
}
template <int OP>
NearFn near_k(bool helm, int dt, int dr, bool apply) {
    return helm ? near_fn<OP, true>(dt, dr, apply) : near_fn<OP, false>(dt, dr, apply);
}
NearFn near_any(int op, bool helm, int dt, int dr, bool apply) {
    switch (op) {
        case BEM_OP_V: return near_k<BEM_OP_V>(helm, dt, dr, apply);
        case BEM_OP_K: return near_k<BEM_OP_K>(helm, dt, dr, apply);
        case BEM_OP_KADJ: return near_k<BEM_OP_KADJ>(helm, dt, dr, apply);
        default: return near_k<BEM_OP_W>(helm, dt, dr, apply);
    }
}

template <int OP>
NearFn far_apply_dk_fn(int dt, int dr) {
    if (dt == 0 && dr == 0) return far_apply_kernel<double, OP, true, 0, 0, true>;
    if (dt == 0 && dr == 1) return far_apply_kernel<double, OP, true, 0, 1, true>;
    if (dt == 1 && dr == 0) return far_apply_kernel<double, OP, true, 1, 0, true>;
    return far_apply_kernel<double, OP, true, 1, 1, true>;
}
template <int OP>
NearFn near_apply_dk_fn(int dt, int dr) {
    if (dt == 0 && dr == 0) return near_apply_kernel<OP, true, 0, 0, true>;
    if (dt == 0 && dr == 1) return near_apply_kernel<OP, true, 0, 1, true>;
    if (dt == 1 && dr == 0) return near_apply_kernel<OP, true, 1, 0, true>;
    return near_apply_kernel<OP, true, 1, 1, true>;
}
NearFn far_apply_dk_any(int op, int dt, int dr) {
    switch (op) {
        case BEM_OP_V: return far_apply_dk_fn<BEM_OP_V>(dt, dr);
        case BEM_OP_K: return far_apply_dk_fn<BEM_OP_K>(dt, dr);
        case BEM_OP_KADJ: return far_apply_dk_fn<BEM_OP_KADJ>(dt, dr);
        default: return far_apply_dk_fn<BEM_OP_W>(dt, dr);
    }
}
NearFn near_apply_dk_any(int op, int dt, int dr) {
    switch (op) {
        case BEM_OP_V: return near_apply_dk_fn<BEM_OP_V>(dt, dr);
        case BEM_OP_K: return near_apply_dk_fn<BEM_OP_K>(dt, dr);
        case BEM_OP_KADJ: return near_apply_dk_fn<BEM_OP_KADJ>(dt, dr);
        default: return near_apply_dk_fn<BEM_OP_W>(dt, dr);
    }
}

template <int OP, bool HELM>
void apply_t_fns(int dt, int dr, FarFn& far, NearFn& nearf) {
    if (dt == 0 && dr == 0) { far = far_apply_t_kernel<OP, HELM, 0, 0>; nearf = near_apply_t_kernel<OP, HELM, 0, 0>; }
    else if (dt == 0 && dr == 1) { far = far_apply_t_kernel<OP, HELM, 0, 1>; nearf = near_apply_t_kernel<OP, HELM, 0, 1>; }
    else if (dt == 1 && dr == 0) { far = far_apply_t_kernel<OP, HELM, 1, 0>; nearf = near_apply_t_kernel<OP, HELM, 1, 0>; }
    else { far = far_apply_t_kernel<OP, HELM, 1, 1>; nearf = near_apply_t_kernel<OP, HELM, 1, 1>; }
}
template <int OP>
void apply_t_k(bool helm, int dt, int dr, FarFn& far, NearFn& nearf) {
    if (helm) apply_t_fns<OP, true>(dt, dr, far, nearf);
    else apply_t_fns<OP, false>(dt, dr, far, nearf);
}
void apply_t_any(int op, bool helm, int dt, int dr, FarFn& far, NearFn& nearf) {
    switch (op) {
        case BEM_OP_V: apply_t_k<BEM_OP_V>(helm, dt, dr, far, nearf); break;
        case BEM_OP_K: apply_t_k<BEM_OP_K>(helm, dt, dr, far, nearf); break;
        case BEM_OP_KADJ: apply_t_k<BEM_OP_KADJ>(helm, dt, dr, far, nearf); break;
        default: apply_t_k<BEM_OP_W>(helm, dt, dr, far, nearf); break;
    }
}

int validate(Context* ctx, const bem_op_desc* d) {
    if (!ctx->mesh.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
    if (!d) return set_error(BEM_ERR_INVALID, "null operator descriptor");
    if (d->op < BEM_OP_V || d->op > BEM_OP_W) return set_error(BEM_ERR_INVALID, "unknown operator");
    if (d->kernel < BEM_LAPLACE || d->kernel > BEM_HELMHOLTZ_TRANSMISSION)
        return set_error(BEM_ERR_INVALID, "Kernel Type only support Laplace and Helmholtz for now");
    if ((d->deg_test | d->deg_trial) & ~1) return set_error(BEM_ERR_INVALID, "basis degree must be 0 or 1");
    if (d->dtype < BEM_F64 || d->dtype > BEM_C64) return set_error(BEM_ERR_INVALID, "unknown dtype");
    if (d->parts & ~(BEM_PART_FAR | BEM_PART_NEAR)) return set_error(BEM_ERR_INVALID, "unknown parts mask");
    if (d->kernel != BEM_LAPLACE && !dtype_complex(d->dtype))
        return set_error(BEM_ERR_INVALID, "Helmholtz operators need a complex dtype");
    if (d->precision != BEM_PREC_FP64 && d->precision != BEM_PREC_FP32)
        return set_error(BEM_ERR_INVALID, "unknown precision");
    if (d->flags & ~(BEM_FLAG_ROTATE_NEAR | BEM_FLAG_SYM_COINCIDE | BEM_FLAG_DK | BEM_FLAG_TRANSPOSE))
        return set_error(BEM_ERR_INVALID, "unknown flags");
    return BEM_OK;
}

}  // namespace

int assemble_op(Context* ctx, const bem_op_desc* d, void* out, int64_t ld, double ar, double ai,
                hipStream_t s) {
    int rc = validate(ctx, d);
    if (rc) return rc;
    if (!out) return set_error(BEM_ERR_INVALID, "null output");
    if (d->flags & BEM_FLAG_DK)
        return set_error(BEM_ERR_UNSUPPORTED, "k-derivative operators are matrix-free only (bem_apply)");
    if (d->flags & BEM_FLAG_TRANSPOSE)
        return set_error(BEM_ERR_INVALID, "BEM_FLAG_TRANSPOSE is a bem_apply flag");
    rc = mesh_build_near(ctx, d->flags, s);
    if (rc) return rc;
    OpParams p = make_params(ctx, d);
    p.out = out; p.ld = ld; p.dtype = d->dtype; p.ar = ar; p.ai = ai;
    const bool helm = d->kernel != BEM_LAPLACE;
    const int nf = p.nf, njb = (nf + 255) / 256;
    const int parts = d->parts ? d->parts : (BEM_PART_FAR | BEM_PART_NEAR);
    if (parts & BEM_PART_FAR) {
        FarFn far = d->precision == BEM_PREC_FP32 ? far_assemble_any<float>(d->op, helm, d->deg_test, d->deg_trial)
                                                  : far_assemble_any<double>(d->op, helm, d->deg_test, d->deg_trial);
        hipLaunchKernelGGL(far, dim3((unsigned)((int64_t)nf * njb)), dim3(256), 0, s, p, njb);
        rc = check_hip(hipGetLastError(), "far-field assembly launch");
        if (rc) return rc;
    }
    if (p.n_near > 0 && (parts & BEM_PART_NEAR)) {
        NearFn nearf = near_any(d->op, helm, d->deg_test, d->deg_trial, false);
        hipLaunchKernelGGL(nearf, dim3((unsigned)((p.n_near + 3) / 4)), dim3(256), 0, s, p);
        rc = check_hip(hipGetLastError(), "near-field assembly launch");
    }
    return rc;
}

int apply_op(Context* ctx, const bem_op_desc* d, const void* x, int conj_x, void* y, double ar,
             double ai, hipStream_t s) {
    int rc = validate(ctx, d);
    if (rc) return rc;
    if (!x || !y) return set_error(BEM_ERR_INVALID, "null vector");
    const bool dk = (d->flags & BEM_FLAG_DK) != 0;
    const bool helm = d->kernel != BEM_LAPLACE;
    if (dk && d->precision != BEM_PREC_FP64)
        return set_error(BEM_ERR_UNSUPPORTED, "k-derivative products run in FP64 only");
    const bool tr = (d->flags & BEM_FLAG_TRANSPOSE) != 0;
    if (tr && dk) return set_error(BEM_ERR_UNSUPPORTED, "transposed k-derivative products are not provided");
    if (tr && d->precision != BEM_PREC_FP64)
        return set_error(BEM_ERR_UNSUPPORTED, "transposed products run in FP64 only");
    if (dk && !helm) return BEM_OK;  // the Laplace kernel does not depend on k
    rc = mesh_build_near(ctx, d->flags, s);
    if (rc) return rc;
    OpParams p = make_params(ctx, d);
    p.x = x; p.conj_x = conj_x; p.dtype_x = d->dtype; p.y = y; p.ar = ar; p.ai = ai;
    const int parts = d->parts ? d->parts : (BEM_PART_FAR | BEM_PART_NEAR);
    if (tr) {
        FarFn far = nullptr;
        NearFn nearf = nullptr;
        apply_t_any(d->op, helm, d->deg_test, d->deg_trial, far, nearf);
        const int njb = (p.nf + 255) / 256, nib = (p.nf + TCH - 1) / TCH;
        if (parts & BEM_PART_FAR) {
            hipLaunchKernelGGL(far, dim3((unsigned)((int64_t)nib * njb)), dim3(256), 0, s, p, njb);
            rc = check_hip(hipGetLastError(), "transposed far-field apply launch");
            if (rc) return rc;
        }
        if (p.n_near > 0 && (parts & BEM_PART_NEAR)) {
            hipLaunchKernelGGL(nearf, dim3((unsigned)((p.n_near + 3) / 4)), dim3(256), 0, s, p);
            const int ncol = d->deg_trial ? p.nv : p.nf;
            hipLaunchKernelGGL(trial_gather_kernel, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, s, p,
                               d->deg_trial);
            rc = check_hip(hipGetLastError(), "transposed near-field apply launch");
        }
        return rc;
    }
    const bool use_far = (parts & BEM_PART_FAR) != 0;
    const bool use_near = p.n_near > 0 && (parts & BEM_PART_NEAR);
    if (use_far) {
        NearFn far = dk ? far_apply_dk_any(d->op, d->deg_test, d->deg_trial)
                        : d->precision == BEM_PREC_FP32 ? far_apply_any<float>(d->op, helm, d->deg_test, d->deg_trial)
                                                        : far_apply_any<double>(d->op, helm, d->deg_test, d->deg_trial);
        hipLaunchKernelGGL(far, dim3(p.nf), dim3(256), 0, s, p);
        rc = check_hip(hipGetLastError(), "far-field apply launch");
        if (rc) return rc;
    }
    if (use_near) {
        NearFn nearf = dk ? near_apply_dk_any(d->op, d->deg_test, d->deg_trial)
                          : near_any(d->op, helm, d->deg_test, d->deg_trial, true);
        hipLaunchKernelGGL(nearf, dim3((unsigned)((p.n_near + 3) / 4)), dim3(256), 0, s, p);
        if (!d->deg_test)
            hipLaunchKernelGGL(near_reduce_kernel, dim3((unsigned)((p.nf + 255) / 256)), dim3(256), 0, s, p);
        rc = check_hip(hipGetLastError(), "near-field apply launch");
        if (rc) return rc;
    }
    if (d->deg_test && (use_far || use_near)) {
        hipLaunchKernelGGL(vertex_rows_kernel, dim3((unsigned)((p.nv + 255) / 256)), dim3(256), 0, s, p,
                           (int)use_far, (int)use_near);
        rc = check_hip(hipGetLastError(), "vertex-row sum launch");
    }
    return rc;
}

int assemble_dirichlet(Context* ctx, const bem_op_desc* d, void* A, int64_t ld, double ar, double ai,
                       const void* g, int conj_g, void* rhs, int dtype_rhs, double alpha_k, double alpha_m,
                       hipStream_t s) {
    int rc = validate(ctx, d);
    if (rc) return rc;
    if (!A || !g || !rhs) return set_error(BEM_ERR_INVALID, "null argument");
    if (d->op != BEM_OP_V || d->deg_test != 0 || d->deg_trial != 0)
        return set_error(BEM_ERR_INVALID, "the Dirichlet system takes the P0 single-layer descriptor");
    if (d->flags & BEM_FLAG_DK) return set_error(BEM_ERR_UNSUPPORTED, "k-derivative: use bem_apply");
    if (d->flags & BEM_FLAG_TRANSPOSE) return set_error(BEM_ERR_INVALID, "BEM_FLAG_TRANSPOSE is a bem_apply flag");
    if (dtype_rhs < BEM_F64 || dtype_rhs > BEM_C64) return set_error(BEM_ERR_INVALID, "unknown rhs dtype");
    if (dtype_complex(d->dtype) != dtype_complex(dtype_rhs))
        return set_error(BEM_ERR_INVALID, "matrix and right-hand side must both be real or both complex");
    rc = mesh_build_near(ctx, d->flags, s);
    if (rc) return rc;
    Mesh& m = ctx->mesh;
    const int nf = m.nf, njb = (nf + 255) / 256;
    // row-i partials [nf][njb], then the row-j partials: [nf] (FP32, atomics)
    // or [row slots][nf] (FP64, one store per slot)
    const size_t jrows = d->precision == BEM_PREC_FP32 ? 1 : (size_t)njb * DSUB;
    const size_t kbytes = ((size_t)nf * njb + jrows * nf) * 2 * sizeof(double);
    if (m.kpart_bytes < kbytes) {
        if (m.kpart) hipFree(m.kpart);
        m.kpart = nullptr;
        m.kpart_bytes = 0;
        rc = check_hip(hipMalloc((void**)&m.kpart, kbytes), "K partials");
        if (rc) return rc;
        m.kpart_bytes = kbytes;
    }
    OpParams p = make_params(ctx, d);
    p.out = A; p.ld = ld; p.dtype = d->dtype; p.ar = ar; p.ai = ai;
    p.x = g; p.conj_x = conj_g; p.dtype_x = dtype_rhs; p.y = rhs;
    const bool helm = d->kernel != BEM_LAPLACE;
    const int parts = d->parts ? d->parts : (BEM_PART_FAR | BEM_PART_NEAR);
    if (parts & BEM_PART_FAR) {
        const int ngroups = (nf + 255) / 256;
        if (d->precision == BEM_PREC_FP32) {
            auto fk = helm ? far_dirichlet_kernel<float, true> : far_dirichlet_kernel<float, false>;
            double* krow = m.kpart + (size_t)nf * njb * 2;
            hipLaunchKernelGGL(zero_kernel, dim3((2 * nf + 255) / 256), dim3(256), 0, s, krow, (int64_t)2 * nf);
            // upper-triangle blocks: DTG tiles of group g pair with njb - g trial blocks
            const int64_t nblocks = (int64_t)DTG * ((int64_t)ngroups * njb - (int64_t)ngroups * (ngroups - 1) / 2);
            hipLaunchKernelGGL(fk, dim3((unsigned)nblocks), dim3(256), 0, s, p, njb, ngroups, m.kpart, krow);
            hipLaunchKernelGGL(kpart_reduce_kernel, dim3((nf + 255) / 256), dim3(256), 0, s, (const double*)m.kpart,
                               (const double*)krow, nf, njb, rhs, dtype_rhs, alpha_k, 0.0);
        } else {
            // FP64: moment form, fixed-order partials for both rows
            // Helmholtz at kap = 0 is the Laplace kernel (the Helmholtz instance scales by 1/|kap|)
            const bool hk = helm && p.kap != 0.0;
            auto fk = m.q2 == 25 ? (hk ? far_dirichlet_f64_kernel<true, 25> : far_dirichlet_f64_kernel<false, 25>)
                                 : (hk ? far_dirichlet_f64_kernel<true, 0> : far_dirichlet_f64_kernel<false, 0>);
            double* jpart = m.kpart + (size_t)nf * njb * 2;
            const int64_t nblocks = (int64_t)DSUB * ((int64_t)ngroups * njb - (int64_t)ngroups * (ngroups - 1) / 2);
            hipLaunchKernelGGL(fk, dim3((unsigned)nblocks), dim3(256), 0, s, p, njb, ngroups, m.kpart, jpart);
            hipLaunchKernelGGL(kpart_jpart_reduce_kernel, dim3((nf + 255) / 256), dim3(256), 0, s,
                               (const double*)m.kpart, (const double*)jpart, nf, njb, rhs, dtype_rhs, alpha_k, 0.0);
        }
        rc = check_hip(hipGetLastError(), "far-field Dirichlet system launch");
        if (rc) return rc;
    }
    if (parts & BEM_PART_NEAR) {
        // near-field V into the matrix, near-field K conj(g) into the rhs,
        // and alpha_m M conj(g)
        bem_op_desc dk = *d;
        dk.op = BEM_OP_K; dk.deg_trial = 1; dk.dtype = dtype_rhs; dk.parts = BEM_PART_NEAR;
        if (p.n_near > 0) {
            NearFn nv = near_any(BEM_OP_V, helm, 0, 0, false);
            hipLaunchKernelGGL(nv, dim3((unsigned)((p.n_near + 3) / 4)), dim3(256), 0, s, p);
        }
        rc = check_hip(hipGetLastError(), "near-field V launch");
        if (rc) return rc;
        rc = apply_op(ctx, &dk, g, conj_g, rhs, alpha_k, 0.0, s);
        if (rc) return rc;
        if (alpha_m != 0.0) rc = mass_apply(ctx, 0, 1, dtype_rhs, g, conj_g, rhs, alpha_m, s);
    }
    return rc;
}

namespace {
int zero_matrix(void* out, int64_t rows, int64_t cols, int64_t ld, int dtype, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return BEM_OK;
    const int64_t es = (int64_t)dtype_size(dtype);
    const int64_t row_bytes = cols * es, ld_bytes = ld * es;
    const int wide = ((uintptr_t)out % 16 == 0 && row_bytes % 16 == 0 && ld_bytes % 16 == 0) ? 1 : 0;
    hipLaunchKernelGGL(zero_rows_kernel, dim3((unsigned)rows), dim3(256), 0, s, (char*)out, row_bytes, ld_bytes, wide);
    return check_hip(hipGetLastError(), "matrix zero launch");
}

template <typename E>
int mirror_upper(void* W, int64_t ld, int n, hipStream_t s) {
    const int ntiles = (n + 31) / 32;
    const int64_t nblocks = (int64_t)ntiles * (ntiles + 1) / 2;
    hipLaunchKernelGGL(mirror_upper_kernel<E>, dim3((unsigned)nblocks), dim3(256), 0, s, (E*)W, ld, n, ntiles);
    return check_hip(hipGetLastError(), "W mirror launch");
}
}  // namespace

int assemble_vkw(Context* ctx, const bem_op_desc* d, void* V, int64_t ldv, void* K, int64_t ldk, void* W,
                 int64_t ldw, hipStream_t s) {
    int rc = validate(ctx, d);
    if (rc) return rc;
    if (!V || !K || !W) return set_error(BEM_ERR_INVALID, "null output");
    if (d->op != BEM_OP_V || d->deg_test != 0 || d->deg_trial != 0)
        return set_error(BEM_ERR_INVALID, "the V/K/W pass takes the P0 single-layer descriptor");
    if (d->flags & BEM_FLAG_DK)
        return set_error(BEM_ERR_UNSUPPORTED, "k-derivative operators are matrix-free only (bem_apply)");
    if (d->flags & BEM_FLAG_TRANSPOSE) return set_error(BEM_ERR_INVALID, "BEM_FLAG_TRANSPOSE is a bem_apply flag");
    Mesh& m = ctx->mesh;
    if (ldv < m.nf || ldk < m.nv || ldw < m.nv)
        return set_error(BEM_ERR_INVALID, "leading dimension smaller than the block width");
    rc = mesh_build_near(ctx, d->flags, s);
    if (rc) return rc;
    const int parts = d->parts ? d->parts : (BEM_PART_FAR | BEM_PART_NEAR);
    rc = zero_matrix(V, m.nf, m.nf, ldv, d->dtype, s);
    if (!rc) rc = zero_matrix(K, m.nf, m.nv, ldk, d->dtype, s);
    if (!rc) rc = zero_matrix(W, m.nv, m.nv, ldw, d->dtype, s);
    if (rc) return rc;
    OpParams p = make_params(ctx, d);
    p.dtype = d->dtype; p.ar = 1.0; p.ai = 0.0;
    const bool helm = d->kernel != BEM_LAPLACE;
    const int nf = m.nf, njb = (nf + 255) / 256, ngroups = (nf + 255) / 256;
    if (parts & BEM_PART_FAR) {
        const bool q25 = p.q2 == 25;
        // kap = 0 takes the Laplace instance (the FP64 Helmholtz one scales by 1/|kap|)
        const bool hk = helm && p.kap != 0.0;
        auto fk = d->precision == BEM_PREC_FP32
                      ? (helm ? (q25 ? far_vkw_kernel<float, true, 25> : far_vkw_kernel<float, true>)
                              : (q25 ? far_vkw_kernel<float, false, 25> : far_vkw_kernel<float, false>))
                      : (hk ? (q25 ? far_vkw_kernel<double, true, 25> : far_vkw_kernel<double, true>)
                            : (q25 ? far_vkw_kernel<double, false, 25> : far_vkw_kernel<double, false>));
        const int64_t nblocks = (int64_t)DTG * ((int64_t)ngroups * njb - (int64_t)ngroups * (ngroups - 1) / 2);
        VkwOut o{V, ldv, K, ldk, W, ldw};
        hipLaunchKernelGGL(fk, dim3((unsigned)nblocks), dim3(256), 0, s, p, o, njb, ngroups);
        rc = check_hip(hipGetLastError(), "far-field V/K/W launch");
        if (rc) return rc;
        switch (d->dtype) {
            case BEM_C128: rc = mirror_upper<double2>(W, ldw, m.nv, s); break;
            case BEM_C64:
            case BEM_F64: rc = mirror_upper<double>(W, ldw, m.nv, s); break;  // 8-byte elements
            default: rc = mirror_upper<float>(W, ldw, m.nv, s); break;
        }
        if (rc) return rc;
    }
    if (p.n_near > 0 && (parts & BEM_PART_NEAR)) {
        const unsigned nb = (unsigned)((p.n_near + 3) / 4);
        p.out = V; p.ld = ldv;
        hipLaunchKernelGGL(near_any(BEM_OP_V, helm, 0, 0, false), dim3(nb), dim3(256), 0, s, p);
        p.out = K; p.ld = ldk;
        hipLaunchKernelGGL(near_any(BEM_OP_K, helm, 0, 1, false), dim3(nb), dim3(256), 0, s, p);
        p.out = W; p.ld = ldw;
        hipLaunchKernelGGL(near_any(BEM_OP_W, helm, 1, 1, false), dim3(nb), dim3(256), 0, s, p);
        rc = check_hip(hipGetLastError(), "near-field V/K/W launch");
    }
    return rc;
}

int vkw_dk_forms(Context* ctx, const bem_op_desc* d, const void* aN, const void* bN, const void* aD,
                 const void* bD, double* out, hipStream_t s) {
    int rc = validate(ctx, d);
    if (rc) return rc;
    if (!aN || !bN || !aD || !bD || !out) return set_error(BEM_ERR_INVALID, "null argument");
    if (d->op != BEM_OP_V || d->deg_test != 0 || d->deg_trial != 0)
        return set_error(BEM_ERR_INVALID, "the k-derivative forms take the P0 single-layer descriptor");
    if (d->kernel == BEM_LAPLACE) return set_error(BEM_ERR_INVALID, "the Laplace kernel does not depend on k");
    if (d->precision != BEM_PREC_FP64 || d->dtype != BEM_C128)
        return set_error(BEM_ERR_UNSUPPORTED, "k-derivative forms run in FP64 on complex128 vectors");
    if (d->flags & BEM_FLAG_TRANSPOSE) return set_error(BEM_ERR_INVALID, "BEM_FLAG_TRANSPOSE is a bem_apply flag");
    rc = mesh_build_near(ctx, d->flags & ~BEM_FLAG_DK, s);
    if (rc) return rc;
    Mesh& m = ctx->mesh;
    const int nf = m.nf, nv = m.nv, njb = (nf + 255) / 256, ngroups = (nf + 255) / 256;
    const int64_t nblocks = (int64_t)DTG * ((int64_t)ngroups * njb - (int64_t)ngroups * (ngroups - 1) / 2);
    const size_t ylen = 2 * ((size_t)3 * nf + nv);  // doubles of the near-field vectors
    const size_t bytes = ((size_t)nblocks * 8 + ylen) * sizeof(double);
    if (m.kpart_bytes < bytes) {
        if (m.kpart) hipFree(m.kpart);
        m.kpart = nullptr;
        m.kpart_bytes = 0;
        rc = check_hip(hipMalloc((void**)&m.kpart, bytes), "k-derivative partials");
        if (rc) return rc;
        m.kpart_bytes = bytes;
    }
    double* part = m.kpart;
    double* y = part + (size_t)nblocks * 8;
    double* yv[4] = {y, y + 2 * (size_t)nf, y + 4 * (size_t)nf, y + 6 * (size_t)nf};
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((ylen + 255) / 256)), dim3(256), 0, s, y, (int64_t)ylen);
    OpParams p = make_params(ctx, d);
    DkVecs vv{(const double*)aN, (const double*)bN, (const double*)aD, (const double*)bD};
    if (p.q2 == 25)
        hipLaunchKernelGGL(far_vkw_dk_kernel<25>, dim3((unsigned)nblocks), dim3(256), 0, s, p, vv, njb, ngroups, part);
    else
        hipLaunchKernelGGL(far_vkw_dk_kernel<0>, dim3((unsigned)nblocks), dim3(256), 0, s, p, vv, njb, ngroups, part);
    rc = check_hip(hipGetLastError(), "far-field k-derivative forms launch");
    if (rc) return rc;
    // near field: the matrix-free k-derivative products, then plain dots
    const int ops[4][3] = {{BEM_OP_V, 0, 0}, {BEM_OP_K, 0, 1}, {BEM_OP_K, 0, 1}, {BEM_OP_W, 1, 1}};
    const void* xin[4] = {bN, bD, aD, bD};
    for (int f = 0; f < 4; ++f) {
        bem_op_desc dn = *d;
        dn.op = ops[f][0]; dn.deg_test = ops[f][1]; dn.deg_trial = ops[f][2];
        dn.flags = d->flags | BEM_FLAG_DK; dn.parts = BEM_PART_NEAR;
        rc = apply_op(ctx, &dn, xin[f], 0, yv[f], 1.0, 0.0, s);
        if (rc) return rc;
    }
    DkNear nr{{(const double*)aN, (const double*)aN, (const double*)bN, (const double*)aD},
              {yv[0], yv[1], yv[2], yv[3]},
              {nf, nf, nf, nv}};
    hipLaunchKernelGGL(dk_finish_kernel, dim3(1), dim3(256), 0, s, (const double*)part, nblocks, nr, out);
    return check_hip(hipGetLastError(), "k-derivative forms reduction launch");
}

int mass_assemble(Context* ctx, int deg_test, int deg_trial, int dtype, void* out, int64_t ld,
                  double alpha, int transpose, hipStream_t s) {
    if (!ctx->mesh.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
    if (!out) return set_error(BEM_ERR_INVALID, "null output");
    if ((deg_test | deg_trial) & ~1) return set_error(BEM_ERR_INVALID, "basis degree must be 0 or 1");
    if (dtype < BEM_F64 || dtype > BEM_C64) return set_error(BEM_ERR_INVALID, "unknown dtype");
    OpParams p = make_params(ctx, nullptr);
    p.out = out; p.ld = ld; p.dtype = dtype; p.ar = alpha;
    hipLaunchKernelGGL(mass_assemble_kernel, dim3((p.nf + 255) / 256), dim3(256), 0, s, p,
                       ctx->mesh.mass_w, deg_test, deg_trial, transpose);
    return check_hip(hipGetLastError(), "mass assembly launch");
}

int mass_apply(Context* ctx, int deg_test, int deg_trial, int dtype, const void* x, int conj_x,
               void* y, double alpha, hipStream_t s) {
    if (!ctx->mesh.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
    if (!x || !y) return set_error(BEM_ERR_INVALID, "null vector");
    if ((deg_test | deg_trial) & ~1) return set_error(BEM_ERR_INVALID, "basis degree must be 0 or 1");
    if (dtype < BEM_F64 || dtype > BEM_C64) return set_error(BEM_ERR_INVALID, "unknown dtype");
    OpParams p = make_params(ctx, nullptr);
    p.x = x; p.conj_x = conj_x; p.dtype_x = dtype; p.y = y; p.ar = alpha;
    if (deg_test)
        hipLaunchKernelGGL(mass_apply_vertex_kernel, dim3((p.nv + 255) / 256), dim3(256), 0, s, p,
                           ctx->mesh.mass_w, deg_trial);
    else
        hipLaunchKernelGGL(mass_apply_kernel, dim3((p.nf + 255) / 256), dim3(256), 0, s, p, ctx->mesh.mass_w,
                           deg_trial);
    return check_hip(hipGetLastError(), "mass apply launch");
}

int assemble_combined_batched(Context* ctx, int count, const CfInput* probs, int dtype, int precision,
                              int flags, const bem_tuning* tune, hipStream_t s) {
    if (count <= 0) return BEM_OK;
    if (!probs) return set_error(BEM_ERR_INVALID, "null problem array");
    if (dtype != BEM_C128 && dtype != BEM_C64)
        return set_error(BEM_ERR_INVALID, "the combined-field operator is complex: dtype must be c128 or c64");
    if (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32) return set_error(BEM_ERR_INVALID, "unknown precision");
    if (flags & ~(BEM_FLAG_ROTATE_NEAR | BEM_FLAG_SYM_COINCIDE)) return set_error(BEM_ERR_INVALID, "unknown flags");
    int ti = tune && tune->test_panels_per_block > 0 ? tune->test_panels_per_block : 2;
    int bpc = tune && tune->blocks_per_cu > 0 ? tune->blocks_per_cu : 4;
    if (ti != 1 && ti != 2 && ti != 4 && ti != 8 && ti != 16)
        return set_error(BEM_ERR_INVALID, "test_panels_per_block must be 1, 2, 4, 8 or 16");
    if (bpc > 64) return set_error(BEM_ERR_INVALID, "blocks_per_cu must be at most 64");
    std::vector<CfProb> tbl(count);
    int64_t nfar = 0, nnear = 0, nmass = 0;
    for (int q = 0; q < count; ++q) {
        const CfInput& pr = probs[q];
        if (!pr.ctx || !pr.A) return set_error(BEM_ERR_INVALID, "null context or matrix in the batch");
        Context* c = pr.ctx;
        if (!c->mesh.nf) return set_error(BEM_ERR_INVALID, "a problem's context has no mesh");
        if (pr.ld < c->mesh.nv) return set_error(BEM_ERR_INVALID, "matrix row stride below the vertex count");
        int rc = mesh_build_near(c, flags, s);
        if (rc) return rc;
        bem_op_desc d{};
        d.op = BEM_OP_V; d.kernel = BEM_HELMHOLTZ; d.k = pr.k; d.sqrt_n = 1.0; d.flags = flags;
        CfProb& P = tbl[q];
        P.p = make_params(c, &d);
        P.p.out = pr.A; P.p.ld = pr.ld; P.p.dtype = dtype;
        P.mass_w = c->mesh.mass_w;
        P.eta = pr.eta;
        P.nib = (c->mesh.nf + ti - 1) / ti;
        P.njb = (c->mesh.nf + 255) / 256;
        P.far_begin = nfar;
        P.near_begin = nnear;
        P.mass_begin = nmass;
        {
            // upper-triangle items: test groups of 256-row group g pair with trial blocks g .. njb-1
            const int gp = 256 / ti;
            for (int ib0 = 0; ib0 < P.nib; ib0 += gp) {
                const int g = ib0 / gp;
                nfar += (int64_t)std::min(gp, P.nib - ib0) * (P.njb - g);
            }
        }
        nnear += c->mesh.n_near;
        nmass += c->mesh.nf;
    }
    const size_t tb = sizeof(CfProb) * count;
    if (ctx->cf_bytes < tb) {
        if (ctx->cf_ws) hipFree(ctx->cf_ws);
        ctx->cf_ws = nullptr;
        ctx->cf_bytes = 0;
        ctx->cf_table.clear();
        int rc = check_hip(hipMalloc(&ctx->cf_ws, tb), "combined-field table");
        if (rc) return rc;
        ctx->cf_bytes = tb;
    }
    if (ctx->cf_table.size() != tb || memcmp(ctx->cf_table.data(), tbl.data(), tb) != 0) {
        int rc = check_hip(hipMemcpyAsync(ctx->cf_ws, tbl.data(), tb, hipMemcpyHostToDevice, s), "combined-field table");
        if (rc) return rc;
        rc = check_hip(hipStreamSynchronize(s), "combined-field table sync");
        if (rc) return rc;
        ctx->cf_table.assign((const char*)tbl.data(), (const char*)tbl.data() + tb);
    }
    if (ctx->num_cu <= 0) {
        hipDeviceProp_t prop;
        int rc = check_hip(hipGetDeviceProperties(&prop, ctx->device), "device properties");
        if (rc) return rc;
        ctx->num_cu = prop.multiProcessorCount;
    }
    const CfProb* d_tbl = (const CfProb*)ctx->cf_ws;
    const int64_t grid_cap = (int64_t)ctx->num_cu * bpc;
    auto grid_of = [&](int64_t items) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(items, grid_cap)); };
    using CfFn = void (*)(const CfProb*, int, int64_t);
    CfFn far = nullptr;
    const bool f32 = precision == BEM_PREC_FP32;
    // FP32 with GaussQR 5 on every mesh of the batch: the unrolled 25-point loop
    bool q25 = true;
    for (int q = 0; q < count; ++q) q25 = q25 && probs[q].ctx->mesh.q2 == 25;
    switch (ti) {
        case 1: far = !f32 ? cf_far_kernel<double, 1> : q25 ? cf_far_kernel<float, 1, 25> : cf_far_kernel<float, 1>; break;
        case 2: far = !f32 ? cf_far_kernel<double, 2> : q25 ? cf_far_kernel<float, 2, 25> : cf_far_kernel<float, 2>; break;
        case 4: far = !f32 ? cf_far_kernel<double, 4> : q25 ? cf_far_kernel<float, 4, 25> : cf_far_kernel<float, 4>; break;
        case 8: far = !f32 ? cf_far_kernel<double, 8> : q25 ? cf_far_kernel<float, 8, 25> : cf_far_kernel<float, 8>; break;
        default:
            far = !f32 ? cf_far_kernel<double, 16> : q25 ? cf_far_kernel<float, 16, 25> : cf_far_kernel<float, 16>;
            break;
    }
    hipLaunchKernelGGL(far, dim3(grid_of(nfar)), dim3(256), 0, s, d_tbl, count, nfar);
    if (nnear > 0)
        hipLaunchKernelGGL(cf_near_kernel, dim3(grid_of((nnear + 3) / 4)), dim3(256), 0, s, d_tbl, count, nnear);
    hipLaunchKernelGGL(cf_mass_kernel, dim3(grid_of((nmass + 255) / 256)), dim3(256), 0, s, d_tbl, count, nmass);
    return check_hip(hipGetLastError(), "combined-field launch");
}

}  // namespace bem
