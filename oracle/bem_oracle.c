/*
 * bem_oracle.c -- CPU FP64 restatement of the reference BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and the reported CPU baseline).
 * The product path never links or calls this file.
 *
 * Loop nest kept from the reference Taichi kernels: parallel over test panels,
 * serial over trial panels, basis pairs and the GaussQR^4 tensor points
 * (e.g. src/BEM_functions/single_layer/single_layer_3d.py:385-433).  Unlike
 * the reference, one kernel evaluation is shared by all basis pairs of a
 * quadrature point (the reference re-evaluates it per basis pair); the sum is
 * the same integral.
 */
#include "bem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BO_PI 3.14159265358979323846
#define BO_QMAX 12

/* ------------------------------------------------------------------ */
/* Gauss-Legendre rule: src/BEM_functions/utils/functions.py:5-85.     */
/* The reference stores 16-digit literals for q in {1..8,12} and maps  */
/* them to [0,1] with (p+1)/2 and w/2; here the nodes are the roots of */
/* P_q found by Newton iteration from the Chebyshev guess.             */
/* ------------------------------------------------------------------ */
int bo_gauss01(int q, double *pts, double *wts)
{
    if (!((q >= 1 && q <= 8) || q == 12)) return -1;
    for (int i = 0; i < q; ++i) {
        /* root i of P_q in ascending order */
        double x = -cos(BO_PI * (i + 0.75) / (q + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            if (q == 1) { p1 = x; p0 = 1.0; }
            for (int n = 2; n <= q; ++n) {
                double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
                p0 = p1; p1 = p2;
            }
            /* p1 = P_q(x), p0 = P_{q-1}(x) */
            dp = q * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        {
            double p0 = 1.0, p1 = x;
            for (int n = 2; n <= q; ++n) {
                double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
                p0 = p1; p1 = p2;
            }
            dp = q * (x * p1 - p0) / (x * x - 1.0);
        }
        double w = 2.0 / ((1.0 - x * x) * dp * dp);
        if (q == 1) { x = 0.0; w = 2.0; }
        pts[i] = (x + 1.0) / 2.0;
        wts[i] = w / 2.0;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Geometry: src/managers/mesh_manager.py:114-134                      */
/* ------------------------------------------------------------------ */
static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

void bo_geometry(int nv, int nf, const double *verts, const int32_t *tris,
                 double *area, double *normal, double *vert_area, double *vert_normal)
{
    for (int v = 0; v < nv; ++v) {
        if (vert_area) vert_area[v] = 0.0;
        if (vert_normal) vert_normal[3 * v] = vert_normal[3 * v + 1] = vert_normal[3 * v + 2] = 0.0;
    }
    for (int i = 0; i < nf; ++i) {
        const double *x1 = verts + 3 * tris[3 * i], *x2 = verts + 3 * tris[3 * i + 1], *x3 = verts + 3 * tris[3 * i + 2];
        double e1[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
        double e2[3] = {x3[0] - x2[0], x3[1] - x2[1], x3[2] - x2[2]};
        double c[3];
        cross3(e1, e2, c);
        double len = sqrt(dot3(c, c));
        area[i] = 0.5 * len;
        normal[3 * i] = c[0] / len; normal[3 * i + 1] = c[1] / len; normal[3 * i + 2] = c[2] / len;
    }
    for (int f = 0; f < 3 * nf; ++f) {
        int v = tris[f], i = f / 3;
        if (vert_area) vert_area[v] += area[i] / 3.0;
        if (vert_normal) for (int d = 0; d < 3; ++d) vert_normal[3 * v + d] += normal[3 * i + d];
    }
    if (vert_normal)
        for (int v = 0; v < nv; ++v) {
            double l = sqrt(dot3(vert_normal + 3 * v, vert_normal + 3 * v));
            for (int d = 0; d < 3; ++d) vert_normal[3 * v + d] /= l;
        }
}

/* ------------------------------------------------------------------ */
/* Panel relation: src/managers/mesh_manager.py:538-573                */
/* ------------------------------------------------------------------ */
int bo_relation(const int32_t *tris, int i, int j)
{
    if (i == j) return BO_COINCIDE;
    int rel = BO_SEPARATE, common_vertex = 0;
    for (int a = 0; a < 3; ++a) {
        int p1 = tris[3 * i + a], p2 = tris[3 * i + (a + 1) % 3];
        for (int b = 0; b < 3; ++b) {
            int q1 = tris[3 * j + b], q2 = tris[3 * j + (b + 1) % 3];
            if (p1 == q1 && p2 == q2) rel = BO_COMMON_EDGE;
            if (p1 == q2 && p2 == q1) rel = BO_COMMON_EDGE;
            if (p1 == q1) common_vertex = 1;
        }
    }
    if (common_vertex && rel != BO_COMMON_EDGE) rel = BO_COMMON_VERTEX;
    return rel;
}

void bo_relation_counts(int nf, const int32_t *tris, int64_t counts[4])
{
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nf; ++j) counts[bo_relation(tris, i, j)]++;
}

/* ------------------------------------------------------------------ */
/* Quadrature sub-samples of one tensor point (xi, e1, e2, e3).        */
/* SEPARATE   single_layer_3d.py:155-177                               */
/* COINCIDE   single_layer_3d.py:178-206 (maps BEM_manager.py:141-165) */
/* COMMON_VERTEX single_layer_3d.py:224-264                            */
/* COMMON_EDGE   single_layer_3d.py:265-367                            */
/* ------------------------------------------------------------------ */
typedef struct { double r1x, r2x, r1y, r2y, jac; } sub_t;

static int subsamples(int rel, double xi, double e1, double e2, double e3, sub_t *s)
{
    if (rel == BO_SEPARATE) {
        s[0].r1x = xi; s[0].r2x = e1 * xi; s[0].r1y = e2; s[0].r2y = e3 * e2;
        s[0].jac = xi * e2;
        return 1;
    }
    if (rel == BO_COINCIDE) {
        double jac = xi * xi * xi * e1 * e1 * e2;
        double w0 = xi, w1 = xi * e1, w2 = xi * e1 * e2, w3 = xi * e1 * e2 * e3;
        double z[3][4] = {
            {w0, w0 - w1 + w2, w3, w2},            /* rows of map 0 applied to w */
            {w0, w1 - w2 + w3, w2, w3},            /* map 1 */
            {w0 - w3, w1 - w3, -w3, w2 - w3},      /* map 2 */
        };
        for (int m = 0; m < 3; ++m) {
            s[m].r1x = z[m][0]; s[m].r2x = z[m][1];
            s[m].r1y = z[m][0] - z[m][2]; s[m].r2y = z[m][1] - z[m][3];
            s[m].jac = jac;
        }
        return 3;
    }
    if (rel == BO_COMMON_VERTEX) {
        double jac = xi * xi * xi * e2;
        s[0].r1x = xi; s[0].r2x = xi * e1; s[0].r1y = xi * e2; s[0].r2y = xi * e2 * e3; s[0].jac = jac;
        s[1].r1x = xi * e2; s[1].r2x = xi * e2 * e3; s[1].r1y = xi; s[1].r2y = xi * e1; s[1].jac = jac;
        return 2;
    }
    /* COMMON_EDGE: w = (w0, w1, w2, w3); x = (w0, w3); y = (w0 + w1, w2) */
    double w[5][4] = {
        {xi, -xi * e1 * e2, xi * e1 * (1.0 - e2), xi * e1 * e3},
        {xi, -xi * e1 * e2 * e3, xi * e1 * e2 * (1.0 - e3), xi * e1},
        {xi * (1.0 - e1 * e2), xi * e1 * e2, xi * e1 * e2 * e3, xi * e1 * (1.0 - e2)},
        {xi * (1.0 - e1 * e2 * e3), xi * e1 * e2 * e3, xi * e1, xi * e1 * e2 * (1.0 - e3)},
        {xi * (1.0 - e1 * e2 * e3), xi * e1 * e2 * e3, xi * e1 * e2, xi * e1 * (1.0 - e2 * e3)},
    };
    for (int m = 0; m < 5; ++m) {
        s[m].r1x = w[m][0]; s[m].r2x = w[m][3];
        s[m].r1y = w[m][0] + w[m][1]; s[m].r2y = w[m][2];
        s[m].jac = (m == 0) ? xi * xi * xi * e1 * e1 : xi * xi * xi * e1 * e1 * e2;
    }
    return 5;
}

/* P1 shape function, BEM_manager.py:404-425; slot < 0 is the P0 constant. */
static double shape(double r1, double r2, int slot)
{
    if (slot == 0) return 1.0 - r1;
    if (slot == 1) return r1 - r2;
    if (slot == 2) return r2;
    return 1.0;
}

static void interp(double r1, double r2, const double *p1, const double *p2, const double *p3, double *x)
{
    for (int d = 0; d < 3; ++d) x[d] = (1.0 - r1) * p1[d] + (r1 - r2) * p2[d] + r2 * p3[d];
}

/* G, BEM_manager.py:452-485 (3D branch).  out = (re, im). */
static void kern_G(int kernel, double k, double sqrt_n, const double *x, const double *y, double *g)
{
    double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double r = sqrt(dot3(d, d));
    double c = 1.0 / 4.0 / BO_PI / r;
    if (kernel == BO_LAPLACE) { g[0] = c; g[1] = 0.0; return; }
    double kap = (kernel == BO_HELMHOLTZ_TRANSMISSION) ? k * sqrt_n : k;
    g[0] = c * cos(kap * r);
    g[1] = c * sin(kap * r);
}

/* grad_G_y, BEM_manager.py:487-525 (3D branch). */
static void kern_dG(int kernel, double k, double sqrt_n, const double *x, const double *y,
                    const double *ny, double *g)
{
    double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double r = sqrt(dot3(d, d));
    double dir[3] = {d[0] / r, d[1] / r, d[2] / r};
    double base = (1.0 / 4.0 / BO_PI) * dot3(dir, ny) / (r * r);
    if (kernel == BO_LAPLACE) { g[0] = base; g[1] = 0.0; return; }
    double kap = (kernel == BO_HELMHOLTZ_TRANSMISSION) ? k * sqrt_n : k;
    double c = cos(kap * r), s = sin(kap * r);
    /* (1 - i kap r) * exp(i kap r) */
    g[0] = base * (c + kap * r * s);
    g[1] = base * (s - kap * r * c);
}

/* surface gradient of the P1 basis `slot`, hypersingular_layer_3d.py:56-64,
 * and its curl n x grad (:163-169). */
static void basis_curl(const double *p1, const double *p2, const double *p3, const double *n,
                       double area, int slot, double *curl)
{
    double u[3] = {slot == 0 ? 1.0 : 0.0, slot == 1 ? 1.0 : 0.0, slot == 2 ? 1.0 : 0.0};
    double e32[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
    double e13[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
    double e21[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double c1[3], c2[3], c3[3], g[3];
    cross3(n, e32, c1); cross3(n, e13, c2); cross3(n, e21, c3);
    for (int d = 0; d < 3; ++d) g[d] = (u[0] * c1[d] + u[1] * c2[d] + u[2] * c3[d]) / (2.0 * area);
    cross3(n, g, curl);
}

/* Vertex slot order of panel i, after the optional near-field rotation. */
static void near_perm(const int32_t *tris, int i, int j, int rel, int flags, int px[3], int py[3])
{
    px[0] = 0; px[1] = 1; px[2] = 2; py[0] = 0; py[1] = 1; py[2] = 2;
    if (!(flags & BO_FLAG_ROTATE_NEAR)) return;
    const int32_t *ti = tris + 3 * i, *tj = tris + 3 * j;
    if (rel == BO_COMMON_VERTEX) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (ti[a] == tj[b]) {
                    for (int m = 0; m < 3; ++m) { px[m] = (a + m) % 3; py[m] = (b + m) % 3; }
                    return;
                }
    } else if (rel == BO_COMMON_EDGE) {
        /* rotate x so its shared edge runs slot0 -> slot1 ... */
        for (int a = 0; a < 3; ++a) {
            int u = ti[a], v = ti[(a + 1) % 3];
            int bu = -1, bv = -1;
            for (int b = 0; b < 3; ++b) { if (tj[b] == u) bu = b; if (tj[b] == v) bv = b; }
            if (bu >= 0 && bv >= 0) {
                px[0] = a; px[1] = (a + 1) % 3; px[2] = (a + 2) % 3;
                /* ... and order y as (u, v, rest), the same direction */
                py[0] = bu; py[1] = bv; py[2] = 3 - bu - bv;
                return;
            }
        }
    }
}

/* One panel-pair integral for every basis pair (see header). */
static int pair_impl(const bo_op *op, const double *gp, const double *gw, const double *verts,
                     const int32_t *tris, const double *area, const double *normal,
                     int i, int j, int rel, double *out)
{
    const int q = op->q;
    const int ns = op->deg_test ? 3 : 1, nt = op->deg_trial ? 3 : 1;
    int px[3], py[3];
    near_perm(tris, i, j, rel, op->flags, px, py);
    const double *x1 = verts + 3 * tris[3 * i + px[0]], *x2 = verts + 3 * tris[3 * i + px[1]],
                 *x3 = verts + 3 * tris[3 * i + px[2]];
    const double *y1 = verts + 3 * tris[3 * j + py[0]], *y2 = verts + 3 * tris[3 * j + py[1]],
                 *y3 = verts + 3 * tris[3 * j + py[2]];
    const double *nx = normal + 3 * i, *ny = normal + 3 * j;
    const double ax = area[i], ay = area[j];
    const double k2 = op->k * op->k * op->sqrt_n * op->sqrt_n; /* hypersingular_layer_3d.py:171 */
    const double nxny = dot3(nx, ny);

    /* slot (in stored order) of the test/trial basis s/t after permutation:
       basis s of the output refers to stored slot s; in the permuted
       parametrisation that vertex sits at position inv[s]. */
    int invx[3], invy[3];
    for (int m = 0; m < 3; ++m) { invx[px[m]] = m; invy[py[m]] = m; }

    double curlcurl[9];
    if (op->op == BO_OP_W) {
        double cx[3][3], cy[3][3];
        for (int s = 0; s < 3; ++s) {
            basis_curl(verts + 3 * tris[3 * i], verts + 3 * tris[3 * i + 1], verts + 3 * tris[3 * i + 2], nx, ax, s, cx[s]);
            basis_curl(verts + 3 * tris[3 * j], verts + 3 * tris[3 * j + 1], verts + 3 * tris[3 * j + 2], ny, ay, s, cy[s]);
        }
        for (int s = 0; s < 3; ++s)
            for (int t = 0; t < 3; ++t) curlcurl[s * 3 + t] = op->deg_test && op->deg_trial ? dot3(cx[s], cy[t]) : 0.0;
    }

    for (int m = 0; m < ns * nt * 2; ++m) out[m] = 0.0;
    sub_t sub[5];
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b)
            for (int c = 0; c < q; ++c)
                for (int d = 0; d < q; ++d) {
                    const double weight = (gw[a] * gw[b] * (ax * 2.0)) * (gw[c] * gw[d] * (ay * 2.0));
                    const int nsub = subsamples(rel, gp[a], gp[b], gp[c], gp[d], sub);
                    const int twice = (rel == BO_COINCIDE);
                    for (int m = 0; m < nsub; ++m) {
                        for (int pass = 0; pass <= twice; ++pass) {
                            double rx1 = sub[m].r1x, rx2 = sub[m].r2x, ry1 = sub[m].r1y, ry2 = sub[m].r2y;
                            double x[3], y[3], f[2];
                            interp(rx1, rx2, x1, x2, x3, x);
                            interp(ry1, ry2, y1, y2, y3, y);
                            /* basis coordinates: the second coincident pass swaps
                               them only under BO_FLAG_SYM_COINCIDE */
                            double bx1 = rx1, bx2 = rx2, by1 = ry1, by2 = ry2;
                            if (pass == 1 && (op->flags & BO_FLAG_SYM_COINCIDE)) {
                                bx1 = ry1; bx2 = ry2; by1 = rx1; by2 = rx2;
                            }
                            const double *xe = pass ? y : x, *ye = pass ? x : y;
                            if (op->op == BO_OP_V || op->op == BO_OP_W) kern_G(op->kernel, op->k, op->sqrt_n, xe, ye, f);
                            else if (op->op == BO_OP_K) kern_dG(op->kernel, op->k, op->sqrt_n, xe, ye, ny, f);
                            else kern_dG(op->kernel, op->k, op->sqrt_n, ye, xe, nx, f);
                            const double wj = weight * sub[m].jac;
                            for (int s = 0; s < ns; ++s) {
                                double phx = shape(bx1, bx2, op->deg_test ? invx[s] : -1);
                                for (int t = 0; t < nt; ++t) {
                                    double phy = shape(by1, by2, op->deg_trial ? invy[t] : -1);
                                    double fac;
                                    if (op->op == BO_OP_W)
                                        fac = (op->deg_test && op->deg_trial ? curlcurl[s * 3 + t] : 0.0) - k2 * phx * phy * nxny;
                                    else
                                        fac = phx * phy;
                                    out[(s * nt + t) * 2 + 0] += f[0] * fac * wj;
                                    out[(s * nt + t) * 2 + 1] += f[1] * fac * wj;
                                }
                            }
                        }
                    }
                }
    return 0;
}

int bo_pair(const bo_op *op, const double *verts, const int32_t *tris,
            const double *area, const double *normal, int i, int j, int rel, double *out)
{
    double gp[BO_QMAX], gw[BO_QMAX];
    if (bo_gauss01(op->q, gp, gw)) return -1;
    return pair_impl(op, gp, gw, verts, tris, area, normal, i, j, rel, out);
}

int bo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int bo_assemble_rows(const bo_op *op, int nv, int nf, const double *verts,
                     const int32_t *tris, const double *area, const double *normal,
                     int row_begin, int row_end, double *out, int nthreads)
{
    double gp[BO_QMAX], gw[BO_QMAX];
    if (bo_gauss01(op->q, gp, gw)) return -1;
    const int ns = op->deg_test ? 3 : 1, nt = op->deg_trial ? 3 : 1;
    const int64_t cols = op->deg_trial ? nv : nf;
    set_threads(nthreads);
    int i;
#pragma omp parallel for schedule(dynamic)
    for (i = row_begin; i < row_end; ++i) {
        double blk[18];
        for (int j = 0; j < nf; ++j) {
            int rel = bo_relation(tris, i, j);
            pair_impl(op, gp, gw, verts, tris, area, normal, i, j, rel, blk);
            for (int s = 0; s < ns; ++s) {
                /* P0 test: out holds rows [row_begin, row_end); P1 test: all
                   vertex rows (a vertex row collects from several panels) */
                int64_t row = op->deg_test ? tris[3 * i + s] : (int64_t)(i - row_begin);
                for (int t = 0; t < nt; ++t) {
                    int64_t col = op->deg_trial ? tris[3 * j + t] : j;
                    double *o = out + 2 * (row * cols + col);
                    if (op->deg_test) {
#pragma omp atomic
                        o[0] += blk[(s * nt + t) * 2];
#pragma omp atomic
                        o[1] += blk[(s * nt + t) * 2 + 1];
                    } else {
                        o[0] += blk[(s * nt + t) * 2];
                        o[1] += blk[(s * nt + t) * 2 + 1];
                    }
                }
            }
        }
    }
    (void)nv;
    return 0;
}

int bo_assemble(const bo_op *op, int nv, int nf, const double *verts,
                const int32_t *tris, const double *area, const double *normal,
                double *out, int nthreads)
{
    return bo_assemble_rows(op, nv, nf, verts, tris, area, normal, 0, nf, out, nthreads);
}

int bo_apply(const bo_op *op, int nv, int nf, const double *verts,
             const int32_t *tris, const double *area, const double *normal,
             const double *x, int conj_x, double alpha, double *y, int nthreads)
{
    return bo_apply_rows(op, nv, nf, verts, tris, area, normal, x, conj_x, alpha, 0, nf, y, nthreads);
}

int bo_apply_rows(const bo_op *op, int nv, int nf, const double *verts,
                  const int32_t *tris, const double *area, const double *normal,
                  const double *x, int conj_x, double alpha, int row_begin, int row_end,
                  double *y, int nthreads)
{
    double gp[BO_QMAX], gw[BO_QMAX];
    if (bo_gauss01(op->q, gp, gw)) return -1;
    const int ns = op->deg_test ? 3 : 1, nt = op->deg_trial ? 3 : 1;
    const double sg = conj_x ? -1.0 : 1.0;
    set_threads(nthreads);
    int i;
#pragma omp parallel for schedule(dynamic)
    for (i = row_begin; i < row_end; ++i) {
        double blk[18], acc[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < nf; ++j) {
            int rel = bo_relation(tris, i, j);
            pair_impl(op, gp, gw, verts, tris, area, normal, i, j, rel, blk);
            for (int s = 0; s < ns; ++s)
                for (int t = 0; t < nt; ++t) {
                    int col = op->deg_trial ? tris[3 * j + t] : j;
                    double br = blk[(s * nt + t) * 2], bi = blk[(s * nt + t) * 2 + 1];
                    double xr = x[2 * col], xi = sg * x[2 * col + 1];
                    acc[2 * s] += br * xr - bi * xi;
                    acc[2 * s + 1] += br * xi + bi * xr;
                }
        }
        for (int s = 0; s < ns; ++s) {
            int row = op->deg_test ? tris[3 * i + s] : i;
#pragma omp atomic
            y[2 * row] += alpha * acc[2 * s];
#pragma omp atomic
            y[2 * row + 1] += alpha * acc[2 * s + 1];
        }
    }
    (void)nv;
    return 0;
}

int bo_mass(int q, int deg_test, int deg_trial, int nv, int nf,
            const int32_t *tris, const double *area, double alpha, double *out)
{
    double gp[BO_QMAX], gw[BO_QMAX];
    if (bo_gauss01(q, gp, gw)) return -1;
    const int ns = deg_test ? 3 : 1, nt = deg_trial ? 3 : 1;
    const int64_t cols = deg_trial ? nv : nf;
    for (int i = 0; i < nf; ++i)
        for (int s = 0; s < ns; ++s)
            for (int t = 0; t < nt; ++t) {
                double v = 0.0;
                for (int a = 0; a < q; ++a)
                    for (int b = 0; b < q; ++b) {
                        double r1 = gp[a], r2 = gp[b] * r1;
                        double w = gw[a] * gw[b] * (area[i] * 2.0);
                        v += shape(r1, r2, deg_test ? s : -1) * shape(r1, r2, deg_trial ? t : -1) * w * r1;
                    }
                int64_t row = deg_test ? tris[3 * i + s] : i;
                int64_t col = deg_trial ? tris[3 * i + t] : i;
                out[2 * (row * cols + col)] += alpha * v;
            }
    return 0;
}

int bo_rhs_dirichlet(const bo_op *kop, int nv, int nf, const double *verts,
                     const int32_t *tris, const double *area, const double *normal,
                     const double *g, int conj_g, double s, double *rhs, int nthreads)
{
    /* s/2 * M conj(g): multiply_identity_to_vector, rhs_constructor_3d.py:256-311 */
    double gp[BO_QMAX], gw[BO_QMAX];
    if (bo_gauss01(kop->q, gp, gw)) return -1;
    const double sg = conj_g ? -1.0 : 1.0;
    for (int i = 0; i < nf; ++i) {
        for (int t = 0; t < 3; ++t) {
            double v = 0.0;
            for (int a = 0; a < kop->q; ++a)
                for (int b = 0; b < kop->q; ++b) {
                    double r1 = gp[a], r2 = gp[b] * r1;
                    v += shape(r1, r2, -1) * shape(r1, r2, t) * (gw[a] * gw[b] * (area[i] * 2.0)) * r1;
                }
            int vtx = tris[3 * i + t];
            rhs[2 * i] += 0.5 * s * v * g[2 * vtx];
            rhs[2 * i + 1] += 0.5 * s * v * sg * g[2 * vtx + 1];
        }
    }
    /* + K conj(g): apply_K_dot_D_boundary with multiplier 1 (:361-372) */
    bo_op k = *kop;
    k.op = BO_OP_K; k.deg_test = 0; k.deg_trial = 1;
    return bo_apply(&k, nv, nf, verts, tris, area, normal, g, conj_g, 1.0, rhs, nthreads);
}
