// bem_geometry.hip -- mesh upload, panel geometry, quadrature tables and the
// near-field pair list.
//
// Geometry restates src/managers/mesh_manager.py:114-134 on the device;
// the Gauss tables restate src/BEM_functions/utils/functions.py:5-85 and the
// relation classes src/managers/mesh_manager.py:538-573.
#include <math.h>

#include <algorithm>
#include <unordered_map>

#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

namespace {

// One thread per panel: area = 1/2 |(x2-x1) x (x3-x2)|, unit normal of the
// same cross product, centroid, and the curls n x grad(phi_s) of the three
// P1 bases (hypersingular_layer_3d.py:56-64,163-169).
__global__ void panel_geometry_kernel(int nf, const double* __restrict__ v,
                                      const int32_t* __restrict__ t, double* area, double* normal,
                                      double* cen, double* curl) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const double* p1 = v + 3 * t[3 * i];
    const double* p2 = v + 3 * t[3 * i + 1];
    const double* p3 = v + 3 * t[3 * i + 2];
    double e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double e2[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
    double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double a = 0.5 * len;
    double n[3] = {c[0] / len, c[1] / len, c[2] / len};
    area[i] = a;
    for (int d = 0; d < 3; ++d) {
        normal[3 * i + d] = n[d];
        cen[3 * i + d] = (p1[d] + p2[d] + p3[d]) / 3.0;
    }
    // opposite edges of slots 0,1,2: x3-x2, x1-x3, x2-x1
    const double* ea[3] = {p3, p1, p2};
    const double* eb[3] = {p2, p3, p1};
    for (int s = 0; s < 3; ++s) {
        double e[3] = {ea[s][0] - eb[s][0], ea[s][1] - eb[s][1], ea[s][2] - eb[s][2]};
        double g[3] = {(n[1] * e[2] - n[2] * e[1]) / (2.0 * a), (n[2] * e[0] - n[0] * e[2]) / (2.0 * a),
                       (n[0] * e[1] - n[1] * e[0]) / (2.0 * a)};
        double cu[3] = {n[1] * g[2] - n[2] * g[1], n[2] * g[0] - n[0] * g[2], n[0] * g[1] - n[1] * g[0]};
        for (int d = 0; d < 3; ++d) curl[9 * i + 3 * s + d] = cu[d];
    }
}

// Vertex areas (sum of A/3) and normals (normalised sum of panel normals).
// One thread per vertex over a vertex->panel CSR, so the sums are ordered
// (panel order) and deterministic.
__global__ void vertex_geometry_kernel(int nv, const int32_t* __restrict__ vp_row,
                                       const int32_t* __restrict__ vp_el,
                                       const double* __restrict__ area,
                                       const double* __restrict__ normal, double* va, double* vn) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double a = 0.0, n[3] = {0.0, 0.0, 0.0};
    for (int e = vp_row[v]; e < vp_row[v + 1]; ++e) {
        int i = vp_el[e] / 3;
        a += area[i] / 3.0;
        for (int d = 0; d < 3; ++d) n[d] += normal[3 * i + d];
    }
    double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    va[v] = a;
    for (int d = 0; d < 3; ++d) vn[3 * v + d] = n[d] / l;
}

// Regular-rule points of each panel: (r1, r2) = (xi_a, xi_a eta_b),
// x = (1-r1) x1 + (r1-r2) x2 + r2 x3 (BEM_manager.py:427-450), weight
// w_a w_b xi_a (2A) (single_layer_3d.py:150-164).  Writes the FP64 AoS and
// SoA tables and the FP32 centroid-offset tables.
// Local frame of each panel: C, U = x1 - C, E1 = x2 - x1, E2 = x3 - x2, the
// normal and D = -2 U (Mesh::frame).  A point of the panel is
// x(r1, r2) - C = U + r1 E1 + r2 E2, the element map
// (1 - r1) x1 + (r1 - r2) x2 + r2 x3 (BEM_manager.py:427-450) moved to C.
__global__ void panel_frame_kernel(int nf, const double* __restrict__ v, const int32_t* __restrict__ t,
                                   const double* __restrict__ normal, const double* __restrict__ cen,
                                   double* frame) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const double* p1 = v + 3 * t[3 * i];
    const double* p2 = v + 3 * t[3 * i + 1];
    const double* p3 = v + 3 * t[3 * i + 2];
    double* f = frame + (int64_t)i * PANEL_FRAME;
    for (int d = 0; d < 3; ++d) {
        const double c = cen[3 * i + d], u = p1[d] - c;
        f[d] = c;
        f[3 + d] = u;
        f[6 + d] = p2[d] - p1[d];
        f[9 + d] = p3[d] - p2[d];
        f[12 + d] = normal[3 * i + d];
        f[15 + d] = -2.0 * u;
    }
    f[18] = 0.0;
    f[19] = 0.0;
}

__global__ void panel_points_kernel(int nf, int q2, const double* __restrict__ v,
                                    const int32_t* __restrict__ t, const double* __restrict__ area,
                                    const double* __restrict__ cen, const double* __restrict__ r12w,
                                    const double* __restrict__ frame, double* pu, double* pl, float* pu32,
                                    float* pl32, double* pq, float* pq32) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf * q2) return;
    int i = idx / q2, p = idx % q2;
    double r1 = r12w[3 * p], r2 = r12w[3 * p + 1], w = r12w[3 * p + 2] * (area[i] * 2.0);
    {
        // |x - C|^2 with x - C = U + r1 E1 + r2 E2, as the moment-form kernels rebuild it
        const double* f = frame + (int64_t)i * PANEL_FRAME;
        double s2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double o = f[3 + d] + r1 * f[6 + d] + r2 * f[9 + d];
            s2 += o * o;
        }
        const int64_t b = ((int64_t)i * q2 + p) * 4;
        pq[b + 0] = s2;
        pq[b + 1] = -2.0 * r1;
        pq[b + 2] = -2.0 * r2;
        pq[b + 3] = w;
        pq32[b + 0] = (float)s2;
        pq32[b + 1] = (float)(-2.0 * r1);
        pq32[b + 2] = (float)(-2.0 * r2);
        pq32[b + 3] = (float)w;
    }
    const double* p1 = v + 3 * t[3 * i];
    const double* p2 = v + 3 * t[3 * i + 1];
    const double* p3 = v + 3 * t[3 * i + 2];
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = (1.0 - r1) * p1[d] + (r1 - r2) * p2[d] + r2 * p3[d];
    int64_t base = ((int64_t)i * q2 + p) * 4;
    pu[base + 0] = x[0]; pu[base + 1] = x[1]; pu[base + 2] = x[2]; pu[base + 3] = w;
    for (int c = 0; c < 3; ++c) pl[((int64_t)p * 4 + c) * nf + i] = x[c];
    pl[((int64_t)p * 4 + 3) * nf + i] = w;
    // FP32: local offset from the centroid, (1-r1)(x1-c) + (r1-r2)(x2-c) + r2(x3-c)
    float o[3];
    for (int d = 0; d < 3; ++d) {
        double c = cen[3 * i + d];
        o[d] = (float)((1.0 - r1) * (p1[d] - c) + (r1 - r2) * (p2[d] - c) + r2 * (p3[d] - c));
    }
    pu32[base + 0] = o[0]; pu32[base + 1] = o[1]; pu32[base + 2] = o[2]; pu32[base + 3] = (float)w;
    for (int c = 0; c < 3; ++c) pl32[((int64_t)p * 4 + c) * nf + i] = o[c];
    pl32[((int64_t)p * 4 + 3) * nf + i] = (float)w;
}

// Gauss-Legendre nodes on [0,1] with halved weights, by Newton iteration on
// P_q (the values the reference lists at utils/functions.py:9-79).
bool gauss01(int q, double* pts, double* wts) {
    if (!((q >= 1 && q <= 8) || q == 12)) return false;
    for (int i = 0; i < q; ++i) {
        double x = -cos(M_PI * (i + 0.75) / (q + 0.5));
        double p0 = 1.0, p1 = x, dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            p0 = 1.0; p1 = x;
            for (int n = 2; n <= q; ++n) {
                double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
                p0 = p1; p1 = p2;
            }
            dp = q * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        p0 = 1.0; p1 = x;
        for (int n = 2; n <= q; ++n) {
            double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
            p0 = p1; p1 = p2;
        }
        dp = q * (x * p1 - p0) / (x * x - 1.0);
        double w = 2.0 / ((1.0 - x * x) * dp * dp);
        if (q == 1) { x = 0.0; w = 2.0; }
        pts[i] = (x + 1.0) / 2.0;
        wts[i] = w / 2.0;
    }
    return true;
}

// Near-field rules in parametric coordinates: one entry per (tensor point,
// region), weight = product of the four Gauss weights times the region
// Jacobian.  Regions and Jacobians: single_layer_3d.py:178-367 (coincident
// maps from BEM_manager.py:141-165).
void build_rules(int q, const double* gp, const double* gw, std::vector<NearPt>& out, int32_t off[4],
                 int32_t len[4]) {
    out.clear();
    for (int rel = 1; rel <= 3; ++rel) {
        off[rel] = (int32_t)out.size();
        for (int a = 0; a < q; ++a)
            for (int b = 0; b < q; ++b)
                for (int c = 0; c < q; ++c)
                    for (int d = 0; d < q; ++d) {
                        double xi = gp[a], e1 = gp[b], e2 = gp[c], e3 = gp[d];
                        double w4 = gw[a] * gw[b] * gw[c] * gw[d];
                        if (rel == BEM_COINCIDE) {
                            double jac = xi * xi * xi * e1 * e1 * e2;
                            double w0 = xi, w1 = xi * e1, w2 = xi * e1 * e2, w3 = xi * e1 * e2 * e3;
                            double z[3][4] = {{w0, w0 - w1 + w2, w3, w2},
                                              {w0, w1 - w2 + w3, w2, w3},
                                              {w0 - w3, w1 - w3, -w3, w2 - w3}};
                            for (int m = 0; m < 3; ++m)
                                out.push_back({z[m][0], z[m][1], z[m][0] - z[m][2], z[m][1] - z[m][3], w4 * jac});
                        } else if (rel == BEM_COMMON_VERTEX) {
                            double jac = xi * xi * xi * e2;
                            out.push_back({xi, xi * e1, xi * e2, xi * e2 * e3, w4 * jac});
                            out.push_back({xi * e2, xi * e2 * e3, xi, xi * e1, w4 * jac});
                        } else {
                            double w[5][4] = {
                                {xi, -xi * e1 * e2, xi * e1 * (1.0 - e2), xi * e1 * e3},
                                {xi, -xi * e1 * e2 * e3, xi * e1 * e2 * (1.0 - e3), xi * e1},
                                {xi * (1.0 - e1 * e2), xi * e1 * e2, xi * e1 * e2 * e3, xi * e1 * (1.0 - e2)},
                                {xi * (1.0 - e1 * e2 * e3), xi * e1 * e2 * e3, xi * e1, xi * e1 * e2 * (1.0 - e3)},
                                {xi * (1.0 - e1 * e2 * e3), xi * e1 * e2 * e3, xi * e1 * e2, xi * e1 * (1.0 - e2 * e3)}};
                            for (int m = 0; m < 5; ++m) {
                                double jac = (m == 0) ? xi * xi * xi * e1 * e1 : xi * xi * xi * e1 * e1 * e2;
                                out.push_back({w[m][0], w[m][3], w[m][0] + w[m][1], w[m][2], w4 * jac});
                            }
                        }
                    }
        len[rel] = (int32_t)out.size() - off[rel];
    }
}

// Relation of panels i and j, mesh_manager.py:538-573.
int relation(const int32_t* t, int i, int j) {
    if (i == j) return BEM_COINCIDE;
    int rel = BEM_SEPARATE, cv = 0;
    for (int a = 0; a < 3; ++a) {
        int p1 = t[3 * i + a], p2 = t[3 * i + (a + 1) % 3];
        for (int b = 0; b < 3; ++b) {
            int q1 = t[3 * j + b], q2 = t[3 * j + (b + 1) % 3];
            if ((p1 == q1 && p2 == q2) || (p1 == q2 && p2 == q1)) rel = BEM_COMMON_EDGE;
            if (p1 == q1) cv = 1;
        }
    }
    if (cv && rel != BEM_COMMON_EDGE) rel = BEM_COMMON_VERTEX;
    return rel;
}

int encode_perm(const int* px, const int* py) {
    return px[0] | (px[1] << 2) | (px[2] << 4) | (py[0] << 8) | (py[1] << 10) | (py[2] << 12);
}

// Slot order putting the shared vertex / edge where the Sauter-Schwab rule
// expects it (vertex: slot 0 of both; edge: slots 0->1 of both, same
// direction).  Identity unless BEM_FLAG_ROTATE_NEAR.
void near_perm(const int32_t* t, int i, int j, int rel, int flags, int px[3], int py[3]) {
    px[0] = 0; px[1] = 1; px[2] = 2; py[0] = 0; py[1] = 1; py[2] = 2;
    if (!(flags & BEM_FLAG_ROTATE_NEAR)) return;
    const int32_t *ti = t + 3 * i, *tj = t + 3 * j;
    if (rel == BEM_COMMON_VERTEX) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (ti[a] == tj[b]) {
                    for (int m = 0; m < 3; ++m) { px[m] = (a + m) % 3; py[m] = (b + m) % 3; }
                    return;
                }
    } else if (rel == BEM_COMMON_EDGE) {
        for (int a = 0; a < 3; ++a) {
            int u = ti[a], w = ti[(a + 1) % 3], bu = -1, bw = -1;
            for (int b = 0; b < 3; ++b) { if (tj[b] == u) bu = b; if (tj[b] == w) bw = b; }
            if (bu >= 0 && bw >= 0) {
                px[0] = a; px[1] = (a + 1) % 3; px[2] = (a + 2) % 3;
                py[0] = bu; py[1] = bw; py[2] = 3 - bu - bw;
                return;
            }
        }
    }
}

template <typename T>
int dev_alloc(T** p, size_t n) {
    if (*p) { hipFree(*p); *p = nullptr; }
    if (n == 0) n = 1;
    return check_hip(hipMalloc((void**)p, n * sizeof(T)), "hipMalloc");
}

}  // namespace

void mesh_free(Mesh* m) {
    void* ptrs[] = {m->verts, m->tris, m->area, m->normal, m->vert_area, m->vert_normal, m->curl,
                    m->pu, m->pl, m->cen, m->pu32, m->pl32, m->phi, m->phi32, m->mass_w, m->frame, m->pq,
                    m->near, m->near_row, m->rule, m->near_buf, m->kpart, m->vp_row, m->vp_el,
                    m->near_tr, m->adj, m->panel_buf, m->pq32, m->r12w};
    for (void* p : ptrs)
        if (p) hipFree(p);
    *m = Mesh();
}

namespace {

// The tables that depend on the vertex positions: panel and vertex
// geometry, local frames, regular-rule points.
int mesh_geometry(Mesh& m, hipStream_t s) {
    const int nf = m.nf, nv = m.nv, q2 = m.q2;
    hipLaunchKernelGGL(panel_geometry_kernel, dim3((nf + 255) / 256), dim3(256), 0, s, nf, m.verts,
                       m.tris, m.area, m.normal, m.cen, m.curl);
    hipLaunchKernelGGL(vertex_geometry_kernel, dim3((nv + 255) / 256), dim3(256), 0, s, nv, m.vp_row,
                       m.vp_el, m.area, m.normal, m.vert_area, m.vert_normal);
    hipLaunchKernelGGL(panel_frame_kernel, dim3((nf + 255) / 256), dim3(256), 0, s, nf, m.verts, m.tris,
                       m.normal, m.cen, m.frame);
    hipLaunchKernelGGL(panel_points_kernel, dim3((nf * q2 + 255) / 256), dim3(256), 0, s, nf, q2,
                       m.verts, m.tris, m.area, m.cen, m.r12w, m.frame, m.pu, m.pl, m.pu32, m.pl32, m.pq,
                       m.pq32);
    return check_hip(hipGetLastError(), "geometry kernels");
}

int mesh_build_impl(Context* ctx, const double* d_verts, const int32_t* d_tris, int nv, int nf, int q,
                    hipStream_t s) {
    if (nv <= 0 || nf <= 0) return set_error(BEM_ERR_INVALID, "mesh must have vertices and panels");
    double gp[12], gw[12];
    if (!gauss01(q, gp, gw))
        return set_error(BEM_ERR_UNSUPPORTED, "Please refer N as 2~8 for gaussion interation (GaussQR must be 1..8 or 12)");
    Mesh& m = ctx->mesh;
    int rc = 0;
    // The same panels and rule at new vertex positions (shape derivatives,
    // finite differences, optimisation loops): the buffers, the near-field
    // list and the rule tables stay; the positions and what depends on them
    // are rebuilt.
    if (m.nf == nf && m.nv == nv && m.q == q && m.verts && m.r12w) {
        std::vector<int32_t> t((size_t)nf * 3);
        rc = check_hip(hipMemcpyAsync(t.data(), d_tris, sizeof(int32_t) * nf * 3, hipMemcpyDeviceToHost, s),
                       "panels to host");
        if (rc) return rc;
        rc = check_hip(hipStreamSynchronize(s), "sync");
        if (rc) return rc;
        if (t == m.h_tris) {
            rc = check_hip(hipMemcpyAsync(m.verts, d_verts, sizeof(double) * nv * 3, hipMemcpyDeviceToDevice, s),
                           "copy vertices");
            if (rc) return rc;
            rc = mesh_geometry(m, s);
            if (rc) return rc;
            return check_hip(hipStreamSynchronize(s), "mesh update");
        }
    }
    mesh_free(&m);
    m.nv = nv; m.nf = nf; m.q = q; m.q2 = q * q;
    const int q2 = m.q2;
    rc |= dev_alloc(&m.verts, (size_t)nv * 3);
    rc |= dev_alloc(&m.tris, (size_t)nf * 3);
    rc |= dev_alloc(&m.area, nf);
    rc |= dev_alloc(&m.normal, (size_t)nf * 3);
    rc |= dev_alloc(&m.vert_area, nv);
    rc |= dev_alloc(&m.vert_normal, (size_t)nv * 3);
    rc |= dev_alloc(&m.curl, (size_t)nf * 9);
    rc |= dev_alloc(&m.cen, (size_t)nf * 3);
    rc |= dev_alloc(&m.pu, (size_t)nf * q2 * 4);
    rc |= dev_alloc(&m.pl, (size_t)nf * q2 * 4);
    rc |= dev_alloc(&m.pu32, (size_t)nf * q2 * 4);
    rc |= dev_alloc(&m.pl32, (size_t)nf * q2 * 4);
    rc |= dev_alloc(&m.phi, (size_t)q2 * 3);
    rc |= dev_alloc(&m.phi32, (size_t)q2 * 3);
    rc |= dev_alloc(&m.mass_w, (size_t)q2);
    rc |= dev_alloc(&m.frame, (size_t)nf * PANEL_FRAME);
    rc |= dev_alloc(&m.pq, (size_t)nf * q2 * 4);
    rc |= dev_alloc(&m.pq32, (size_t)nf * q2 * 4);
    rc |= dev_alloc(&m.panel_buf, (size_t)nf * 6);
    if (rc) return BEM_ERR_HIP;
    rc = check_hip(hipMemcpyAsync(m.verts, d_verts, sizeof(double) * nv * 3, hipMemcpyDeviceToDevice, s), "copy vertices");
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(m.tris, d_tris, sizeof(int32_t) * nf * 3, hipMemcpyDeviceToDevice, s), "copy panels");
    if (rc) return rc;
    m.h_tris.resize((size_t)nf * 3);
    rc = check_hip(hipMemcpyAsync(m.h_tris.data(), d_tris, sizeof(int32_t) * nf * 3, hipMemcpyDeviceToHost, s), "panels to host");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    for (int e = 0; e < nf * 3; ++e)
        if (m.h_tris[e] < 0 || m.h_tris[e] >= nv)
            return set_error(BEM_ERR_INVALID, "panel vertex index out of range");

    // tables of the regular rule (host, tiny)
    std::vector<double> r12w((size_t)q2 * 3), phi((size_t)q2 * 3), mw(q2);
    std::vector<float> phi32((size_t)q2 * 3);
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b) {
            int p = a * q + b;
            double r1 = gp[a], r2 = gp[b] * r1;
            r12w[3 * p] = r1; r12w[3 * p + 1] = r2; r12w[3 * p + 2] = gw[a] * gw[b] * r1;
            phi[3 * p] = 1.0 - r1; phi[3 * p + 1] = r1 - r2; phi[3 * p + 2] = r2;
            mw[p] = gw[a] * gw[b] * r1;
        }
    for (int e = 0; e < q2 * 3; ++e) phi32[e] = (float)phi[e];
    if (dev_alloc(&m.r12w, r12w.size())) return BEM_ERR_HIP;
    hipMemcpyAsync(m.r12w, r12w.data(), r12w.size() * sizeof(double), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(m.phi, phi.data(), phi.size() * sizeof(double), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(m.phi32, phi32.data(), phi32.size() * sizeof(float), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(m.mass_w, mw.data(), mw.size() * sizeof(double), hipMemcpyHostToDevice, s);

    // vertex -> panel-corner CSR (host, deterministic panel order)
    std::vector<int32_t> vp_row(nv + 1, 0), vp_el((size_t)nf * 3);
    for (int e = 0; e < nf * 3; ++e) vp_row[m.h_tris[e] + 1]++;
    for (int v = 0; v < nv; ++v) vp_row[v + 1] += vp_row[v];
    {
        std::vector<int32_t> fill(vp_row.begin(), vp_row.end() - 1);
        for (int e = 0; e < nf * 3; ++e) vp_el[fill[m.h_tris[e]]++] = e;
    }
    if (dev_alloc(&m.vp_row, vp_row.size()) || dev_alloc(&m.vp_el, vp_el.size())) return BEM_ERR_HIP;
    hipMemcpyAsync(m.vp_row, vp_row.data(), vp_row.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(m.vp_el, vp_el.data(), vp_el.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);

    rc = mesh_geometry(m, s);
    if (rc) return rc;

    // near-field rules
    std::vector<NearPt> rules;
    build_rules(q, gp, gw, rules, m.rule_off, m.rule_len);
    if (dev_alloc(&m.rule, rules.size())) return BEM_ERR_HIP;
    hipMemcpyAsync(m.rule, rules.data(), rules.size() * sizeof(NearPt), hipMemcpyHostToDevice, s);
    rc = check_hip(hipStreamSynchronize(s), "mesh build");
    if (rc) return rc;
    return mesh_build_near(ctx, 0, s);
}

}  // namespace

// A build that fails leaves no mesh behind (nf = 0), so every later call
// reports "no mesh" instead of running on partly filled tables or on panel
// indices that were found out of range.
int mesh_build(Context* ctx, const double* d_verts, const int32_t* d_tris, int nv, int nf, int q,
               hipStream_t s) {
    const int rc = mesh_build_impl(ctx, d_verts, d_tris, nv, nf, q, s);
    if (rc != BEM_OK) mesh_free(&ctx->mesh);
    return rc;
}

// Near-field list: all ordered pairs sharing at least one vertex, found
// through the vertex->panel adjacency (O(F) instead of the reference's
// per-pair test inside every kernel), classified with the reference rule.
int mesh_build_near(Context* ctx, int flags, hipStream_t s) {
    Mesh& m = ctx->mesh;
    flags &= BEM_FLAG_ROTATE_NEAR;
    if (m.near && m.near_flags == flags) return BEM_OK;
    const int nf = m.nf, nv = m.nv;
    const int32_t* t = m.h_tris.data();
    std::vector<std::vector<int32_t>> vp(nv);
    for (int e = 0; e < nf * 3; ++e) vp[t[e]].push_back(e / 3);
    std::vector<NearPair> list;
    std::vector<int32_t> row(nf + 1, 0);
    std::vector<int32_t> cand;
    for (int i = 0; i < 4; ++i) m.near_counts[i] = 0;
    for (int i = 0; i < nf; ++i) {
        cand.clear();
        for (int a = 0; a < 3; ++a)
            for (int j : vp[t[3 * i + a]]) cand.push_back(j);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (int j : cand) {
            int rel = relation(t, i, j);
            if (rel == BEM_SEPARATE) continue;  // cannot happen: j shares a vertex
            int px[3], py[3];
            near_perm(t, i, j, rel, flags, px, py);
            list.push_back({i, j, rel, encode_perm(px, py)});
            m.near_counts[rel]++;
        }
        row[i + 1] = (int32_t)list.size();
    }
    m.n_near = (int64_t)list.size();
    m.near_counts[BEM_SEPARATE] = (int64_t)nf * nf - m.n_near;
    // the relation is symmetric, so (j, i) is in row j, whose trial panels are sorted
    std::vector<int32_t> tr(list.size());
    for (size_t e = 0; e < list.size(); ++e) {
        const int i = list[e].i, j = list[e].j;
        int lo = row[j], hi = row[j + 1];
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (list[mid].j < i) lo = mid + 1; else hi = mid;
        }
        if (lo == row[j + 1] || list[lo].j != i)
            return set_error(BEM_ERR_INVALID, "near list: a pair without its mirror (inconsistent panels)");
        tr[e] = lo;
    }
    if (dev_alloc(&m.near, list.size()) || dev_alloc(&m.near_row, row.size()) || dev_alloc(&m.near_tr, tr.size()))
        return BEM_ERR_HIP;
    hipMemcpyAsync(m.near, list.data(), list.size() * sizeof(NearPair), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(m.near_row, row.data(), row.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(m.near_tr, tr.data(), tr.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (m.near_buf_len < m.n_near * 6) {
        if (dev_alloc(&m.near_buf, (size_t)m.n_near * 6)) return BEM_ERR_HIP;
        m.near_buf_len = m.n_near * 6;
    }
    int rc = check_hip(hipStreamSynchronize(s), "near list upload");
    if (rc) return rc;
    m.near_flags = flags;
    return BEM_OK;
}

}  // namespace bem
